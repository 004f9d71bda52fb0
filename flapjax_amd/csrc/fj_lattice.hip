// K1 on vortex lattices: the pair sum over the filaments of whole vertex
// grids (the bound and wake surfaces of compute_v_ind, aero/aic.py:248-310),
// sharing the per-vertex half of the kernel between the filaments that meet
// at a vertex.
//
// biot_savart_epsilon (aero/utils.py:606-623) evaluates, per filament,
//   inv_k = 1/sqrt(|r_k|^2 + eps^2),  r_k = x - y_k     (k = 1, 2)
//   d = r1 inv_1 - r2 inv_2,  c = r1 x r2,
//   v = g c (r0 . d) / (|c|^2 + eps |r0|^4)
// r_k and inv_k belong to the VERTEX y_k, and every interior vertex of a
// lattice is an end point of four filaments, so per target a lattice of V
// vertices needs about V reciprocal square roots, not 2 S ~ 4 V.  The kernel
// walks a strip of FJ_LAT_ROWS ring rows column by column, keeping the
// (r, inv) of the strip's vertices in the current and previous column in
// registers: per column it evaluates the FJ_LAT_ROWS + 1 new vertices, the
// chordwise filaments between them and the spanwise filaments back to the
// previous column.  The division by |c|^2 + eps |r0|^4 is a reciprocal with
// one third-order Newton step.
//
// Data.  Everything one column of a strip needs (its vertices, chordwise and
// spanwise filament data r0, eps |r0|^4 or 1/|r0|^2, g) is gathered from the
// ordinary filament table (fj_segments_from_grid order) into one contiguous
// strip-column record, so that a wave fetches it through the scalar unit in
// a few wide loads and waits once per column (every lane of a wave works on
// the same strip).
//
// Work items are (lattice, strip, column block) tasks; a workgroup of
// FJ_BS_WAVES waves takes one 64-target tile and FJ_BS_WAVES consecutive
// tasks of a chunk (one per wave, strided over the chunk), sums its waves in
// LDS in a fixed order and writes one partial per chunk; the partials are
// summed in a fixed order by the common reduction (deterministic).
#include "fj_common.hpp"

#ifndef FJ_LAT_ROWS
#define FJ_LAT_ROWS 2
#endif
#ifndef FJ_LAT_MIN_WAVES
#define FJ_LAT_MIN_WAVES 5
#endif
#define FJ_MAX_LAT 16

// strip-column record: vertices (FJ_LAT_ROWS + 1) x 3, chordwise filaments
// FJ_LAT_ROWS x 5, spanwise filaments (to the previous column)
// (FJ_LAT_ROWS + 1) x 5; a filament's 5 are r0 (3), aux, g
#define FJ_LAT_COFF (3 * (FJ_LAT_ROWS + 1))
#define FJ_LAT_SOFF (FJ_LAT_COFF + 5 * FJ_LAT_ROWS)
#define FJ_LAT_RAW (FJ_LAT_SOFF + 5 * (FJ_LAT_ROWS + 1))
#define FJ_LAT_REC ((FJ_LAT_RAW + 3) / 4 * 4)

struct FjLattice {
  int off;       // first filament of this lattice in the filament table
  int m, n;      // ring rows and columns
  int cblocks;   // column blocks per strip
  int task_off;  // first task of this lattice
  int rec_off;   // first strip-column record of this lattice
};

struct FjLatticeList {
  FjLattice l[FJ_MAX_LAT];
  int count;     // lattices with work (m > 0 and n > 0)
  int n_tasks;
  int n_recs;
  int cb_width;  // chordwise-filament columns per block
};

struct FjVtx {
  double rx, ry, rz, s, inv;
};

__device__ __forceinline__ FjVtx fj_lat_vertex(const double* __restrict__ v, double x, double y,
                                               double z, double e2) {
  FjVtx o;
  o.rx = x - v[0];
  o.ry = y - v[1];
  o.rz = z - v[2];
  o.s = fma(o.rz, o.rz, fma(o.ry, o.ry, fma(o.rx, o.rx, e2)));
  o.inv = fj_rsqrt(o.s);
  return o;
}

// filament A -> B with data f = (r0 = B - A, aux, g)
template <int KIND>
__device__ __forceinline__ void fj_lat_eval(const FjVtx& A, const FjVtx& B,
                                            const double* __restrict__ f, double e2,
                                            double rcut2, double& ax, double& ay, double& az) {
  const double cx = A.ry * B.rz - A.rz * B.ry;
  const double cy = A.rz * B.rx - A.rx * B.rz;
  const double cz = A.rx * B.ry - A.ry * B.rx;
  const double dx = fma(A.rx, A.inv, -(B.rx * B.inv));
  const double dy = fma(A.ry, A.inv, -(B.ry * B.inv));
  const double dz = fma(A.rz, A.inv, -(B.rz * B.inv));
  const double a = f[0] * dx + f[1] * dy + f[2] * dz;
  double D = fma(cz, cz, fma(cy, cy, cx * cx));
  if (KIND == FJ_KERNEL_EPSILON) D += f[3];
  // 1/D: hardware estimate, then y (1 + e + e^2) with e = 1 - D y
  double w = __builtin_amdgcn_rcp(D);
  const double e = fma(-D, w, 1.0);
  w = fma(w, fma(e, e, e), w);
  double q = (f[4] * a) * w;
  if (KIND == FJ_KERNEL_CUTOFF) {
    // squared distance from x to the filament line: |r1|^2 - (r0.r1)^2 / |r0|^2
    const double a1 = f[0] * A.rx + f[1] * A.ry + f[2] * A.rz;
    const double rad2 = (A.s - e2) - a1 * (a1 * f[3]);
    q = (rad2 > rcut2) ? q : 0.0;
  }
  ax = fma(cx, q, ax);
  ay = fma(cy, q, ay);
  az = fma(cz, q, az);
}

// one task: R chordwise filament rows and SR spanwise rows (R, plus the
// closing row on a lattice's last strip) of the strip whose column records
// start at `col`, chordwise columns jb .. je-1
template <int KIND, int R, int SR>
__device__ __forceinline__ void fj_lat_task(const double* __restrict__ col, int jb, int je,
                                            double x, double y, double z, double e2, double rcut2,
                                            double& ax, double& ay, double& az) {
  FjVtx P[R + 1], C[R + 1];
  if (jb > 0) {
    const double* r = col + (size_t)(jb - 1) * FJ_LAT_REC;
#pragma unroll
    for (int k = 0; k <= R; ++k) P[k] = fj_lat_vertex(r + 3 * k, x, y, z, e2);
  }
  for (int j = jb; j < je; ++j) {
    const double* rp = col + (size_t)j * FJ_LAT_REC;
    double r[FJ_LAT_RAW];
#pragma unroll
    for (int k = 0; k < FJ_LAT_RAW; ++k) r[k] = rp[k];
#pragma unroll
    for (int k = 0; k <= R; ++k) C[k] = fj_lat_vertex(r + 3 * k, x, y, z, e2);
#pragma unroll
    for (int k = 0; k < R; ++k)
      fj_lat_eval<KIND>(C[k], C[k + 1], r + FJ_LAT_COFF + 5 * k, e2, rcut2, ax, ay, az);
    if (j > 0) {
#pragma unroll
      for (int k = 0; k < SR; ++k)
        fj_lat_eval<KIND>(P[k], C[k], r + FJ_LAT_SOFF + 5 * k, e2, rcut2, ax, ay, az);
    }
#pragma unroll
    for (int k = 0; k <= R; ++k) P[k] = C[k];
  }
}

// strip-column records from the filament table: one thread per record
// word (consecutive threads write consecutive words)
__global__ void k_lattice_records(const double* __restrict__ seg, FjLatticeList LL,
                                  double* __restrict__ tab) {
  const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= (long long)LL.n_recs * FJ_LAT_REC) return;
  const int q = (int)(w / FJ_LAT_REC), f = (int)(w - (long long)q * FJ_LAT_REC);
  int li = 0;
  while (li + 1 < LL.count && q >= LL.l[li + 1].rec_off) ++li;
  const FjLattice& L = LL.l[li];
  const int local = q - L.rec_off;
  const int s = local / (L.n + 1), j = local - s * (L.n + 1);
  const int i0 = s * FJ_LAT_ROWS;
  const int rows = min(FJ_LAT_ROWS, L.m - i0);
  const int srows = rows + ((i0 + rows == L.m) ? 1 : 0);
  const double* chord = seg + (size_t)L.off * FJ_SEG_STRIDE;
  const double* span = seg + (size_t)(L.off + L.m * (L.n + 1)) * FJ_SEG_STRIDE;
  double v = 0.0;
  if (f < FJ_LAT_COFF) {
    // vertex (i, j): start of chordwise filament (i, j), or the end of
    // chordwise filament (m-1, j) on the last vertex row
    const int k = f / 3, c = f - 3 * k, i = i0 + k;
    if (k <= rows)
      v = (i < L.m) ? chord[(size_t)(i * (L.n + 1) + j) * FJ_SEG_STRIDE + c]
                    : chord[(size_t)((L.m - 1) * (L.n + 1) + j) * FJ_SEG_STRIDE + 3 + c];
  } else if (f < FJ_LAT_SOFF) {
    const int k = (f - FJ_LAT_COFF) / 5, c = f - FJ_LAT_COFF - 5 * k;
    if (k < rows) v = chord[(size_t)((i0 + k) * (L.n + 1) + j) * FJ_SEG_STRIDE + 6 + c];
  } else if (f < FJ_LAT_RAW) {
    const int k = (f - FJ_LAT_SOFF) / 5, c = f - FJ_LAT_SOFF - 5 * k;
    if (j > 0 && k < srows) v = span[(size_t)((i0 + k) * L.n + j - 1) * FJ_SEG_STRIDE + 6 + c];
  }
  tab[w] = v;
}

template <int KIND>
__global__ __launch_bounds__(FJ_BS_THREADS, KIND == FJ_KERNEL_CUTOFF ? 4 : FJ_LAT_MIN_WAVES) void k_lattice_partial(
    const double* __restrict__ tab, FjLatticeList LL, int chunk_len,
    const double* __restrict__ tgt, int n_tgt, int ld, double* __restrict__ part,
    FjKernelParams kp) {
  __shared__ double red[FJ_BS_WAVES - 1][3][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tile, chunk;
  fj_tile_chunk(tile, chunk);
  const int t = tile * 64 + lane;
  const int tc = (t < n_tgt) ? t : (n_tgt - 1);
  const double x = tgt[3 * tc], y = tgt[3 * tc + 1], z = tgt[3 * tc + 2];
  const double e2 = (KIND == FJ_KERNEL_PLAIN) ? 0.0 : kp.eps2;
  const double rcut2 = kp.rcut * kp.rcut;
  double ax = 0.0, ay = 0.0, az = 0.0;
  const int k_begin = chunk * chunk_len;
  const int k_end = min(LL.n_tasks, k_begin + chunk_len);
  for (int task = k_begin + wave; task < k_end; task += FJ_BS_WAVES) {
    int li = 0;
    while (li + 1 < LL.count && task >= LL.l[li + 1].task_off) ++li;
    const FjLattice& L = LL.l[li];
    const int local = task - L.task_off;
    const int s = local / L.cblocks, b = local - s * L.cblocks;
    const int i0 = s * FJ_LAT_ROWS;
    const int rows = min(FJ_LAT_ROWS, L.m - i0);
    const int last = (i0 + rows == L.m) ? 1 : 0;
    const int jb = b * LL.cb_width;
    const int je = min(L.n + 1, jb + LL.cb_width);
    // every strip but a lattice's last has FJ_LAT_ROWS rows
    const double* col = tab + (size_t)(L.rec_off + s * (L.n + 1)) * FJ_LAT_REC;
#define FJ_LAT_ARGS col, jb, je, x, y, z, e2, rcut2, ax, ay, az
    if (rows == FJ_LAT_ROWS && !last) fj_lat_task<KIND, FJ_LAT_ROWS, FJ_LAT_ROWS>(FJ_LAT_ARGS);
    else if (rows == FJ_LAT_ROWS) fj_lat_task<KIND, FJ_LAT_ROWS, FJ_LAT_ROWS + 1>(FJ_LAT_ARGS);
#if FJ_LAT_ROWS > 3
    else if (rows == 3) fj_lat_task<KIND, 3, 4>(FJ_LAT_ARGS);
#endif
#if FJ_LAT_ROWS > 2
    else if (rows == 2) fj_lat_task<KIND, 2, 3>(FJ_LAT_ARGS);
#endif
#if FJ_LAT_ROWS > 1
    else if (rows == 1) fj_lat_task<KIND, 1, 2>(FJ_LAT_ARGS);
#endif
#undef FJ_LAT_ARGS
  }
  if (wave > 0) {
    red[wave - 1][0][lane] = ax;
    red[wave - 1][1][lane] = ay;
    red[wave - 1][2][lane] = az;
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int w = 0; w < FJ_BS_WAVES - 1; ++w) {
      ax += red[w][0][lane];
      ay += red[w][1][lane];
      az += red[w][2][lane];
    }
    double* o = part + (size_t)chunk * 3 * ld;
    o[t] = ax;
    o[ld + t] = ay;
    o[2 * ld + t] = az;
  }
}

// ---------------------------------------------------------------------------
// plan: tasks = (lattice, strip, column block); the column blocks are
// halved until the tasks fill about FJ_BS_TARGET_BLOCKS workgroups of
// FJ_BS_WAVES waves for the given number of target tiles
// ---------------------------------------------------------------------------
struct FjLatPlan {
  FjLatticeList list;
  int tiles, chunks, chunk_len, ld;
};

static int fj_lat_plan(int n_tgt, int n_lat, const fj_lattice_ref* lats, FjLatPlan& p) {
  if (n_lat < 0 || n_lat > FJ_MAX_LAT || (n_lat > 0 && !lats)) return FJ_ERR_ARG;
  p.tiles = n_tgt > 0 ? fj_cdiv(n_tgt, 64) : 1;
  p.ld = p.tiles * 64;
  int max_cols = 1;
  for (int k = 0; k < n_lat; ++k) {
    if (lats[k].m < 0 || lats[k].n < 0 || lats[k].seg_off < 0) return FJ_ERR_ARG;
    if (lats[k].m > 0 && lats[k].n > 0 && lats[k].n + 1 > max_cols) max_cols = lats[k].n + 1;
  }
  const long long want = (long long)FJ_BS_TARGET_BLOCKS * FJ_BS_WAVES / p.tiles;
  int width = max_cols;
  for (;;) {
    long long tasks = 0;
    for (int k = 0; k < n_lat; ++k)
      if (lats[k].m > 0 && lats[k].n > 0)
        tasks += (long long)fj_cdiv(lats[k].m, FJ_LAT_ROWS) * fj_cdiv(lats[k].n + 1, width);
    if (tasks >= want || width <= 8) break;
    width = fj_cdiv(width, 2);
  }
  FjLatticeList& L = p.list;
  L.cb_width = width;
  L.count = 0;
  int tasks = 0, recs = 0;
  for (int k = 0; k < n_lat; ++k) {
    if (lats[k].m <= 0 || lats[k].n <= 0) continue;  // all filament strengths are zero
    FjLattice& e = L.l[L.count++];
    e.off = lats[k].seg_off;
    e.m = lats[k].m;
    e.n = lats[k].n;
    e.cblocks = fj_cdiv(e.n + 1, width);
    e.task_off = tasks;
    e.rec_off = recs;
    tasks += fj_cdiv(e.m, FJ_LAT_ROWS) * e.cblocks;
    recs += fj_cdiv(e.m, FJ_LAT_ROWS) * (e.n + 1);
  }
  L.n_tasks = tasks;
  L.n_recs = recs;
  // one task per wave per workgroup, longer chunks when there is more work
  // than about FJ_BS_TARGET_BLOCKS workgroups
  int per_chunk = FJ_BS_WAVES;
  const long long blocks = (long long)p.tiles * fj_cdiv(tasks > 0 ? tasks : 1, per_chunk);
  if (blocks > 2LL * FJ_BS_TARGET_BLOCKS)
    per_chunk = FJ_BS_WAVES * (int)((blocks + 2LL * FJ_BS_TARGET_BLOCKS - 1) / (2LL * FJ_BS_TARGET_BLOCKS));
  p.chunk_len = per_chunk;
  p.chunks = fj_cdiv(tasks > 0 ? tasks : 1, per_chunk);
  return FJ_OK;
}

// workspace: chunk partials, then the strip-column records
static size_t fj_lat_part_bytes(const FjLatPlan& p) {
  return ((size_t)p.chunks * 3 * p.ld * sizeof(double) + 255) / 256 * 256;
}

extern "C" size_t fj_pair_sum_lattices_workspace_bytes(int n_tgt, int n_lat,
                                                       const fj_lattice_ref* lats) {
  FjLatPlan p;
  if (n_tgt <= 0 || fj_lat_plan(n_tgt, n_lat, lats, p) != FJ_OK) return 256;
  return fj_lat_part_bytes(p) + (size_t)p.list.n_recs * FJ_LAT_REC * sizeof(double) + 256;
}

extern "C" int fj_pair_sum_lattices(const double* seg, int n_lat, const fj_lattice_ref* lats,
                                    const double* targets, int n_tgt, double* out, int kernel,
                                    double eps, double rcut, int mode, const fj_flowfield* flow,
                                    double time, double dt, void* workspace, void* stream) {
  if (n_tgt < 0) return FJ_ERR_ARG;
  if (kernel < FJ_KERNEL_EPSILON || kernel > FJ_KERNEL_PLAIN) return FJ_ERR_ARG;
  if (mode < FJ_OUT_STORE || mode > FJ_OUT_CONVECT) return FJ_ERR_ARG;
  FjLatPlan p;
  const int rc = fj_lat_plan(n_tgt, n_lat, lats, p);
  if (rc != FJ_OK) return rc;
  if (n_tgt == 0) return FJ_OK;
  if (!targets || !out) return FJ_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const FjFlow fl = fj_flow_from_desc(flow);
  if (p.list.n_tasks == 0) return fj_pair_epilogue(nullptr, 0, 0, n_tgt, targets, out, mode, fl, time, dt, st);
  if (!seg || !workspace) return FJ_ERR_ARG;
  FjKernelParams kp;
  kp.eps = eps; kp.eps2 = eps * eps; kp.rcut = rcut;
  double* part = (double*)workspace;
  double* tab = (double*)((char*)workspace + fj_lat_part_bytes(p));
  hipLaunchKernelGGL(k_lattice_records, dim3(fj_cdiv((long long)p.list.n_recs * FJ_LAT_REC, 256)),
                     dim3(256), 0, st, seg, p.list, tab);
  FJ_CHECK_LAUNCH();
  const dim3 grid(p.tiles, p.chunks), blk(FJ_BS_THREADS);
  if (kernel == FJ_KERNEL_EPSILON)
    hipLaunchKernelGGL(k_lattice_partial<FJ_KERNEL_EPSILON>, grid, blk, 0, st, tab, p.list, p.chunk_len, targets, n_tgt, p.ld, part, kp);
  else if (kernel == FJ_KERNEL_CUTOFF)
    hipLaunchKernelGGL(k_lattice_partial<FJ_KERNEL_CUTOFF>, grid, blk, 0, st, tab, p.list, p.chunk_len, targets, n_tgt, p.ld, part, kp);
  else
    hipLaunchKernelGGL(k_lattice_partial<FJ_KERNEL_PLAIN>, grid, blk, 0, st, tab, p.list, p.chunk_len, targets, n_tgt, p.ld, part, kp);
  FJ_CHECK_LAUNCH();
  return fj_pair_epilogue(part, p.chunks, p.ld, n_tgt, targets, out, mode, fl, time, dt, st);
}
