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
// spanwise filament data: g r0 and eps |r0|^4, or r0, 1/|r0|^2 and g for the
// cutoff kernel) is gathered from the
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
#include <mutex>

#include "fj_common.hpp"
#include "fj_lattice_dev.hpp"

// plan: column blocks are halved until about FJ_LAT_WANT_WAVES wave tasks
// exist for the launch's target tiles, but not below FJ_LAT_MIN_WIDTH
// columns (each task re-evaluates its first column's vertices)
#ifndef FJ_LAT_WANT_WAVES
#define FJ_LAT_WANT_WAVES 32768
#endif
#ifndef FJ_LAT_MIN_WIDTH
#define FJ_LAT_MIN_WIDTH 16
#endif
#ifndef FJ_LAT_MIN_WAVES
#define FJ_LAT_MIN_WAVES 5
#endif
#define FJ_MAX_LAT 16

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
  int kernel;    // FJ_KERNEL_*: record layout of the filament data
};

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
  tab[w] = fj_lat_record_word(seg, L.off, L.m, L.n, s, j, f, LL.kernel);
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
    const double* col = tab + (size_t)(L.rec_off + s * (L.n + 1)) * FJ_LAT_REC;
    fj_lat_strip<KIND>(col, rows, last, jb, je, x, y, z, e2, rcut2, ax, ay, az);
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
  // one spare row of 64 past the tiles: a split epilogue (below) reduces its
  // second target range from an offset that is not a multiple of 64, and its
  // last wave's reads then stay inside the partial row
  p.ld = (p.tiles + 1) * 64;
  int max_cols = 1;
  for (int k = 0; k < n_lat; ++k) {
    if (lats[k].m < 0 || lats[k].n < 0 || lats[k].seg_off < 0) return FJ_ERR_ARG;
    if (lats[k].m > 0 && lats[k].n > 0 && lats[k].n + 1 > max_cols) max_cols = lats[k].n + 1;
  }
  const long long want = (long long)FJ_LAT_WANT_WAVES / p.tiles;
  int width = max_cols;
  for (;;) {
    long long tasks = 0;
    for (int k = 0; k < n_lat; ++k)
      if (lats[k].m > 0 && lats[k].n > 0)
        tasks += (long long)fj_cdiv(lats[k].m, FJ_LAT_ROWS) * fj_cdiv(lats[k].n + 1, width);
    if (tasks >= want || width <= FJ_LAT_MIN_WIDTH) break;
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
  L.kernel = FJ_KERNEL_EPSILON;
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

// The body of fj_pair_sum_lattices on a given workspace (partials, then the
// strip-column records), shared with the allocation-free and the split entry
// points below: targets [0, n_a) get epilogue mode_a into out_a, targets
// [n_a, n_tgt) mode_b into out_b (n_a = n_tgt: one range).
static int fj_lat_run(const double* seg, int n_lat, const fj_lattice_ref* lats,
                      const double* targets, int n_tgt, int n_a, double* out_a, int mode_a,
                      double* out_b, int mode_b, int kernel, double eps, double rcut,
                      const fj_flowfield* flow, double time, double dt, void* workspace,
                      size_t ws_limit, void* stream) {
  if (n_tgt < 0 || n_a < 0 || n_a > n_tgt) return FJ_ERR_ARG;
  if (kernel < FJ_KERNEL_EPSILON || kernel > FJ_KERNEL_PLAIN) return FJ_ERR_ARG;
  if (mode_a < FJ_OUT_STORE || mode_a > FJ_OUT_CONVECT) return FJ_ERR_ARG;
  if (n_a < n_tgt && (mode_b < FJ_OUT_STORE || mode_b > FJ_OUT_CONVECT || !out_b)) return FJ_ERR_ARG;
  FjLatPlan p;
  const int rc = fj_lat_plan(n_tgt, n_lat, lats, p);
  if (rc != FJ_OK) return rc;
  if (n_tgt == 0) return FJ_OK;
  if (!targets || (n_a > 0 && !out_a)) return FJ_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const FjFlow fl = fj_flow_from_desc(flow);
  const int n_b = n_tgt - n_a;
  if (p.list.n_tasks == 0) {
    int r = FJ_OK;
    if (n_a > 0) r = fj_pair_epilogue(nullptr, 0, 0, n_a, targets, out_a, mode_a, fl, time, dt, st);
    if (r == FJ_OK && n_b > 0)
      r = fj_pair_epilogue(nullptr, 0, 0, n_b, targets + 3 * (size_t)n_a, out_b, mode_b, fl, time, dt, st);
    return r;
  }
  if (!seg) return FJ_ERR_ARG;
  if (fj_lat_part_bytes(p) + (size_t)p.list.n_recs * FJ_LAT_REC * sizeof(double) > ws_limit)
    return FJ_ERR_SCRATCH;
  if (!workspace) return FJ_ERR_ARG;
  FjKernelParams kp;
  kp.eps = eps; kp.eps2 = eps * eps; kp.rcut = rcut;
  double* part = (double*)workspace;
  double* tab = (double*)((char*)workspace + fj_lat_part_bytes(p));
  p.list.kernel = kernel;
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
  int r = FJ_OK;
  if (n_a > 0) r = fj_pair_epilogue(part, p.chunks, p.ld, n_a, targets, out_a, mode_a, fl, time, dt, st);
  if (r == FJ_OK && n_b > 0)
    r = fj_pair_epilogue(part + n_a, p.chunks, p.ld, n_b, targets + 3 * (size_t)n_a, out_b, mode_b, fl, time,
                         dt, st);
  return r;
}

extern "C" int fj_pair_sum_lattices(const double* seg, int n_lat, const fj_lattice_ref* lats,
                                    const double* targets, int n_tgt, double* out, int kernel,
                                    double eps, double rcut, int mode, const fj_flowfield* flow,
                                    double time, double dt, void* workspace, void* stream) {
  return fj_lat_run(seg, n_lat, lats, targets, n_tgt, n_tgt, out, mode, nullptr, FJ_OUT_STORE, kernel, eps, rcut,
                    flow, time, dt, workspace, (size_t)-1, stream);
}

extern "C" int fj_pair_sum_lattices_split(const double* seg, int n_lat, const fj_lattice_ref* lats,
                                          const double* targets, int n_tgt, int n_tgt_a, double* out_a,
                                          int mode_a, double* out_b, int mode_b, int kernel, double eps,
                                          double rcut, const fj_flowfield* flow, double time, double dt,
                                          void* workspace, void* stream) {
  return fj_lat_run(seg, n_lat, lats, targets, n_tgt, n_tgt_a, out_a, mode_a, out_b, mode_b, kernel, eps, rcut,
                    flow, time, dt, workspace, (size_t)-1, stream);
}

// ---------------------------------------------------------------------------
// Allocation-free variant: the workspace is a file-scope device array of
// FJ_NOALLOC_SCRATCH_BYTES (one copy per GPU, held until the process ends).
// In host code `scratch` names a host-side stand-in, never the device
// memory: its device address is taken once per device with
// hipGetSymbolAddress (device current) and only that address is used.
// Nothing in it needs to start at zero (the records kernel writes every
// record the partial kernel reads, and the partial kernel writes every
// partial the reduction reads), so no clearing launch is needed.  Calls that
// overlap on one GPU share this memory and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[FJ_NOALLOC_SCRATCH_BYTES];

#define FJ_NOALLOC_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[FJ_NOALLOC_MAX_DEVICES];

static int fj_scratch_ptr(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= FJ_NOALLOC_MAX_DEVICES) return FJ_ERR_LAUNCH;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FJ_ERR_LAUNCH;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FJ_OK;
}

extern "C" int fj_pair_sum_lattices_noalloc(const double* seg, int n_lat, const fj_lattice_ref* lats,
                                            const double* targets, int n_tgt, double* out,
                                            int kernel, double eps, double rcut, int mode,
                                            const fj_flowfield* flow, double time, double dt,
                                            void* stream) {
  // sizes first (host-only), so an oversized call fails before any device call
  if (n_tgt > 0) {
    FjLatPlan p;
    if (n_lat >= 0 && fj_lat_plan(n_tgt, n_lat, lats, p) == FJ_OK && p.list.n_tasks > 0 &&
        fj_lat_part_bytes(p) + (size_t)p.list.n_recs * FJ_LAT_REC * sizeof(double) >
            (size_t)FJ_NOALLOC_SCRATCH_BYTES)
      return FJ_ERR_SCRATCH;
  }
  void* ws = nullptr;
  const int rc = fj_scratch_ptr(&ws);
  if (rc != FJ_OK) return rc;
  return fj_lat_run(seg, n_lat, lats, targets, n_tgt, n_tgt, out, mode, nullptr, FJ_OUT_STORE, kernel, eps, rcut,
                    flow, time, dt, ws, (size_t)FJ_NOALLOC_SCRATCH_BYTES, stream);
}
