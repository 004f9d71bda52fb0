// K1: de-duplicated filament tables and tiled Biot-Savart pair sums.
//
// Replaces the lax.map(row) o vmap(kernel) o einsum contraction of
// v_ind_vmap / compute_v_ind (aero/aic.py:208-310).
//
// Launch shape.  Targets sit in registers, TPL per lane, one 64*TPL target
// tile per workgroup.  The filament list is cut into chunks (grid.y) so that
// even 1k-target launches fill the 256 CUs; the four waves of a workgroup
// walk interleaved filaments of the chunk.  Every lane of a wave evaluates the
// same filament at the same time, so the 96-byte filament record is fetched
// once per wave through the scalar unit into SGPRs (no LDS traffic, no
// barriers in the loop) and fed to the FP64 VALU as scalar operands.  The four
// waves are summed through LDS in a fixed order and each workgroup writes one
// partial per target; a second kernel sums the chunk partials in a fixed order
// (deterministic, no atomics) and applies the epilogue (store / add /
// free-wake convection).
#include "fj_common.hpp"
#include "fj_segments_dev.hpp"


// ---------------------------------------------------------------------------
// filament table
// ---------------------------------------------------------------------------
__global__ void k_segments_from_grid(const double* __restrict__ zeta,
                                     const double* __restrict__ gamma, int m, int n, int mirror,
                                     double mpx, double mpy, double mpz, double mnx, double mny,
                                     double mnz, int kernel, double eps,
                                     double* __restrict__ seg) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m * (n + 1) + (m + 1) * n) return;
  const double mp[3] = {mpx, mpy, mpz}, mn[3] = {mnx, mny, mnz};
  fj_segment_record(zeta, gamma, m, n, s, mirror, mp, mn, kernel, eps,
                    seg + (size_t)s * FJ_SEG_STRIDE);
}

// several grids into one table (their records concatenated in list order)
#define FJ_MAX_GRIDS 16
struct FjGridList {
  const double* zeta[FJ_MAX_GRIDS];
  const double* gamma[FJ_MAX_GRIDS];
  int m[FJ_MAX_GRIDS], n[FJ_MAX_GRIDS], mirror[FJ_MAX_GRIDS], off[FJ_MAX_GRIDS + 1];
  int count;
  double mp[3], mn[3];
};

__global__ void k_segments_from_grids(FjGridList gl, int kernel, double eps,
                                      double* __restrict__ seg) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= gl.off[gl.count]) return;
  int k = 0;
  while (s >= gl.off[k + 1]) ++k;
  fj_segment_record(gl.zeta[k], gl.gamma[k], gl.m[k], gl.n[k], s - gl.off[k], gl.mirror[k], gl.mp,
                    gl.mn, kernel, eps, seg + (size_t)s * FJ_SEG_STRIDE);
}

extern "C" int fj_segments_from_grids(int n_grids, const fj_grid_ref* grids,
                                      const double* mirror_point, const double* mirror_normal,
                                      int kernel, double eps, double* seg_out, void* stream) {
  if (n_grids < 0 || (n_grids > 0 && (!grids || !seg_out))) return FJ_ERR_ARG;
  if (kernel < FJ_KERNEL_EPSILON || kernel > FJ_KERNEL_PLAIN) return FJ_ERR_ARG;
  size_t base = 0;
  for (int g0 = 0; g0 < n_grids; g0 += FJ_MAX_GRIDS) {
    FjGridList gl;
    gl.count = (n_grids - g0 < FJ_MAX_GRIDS) ? n_grids - g0 : FJ_MAX_GRIDS;
    gl.off[0] = 0;
    for (int k = 0; k < gl.count; ++k) {
      const fj_grid_ref& r = grids[g0 + k];
      if (r.m < 0 || r.n < 0 || !r.zeta) return FJ_ERR_ARG;
      if (r.mirror && (!mirror_point || !mirror_normal)) return FJ_ERR_ARG;
      gl.zeta[k] = r.zeta;
      gl.gamma[k] = r.gamma;
      gl.m[k] = r.m;
      gl.n[k] = r.n;
      gl.mirror[k] = r.mirror ? 1 : 0;
      gl.off[k + 1] = gl.off[k] + r.m * (r.n + 1) + (r.m + 1) * r.n;
    }
    for (int a = 0; a < 3; ++a) {
      gl.mp[a] = mirror_point ? mirror_point[a] : 0.0;
      gl.mn[a] = mirror_normal ? mirror_normal[a] : 0.0;
    }
    const int total = gl.off[gl.count];
    if (total > 0) {
      hipLaunchKernelGGL(k_segments_from_grids, dim3(fj_cdiv(total, 256)), dim3(256), 0,
                         (hipStream_t)stream, gl, kernel, eps, seg_out + base * FJ_SEG_STRIDE);
      FJ_CHECK_LAUNCH();
    }
    base += (size_t)total;
  }
  return FJ_OK;
}

extern "C" int fj_segments_from_grid(const double* zeta, const double* gamma, int m, int n,
                                     const double* mirror_point, const double* mirror_normal,
                                     int kernel, double eps, double* seg_out, void* stream) {
  if (m < 0 || n < 0 || !seg_out || !zeta) return FJ_ERR_ARG;
  if (kernel < FJ_KERNEL_EPSILON || kernel > FJ_KERNEL_PLAIN) return FJ_ERR_ARG;
  const int n_all = m * (n + 1) + (m + 1) * n;
  if (n_all == 0) return FJ_OK;
  const int mirror = (mirror_point && mirror_normal) ? 1 : 0;
  const double z3[3] = {0, 0, 0};
  const double* mp = mirror ? mirror_point : z3;
  const double* mn = mirror ? mirror_normal : z3;
  hipLaunchKernelGGL(k_segments_from_grid, dim3(fj_cdiv(n_all, 256)), dim3(256), 0,
                     (hipStream_t)stream, zeta, gamma, m, n, mirror, mp[0], mp[1], mp[2], mn[0],
                     mn[1], mn[2], kernel, eps, seg_out);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// filaments given by their end points (S, 2, 3) and circulations (S): the
// filament-grid input of v_ind_vmap (aero/aic.py:208-245)
__global__ void k_segments_from_endpoints(const double* __restrict__ ends,
                                          const double* __restrict__ g, int ns, int kernel,
                                          double eps, double* __restrict__ seg) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const double* e = ends + 6 * (size_t)s;
  const double r0x = e[3] - e[0], r0y = e[4] - e[1], r0z = e[5] - e[2];
  const double r0sq = r0x * r0x + r0y * r0y + r0z * r0z;
  const double aux = (kernel == FJ_KERNEL_CUTOFF) ? 1.0 / r0sq : eps * r0sq * r0sq;
  const double gs = g ? g[s] : 1.0;
  double* o = seg + (size_t)s * FJ_SEG_STRIDE;
  for (int k = 0; k < 6; ++k) o[k] = e[k];
  o[6] = r0x; o[7] = r0y; o[8] = r0z;
  o[9] = aux; o[10] = gs; o[11] = r0sq;
  float* f = (float*)(o + 12);
  f[0] = (float)r0x; f[1] = (float)r0y; f[2] = (float)r0z;
  f[3] = (float)aux; f[4] = (float)gs;
  f[5] = 0.0f; f[6] = 0.0f; f[7] = 0.0f;
}

extern "C" int fj_segments_from_endpoints(const double* ends, const double* gamma, int n_seg,
                                          int kernel, double eps, double* seg_out,
                                          void* stream) {
  if (n_seg < 0 || (n_seg > 0 && (!ends || !seg_out))) return FJ_ERR_ARG;
  if (kernel < FJ_KERNEL_EPSILON || kernel > FJ_KERNEL_PLAIN) return FJ_ERR_ARG;
  if (n_seg == 0) return FJ_OK;
  hipLaunchKernelGGL(k_segments_from_endpoints, dim3(fj_cdiv(n_seg, 256)), dim3(256), 0,
                     (hipStream_t)stream, ends, gamma, n_seg, kernel, eps, seg_out);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---------------------------------------------------------------------------
// launch plan
// ---------------------------------------------------------------------------
static FjPairPlan fj_pair_plan(int n_tgt, int n_seg) {
  FjPairPlan p;
  p.tpl = (n_tgt >= FJ_BS_TPL2_MIN_TARGETS) ? 2 : 1;
  p.tiles = fj_cdiv(n_tgt, 64 * p.tpl);
  if (p.tiles < 1) p.tiles = 1;
  int chunks = fj_cdiv(FJ_BS_TARGET_BLOCKS, p.tiles);
  const int max_chunks = fj_cdiv(n_seg, FJ_BS_WAVES * FJ_BS_MIN_SEG_PER_WAVE);
  if (chunks > max_chunks) chunks = max_chunks;
  if (chunks < 1) chunks = 1;
  p.chunk_len = fj_cdiv(n_seg > 0 ? n_seg : 1, chunks);
  p.chunks = fj_cdiv(n_seg > 0 ? n_seg : 1, p.chunk_len);
  // one spare row of 64 past the tiles: a split epilogue reduces its second
  // target range from an offset that is not a multiple of 64, and its last
  // wave's reads then stay inside the partial row
  p.ld = (p.tiles * p.tpl + 1) * 64;
  return p;
}

extern "C" size_t fj_pair_sum_workspace_bytes(int n_tgt, int n_seg, int precision) {
  (void)precision;
  if (n_tgt <= 0 || n_seg <= 0) return 256;
  const FjPairPlan p = fj_pair_plan(n_tgt, n_seg);
  return ((size_t)p.chunks * 3 * p.ld * sizeof(double) + 255) / 256 * 256;
}

// ---------------------------------------------------------------------------
// FP64 partial sums
// ---------------------------------------------------------------------------
template <int KIND, int TPL>
__global__ __launch_bounds__(FJ_BS_THREADS) void k_pair_partial_f64(
    const double* __restrict__ seg, int n_seg, int chunk_len, const double* __restrict__ tgt,
    int n_tgt, int ld, double* __restrict__ part, FjKernelParams kp) {
  __shared__ double red[FJ_BS_WAVES - 1][3][64 * TPL];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tile, chunk;
  fj_tile_chunk(tile, chunk);
  const int t0 = tile * 64 * TPL;

  double x[TPL], y[TPL], z[TPL], ax[TPL], ay[TPL], az[TPL];
#pragma unroll
  for (int p = 0; p < TPL; ++p) {
    int t = t0 + p * 64 + lane;
    t = (t < n_tgt) ? t : (n_tgt - 1);
    x[p] = tgt[3 * t + 0];
    y[p] = tgt[3 * t + 1];
    z[p] = tgt[3 * t + 2];
    ax[p] = 0.0; ay[p] = 0.0; az[p] = 0.0;
  }

  const int s_begin = chunk * chunk_len;
  const int s_end = min(n_seg, s_begin + chunk_len);
  for (int s = s_begin + wave; s < s_end; s += FJ_BS_WAVES) {
    const double* sp = seg + (size_t)s * FJ_SEG_STRIDE;
    double rec[FJ_SEG_STRIDE];
#pragma unroll
    for (int k = 0; k < FJ_SEG_STRIDE; ++k) rec[k] = sp[k];
#pragma unroll
    for (int p = 0; p < TPL; ++p) {
      double vx, vy, vz;
      fj_bs_eval<KIND>(x[p], y[p], z[p], rec, kp, vx, vy, vz);
      ax[p] += vx; ay[p] += vy; az[p] += vz;
    }
  }

  // fixed-order reduction of the four waves
  if (wave > 0) {
#pragma unroll
    for (int p = 0; p < TPL; ++p) {
      red[wave - 1][0][p * 64 + lane] = ax[p];
      red[wave - 1][1][p * 64 + lane] = ay[p];
      red[wave - 1][2][p * 64 + lane] = az[p];
    }
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int p = 0; p < TPL; ++p) {
      double sx = ax[p], sy = ay[p], sz = az[p];
#pragma unroll
      for (int w = 0; w < FJ_BS_WAVES - 1; ++w) {
        sx += red[w][0][p * 64 + lane];
        sy += red[w][1][p * 64 + lane];
        sz += red[w][2][p * 64 + lane];
      }
      const int t = t0 + p * 64 + lane;
      double* o = part + (size_t)chunk * 3 * ld;
      o[t] = sx;
      o[ld + t] = sy;
      o[2 * ld + t] = sz;
    }
  }
}

// Chunk partials -> result, one workgroup per 64 targets: wave w sums chunks
// w, w + FJ_RED_WAVES, ... (independent loads in flight across the waves),
// then wave 0 adds the wave sums in wave order -- a fixed order for given
// sizes, so results are deterministic.
#define FJ_RED_WAVES 8
template <int MODE>
__global__ __launch_bounds__(64 * FJ_RED_WAVES) void k_pair_reduce(
    const double* __restrict__ part, int chunks, int ld, int n_tgt,
    const double* __restrict__ tgt, double* __restrict__ out, FjFlow flow, double time,
    double dt) {
  __shared__ double red[FJ_RED_WAVES - 1][3][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = blockIdx.x * 64 + lane;  // < ld: every partial row holds ld >= 64 * grid entries
  double vx = 0.0, vy = 0.0, vz = 0.0;
#pragma unroll 4
  for (int c = wave; c < chunks; c += FJ_RED_WAVES) {
    const double* o = part + (size_t)c * 3 * ld;
    vx += o[t];
    vy += o[ld + t];
    vz += o[2 * ld + t];
  }
  if (wave > 0) {
    red[wave - 1][0][lane] = vx;
    red[wave - 1][1][lane] = vy;
    red[wave - 1][2][lane] = vz;
  }
  __syncthreads();
  if (wave != 0 || t >= n_tgt) return;
#pragma unroll
  for (int w = 0; w < FJ_RED_WAVES - 1; ++w) {
    vx += red[w][0][lane];
    vy += red[w][1][lane];
    vz += red[w][2][lane];
  }
  vx *= FJ_INV_4PI; vy *= FJ_INV_4PI; vz *= FJ_INV_4PI;
  if (MODE == FJ_OUT_STORE) {
    out[3 * t + 0] = vx; out[3 * t + 1] = vy; out[3 * t + 2] = vz;
  } else if (MODE == FJ_OUT_ADD) {
    out[3 * t + 0] += vx; out[3 * t + 1] += vy; out[3 * t + 2] += vz;
  } else {
    const double px = tgt[3 * t + 0], py = tgt[3 * t + 1], pz = tgt[3 * t + 2];
    double u[3];
    fj_flow_eval(flow, px, py, pz, time, u);
    out[3 * t + 0] = px + dt * (u[0] + vx);
    out[3 * t + 1] = py + dt * (u[1] + vy);
    out[3 * t + 2] = pz + dt * (u[2] + vz);
  }
}

template <int MODE>
__global__ void k_pair_empty(int n_tgt, const double* __restrict__ tgt, double* __restrict__ out,
                             FjFlow flow, double time, double dt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tgt) return;
  if (MODE == FJ_OUT_STORE) {
    out[3 * t + 0] = 0.0; out[3 * t + 1] = 0.0; out[3 * t + 2] = 0.0;
  } else if (MODE == FJ_OUT_CONVECT) {
    const double px = tgt[3 * t + 0], py = tgt[3 * t + 1], pz = tgt[3 * t + 2];
    double u[3];
    fj_flow_eval(flow, px, py, pz, time, u);
    out[3 * t + 0] = px + dt * (u[0] + 0.0);
    out[3 * t + 1] = py + dt * (u[1] + 0.0);
    out[3 * t + 2] = pz + dt * (u[2] + 0.0);
  }
}

template <int KIND, int TPL>
static void launch_partial_f64(const FjPairPlan& p, const double* seg, int n_seg,
                               const double* tgt, int n_tgt, double* part,
                               const FjKernelParams& kp, hipStream_t st) {
  hipLaunchKernelGGL((k_pair_partial_f64<KIND, TPL>), dim3(p.tiles, p.chunks),
                     dim3(FJ_BS_THREADS), 0, st, seg, n_seg, p.chunk_len, tgt, n_tgt, p.ld, part,
                     kp);
}

// Epilogue of a pair sum: the chunk partials (or nothing, for an empty
// filament set: part == NULL) to out by `mode`.
int fj_pair_epilogue(const double* part, int chunks, int ld, int n_tgt, const double* targets,
                     double* out, int mode, const FjFlow& fl, double time, double dt,
                     hipStream_t st) {
  if (!part) {
    const dim3 g(fj_cdiv(n_tgt, 256));
    if (mode == FJ_OUT_STORE)
      hipLaunchKernelGGL(k_pair_empty<FJ_OUT_STORE>, g, dim3(256), 0, st, n_tgt, targets, out, fl, time, dt);
    else if (mode == FJ_OUT_CONVECT)
      hipLaunchKernelGGL(k_pair_empty<FJ_OUT_CONVECT>, g, dim3(256), 0, st, n_tgt, targets, out, fl, time, dt);
    FJ_CHECK_LAUNCH();
    return FJ_OK;
  }
  const dim3 g(fj_cdiv(n_tgt, 64)), blk(64 * FJ_RED_WAVES);
  if (mode == FJ_OUT_STORE)
    hipLaunchKernelGGL(k_pair_reduce<FJ_OUT_STORE>, g, blk, 0, st, part, chunks, ld, n_tgt, targets, out, fl, time, dt);
  else if (mode == FJ_OUT_ADD)
    hipLaunchKernelGGL(k_pair_reduce<FJ_OUT_ADD>, g, blk, 0, st, part, chunks, ld, n_tgt, targets, out, fl, time, dt);
  else
    hipLaunchKernelGGL(k_pair_reduce<FJ_OUT_CONVECT>, g, blk, 0, st, part, chunks, ld, n_tgt, targets, out, fl, time, dt);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// fj_pair_sum over two target ranges: [0, n_a) with mode_a into out_a,
// [n_a, n_tgt) with mode_b into out_b (n_a = n_tgt: one range)
static int fj_pair_run(const double* seg, int n_seg, const double* targets, int n_tgt, int n_a,
                       double* out_a, int mode_a, double* out_b, int mode_b, int kernel, double eps,
                       double rcut, int precision, const fj_flowfield* flow, double time, double dt,
                       void* workspace, hipStream_t st) {
  if (n_tgt < 0 || n_seg < 0 || n_a < 0 || n_a > n_tgt) return FJ_ERR_ARG;
  if (kernel < FJ_KERNEL_EPSILON || kernel > FJ_KERNEL_PLAIN) return FJ_ERR_ARG;
  if (mode_a < FJ_OUT_STORE || mode_a > FJ_OUT_CONVECT) return FJ_ERR_ARG;
  if (n_a < n_tgt && (mode_b < FJ_OUT_STORE || mode_b > FJ_OUT_CONVECT || !out_b)) return FJ_ERR_ARG;
  if (precision != FJ_PREC_FP64 && precision != FJ_PREC_MIXED) return FJ_ERR_ARG;
  if (n_tgt == 0) return FJ_OK;
  if (!targets || (n_a > 0 && !out_a)) return FJ_ERR_ARG;
  const FjFlow fl = fj_flow_from_desc(flow);
  const int n_b = n_tgt - n_a;
  const double* tgt_b = targets + 3 * (size_t)n_a;
  if (n_seg == 0) {
    int r = FJ_OK;
    if (n_a > 0) r = fj_pair_epilogue(nullptr, 0, 0, n_a, targets, out_a, mode_a, fl, time, dt, st);
    if (r == FJ_OK && n_b > 0) r = fj_pair_epilogue(nullptr, 0, 0, n_b, tgt_b, out_b, mode_b, fl, time, dt, st);
    return r;
  }
  if (!seg || !workspace) return FJ_ERR_ARG;
  FjKernelParams kp;
  kp.eps = eps; kp.eps2 = eps * eps; kp.rcut = rcut;
  const FjPairPlan p = fj_pair_plan(n_tgt, n_seg);
  double* part = (double*)workspace;
  if (precision == FJ_PREC_MIXED) {
    const int rc = fj_pair_partial_mixed(&p, seg, n_seg, targets, n_tgt, part, kernel, &kp, st);
    if (rc != FJ_OK) return rc;
  } else {
#define FJ_DISPATCH_TPL(K)                                                               \
  if (p.tpl == 2) launch_partial_f64<K, 2>(p, seg, n_seg, targets, n_tgt, part, kp, st); \
  else launch_partial_f64<K, 1>(p, seg, n_seg, targets, n_tgt, part, kp, st);
    if (kernel == FJ_KERNEL_EPSILON) { FJ_DISPATCH_TPL(FJ_KERNEL_EPSILON) }
    else if (kernel == FJ_KERNEL_CUTOFF) { FJ_DISPATCH_TPL(FJ_KERNEL_CUTOFF) }
    else { FJ_DISPATCH_TPL(FJ_KERNEL_PLAIN) }
#undef FJ_DISPATCH_TPL
  }
  FJ_CHECK_LAUNCH();
  int r = FJ_OK;
  if (n_a > 0) r = fj_pair_epilogue(part, p.chunks, p.ld, n_a, targets, out_a, mode_a, fl, time, dt, st);
  if (r == FJ_OK && n_b > 0)
    r = fj_pair_epilogue(part + n_a, p.chunks, p.ld, n_b, tgt_b, out_b, mode_b, fl, time, dt, st);
  return r;
}

extern "C" int fj_pair_sum(const double* seg, int n_seg, const double* targets, int n_tgt,
                           double* out, int kernel, double eps, double rcut, int precision,
                           int mode, const fj_flowfield* flow, double time, double dt,
                           void* workspace, void* stream) {
  return fj_pair_run(seg, n_seg, targets, n_tgt, n_tgt, out, mode, nullptr, FJ_OUT_STORE, kernel, eps, rcut,
                     precision, flow, time, dt, workspace, (hipStream_t)stream);
}

extern "C" int fj_pair_sum_split(const double* seg, int n_seg, const double* targets, int n_tgt,
                                 int n_tgt_a, double* out_a, int mode_a, double* out_b, int mode_b,
                                 int kernel, double eps, double rcut, int precision,
                                 const fj_flowfield* flow, double time, double dt, void* workspace,
                                 void* stream) {
  return fj_pair_run(seg, n_seg, targets, n_tgt, n_tgt_a, out_a, mode_a, out_b, mode_b, kernel, eps, rcut,
                     precision, flow, time, dt, workspace, (hipStream_t)stream);
}
