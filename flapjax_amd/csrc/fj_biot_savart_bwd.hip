// Adjoint of the pair sum v_t = (1/4pi) sum_s g_s K(x_t; y0_s, y1_s)
// (epsilon, cutoff and plain kernels), for hand-written reverse mode.
//
// K depends on x only through r1 = x - y0 and r2 = x - y1, so with upstream
// weights w_t the per-pair derivatives reduce to two 3-vectors
//   G1 = d(w.K)/dr1,  G2 = d(w.K)/dr2,
// and  dL/dx_t += g (G1 + G2),  dL/dy0_s -= g G1,  dL/dy1_s -= g G2,
//      dL/dg_s  += w_t . K.
// Writing K = c * phi, c = r1 x r2, phi = N / D with N = r0 . d,
// r0 = r1 - r2, d = a1 r1 - a2 r2, a_i = (|r_i|^2 + eps^2)^(-1/2),
// D = |c|^2 + eps |r0|^4 (epsilon kernel), P = w . c:
//   G1 = phi (r2 x w) + P (dN/dr1 - phi dD/dr1) / D
//   G2 = phi (w x r1) + P (dN/dr2 - phi dD/dr2) / D
//   dN/dr1 =  d + a1 r0 - a1^3 (r0.r1) r1,  dN/dr2 = -d - a2 r0 + a2^3 (r0.r2) r2
//   dD/dr1 = 2 r2 x c + 4 eps |r0|^2 r0,    dD/dr2 = 2 c x r1 - 4 eps |r0|^2 r0
// (cutoff: no eps term in D, zero inside the cut-off radius, the radius test
// itself is piecewise constant; plain: eps = 0).
//
// Two launches: target-major (dL/dx, same tiling as the forward pass, filament
// records through the scalar unit) and filament-major (dL/dy0, dL/dy1, dL/dg:
// one filament per lane, targets and weights read wave-uniformly).
#include "fj_common.hpp"

#define FJ_BWD_THREADS 256

template <int KIND>
__device__ __forceinline__ void fj_bs_vjp(double r1x, double r1y, double r1z, double r2x,
                                          double r2y, double r2z, double wx, double wy, double wz,
                                          const FjKernelParams& kp, double* G1, double* G2,
                                          double& wK) {
  const double r0x = r1x - r2x, r0y = r1y - r2y, r0z = r1z - r2z;
  const double cx = r1y * r2z - r1z * r2y;
  const double cy = r1z * r2x - r1x * r2z;
  const double cz = r1x * r2y - r1y * r2x;
  const double e2 = (KIND == FJ_KERNEL_PLAIN) ? 0.0 : kp.eps2;
  const double s1 = r1x * r1x + r1y * r1y + r1z * r1z + e2;
  const double s2 = r2x * r2x + r2y * r2y + r2z * r2z + e2;
  const double a1 = 1.0 / sqrt(s1), a2 = 1.0 / sqrt(s2);
  const double dx = a1 * r1x - a2 * r2x, dy = a1 * r1y - a2 * r2y, dz = a1 * r1z - a2 * r2z;
  const double N = r0x * dx + r0y * dy + r0z * dz;
  const double r0sq = r0x * r0x + r0y * r0y + r0z * r0z;
  const double epsr = (KIND == FJ_KERNEL_EPSILON) ? kp.eps : 0.0;
  const double D = cx * cx + cy * cy + cz * cz + epsr * r0sq * r0sq;
  bool active = true;
  if (KIND == FJ_KERNEL_CUTOFF) {
    const double sm = (r0x * r1x + r0y * r1y + r0z * r1z) / r0sq;
    const double px = r1x - sm * r0x, py = r1y - sm * r0y, pz = r1z - sm * r0z;
    active = (px * px + py * py + pz * pz) > kp.rcut * kp.rcut;
  }
  if (!active) {
    for (int k = 0; k < 3; ++k) { G1[k] = 0.0; G2[k] = 0.0; }
    wK = 0.0;
    return;
  }
  const double invD = 1.0 / D;
  const double phi = N * invD;
  const double P = wx * cx + wy * cy + wz * cz;
  wK = P * phi;
  const double b1 = a1 * a1 * a1 * (r0x * r1x + r0y * r1y + r0z * r1z);
  const double b2 = a2 * a2 * a2 * (r0x * r2x + r0y * r2y + r0z * r2z);
  const double dN1x = dx + a1 * r0x - b1 * r1x, dN1y = dy + a1 * r0y - b1 * r1y, dN1z = dz + a1 * r0z - b1 * r1z;
  const double dN2x = -dx - a2 * r0x + b2 * r2x, dN2y = -dy - a2 * r0y + b2 * r2y, dN2z = -dz - a2 * r0z + b2 * r2z;
  const double e4 = 4.0 * epsr * r0sq;
  // 2 r2 x c and 2 c x r1
  const double dD1x = 2.0 * (r2y * cz - r2z * cy) + e4 * r0x;
  const double dD1y = 2.0 * (r2z * cx - r2x * cz) + e4 * r0y;
  const double dD1z = 2.0 * (r2x * cy - r2y * cx) + e4 * r0z;
  const double dD2x = 2.0 * (cy * r1z - cz * r1y) - e4 * r0x;
  const double dD2y = 2.0 * (cz * r1x - cx * r1z) - e4 * r0y;
  const double dD2z = 2.0 * (cx * r1y - cy * r1x) - e4 * r0z;
  const double f = P * invD;
  G1[0] = phi * (r2y * wz - r2z * wy) + f * (dN1x - phi * dD1x);
  G1[1] = phi * (r2z * wx - r2x * wz) + f * (dN1y - phi * dD1y);
  G1[2] = phi * (r2x * wy - r2y * wx) + f * (dN1z - phi * dD1z);
  G2[0] = phi * (wy * r1z - wz * r1y) + f * (dN2x - phi * dD2x);
  G2[1] = phi * (wz * r1x - wx * r1z) + f * (dN2y - phi * dD2y);
  G2[2] = phi * (wx * r1y - wy * r1x) + f * (dN2z - phi * dD2z);
}

// dL/dx_t = (1/4pi) sum_s g_s (G1 + G2)
template <int KIND>
__global__ __launch_bounds__(FJ_BWD_THREADS) void k_bwd_targets(
    const double* __restrict__ seg, int n_seg, const double* __restrict__ tgt,
    const double* __restrict__ w, int n_tgt, FjKernelParams kp, double* __restrict__ gx,
    int accumulate) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int tc = t < n_tgt ? t : n_tgt - 1;
  const double x = tgt[3 * tc], y = tgt[3 * tc + 1], z = tgt[3 * tc + 2];
  const double wx = w[3 * tc], wy = w[3 * tc + 1], wz = w[3 * tc + 2];
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int s = 0; s < n_seg; ++s) {
    const double* sp = seg + (size_t)s * FJ_SEG_STRIDE;
    const double g = sp[10];
    double G1[3], G2[3], wK;
    fj_bs_vjp<KIND>(x - sp[0], y - sp[1], z - sp[2], x - sp[3], y - sp[4], z - sp[5], wx, wy, wz,
                    kp, G1, G2, wK);
    ax += g * (G1[0] + G2[0]);
    ay += g * (G1[1] + G2[1]);
    az += g * (G1[2] + G2[2]);
  }
  if (t < n_tgt) {
    if (accumulate) {
      gx[3 * t] += ax * FJ_INV_4PI; gx[3 * t + 1] += ay * FJ_INV_4PI; gx[3 * t + 2] += az * FJ_INV_4PI;
    } else {
      gx[3 * t] = ax * FJ_INV_4PI; gx[3 * t + 1] = ay * FJ_INV_4PI; gx[3 * t + 2] = az * FJ_INV_4PI;
    }
  }
}

// The same sum spread over the whole chip: workgroup (tile, chunk) takes 64
// targets (a lane each) against one contiguous chunk of filaments, its four
// waves taking every fourth filament of the chunk; the waves' sums meet in LDS
// in wave order and the chunk's partial (3 x n_tgt per chunk) is summed over
// chunks in chunk order by k_bwd_targets_reduce, so the result does not depend
// on scheduling.  A few thousand targets (the force midpoints of a lattice)
// otherwise occupy only a few dozen workgroups.
#define FJ_BWD_PART_WAVES 4
#define FJ_BWD_TARGET_BLOCKS 2048
#define FJ_BWD_MIN_CHUNK 256

template <int KIND>
__global__ __launch_bounds__(64 * FJ_BWD_PART_WAVES) void k_bwd_targets_part(
    const double* __restrict__ seg, int n_seg, const double* __restrict__ tgt,
    const double* __restrict__ w, int n_tgt, FjKernelParams kp, int chunk_len,
    double* __restrict__ part) {
  __shared__ double red[FJ_BWD_PART_WAVES][3][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t = blockIdx.x * 64 + lane;
  const int tc = t < n_tgt ? t : n_tgt - 1;
  const double x = tgt[3 * tc], y = tgt[3 * tc + 1], z = tgt[3 * tc + 2];
  const double wx = w[3 * tc], wy = w[3 * tc + 1], wz = w[3 * tc + 2];
  const int s0 = blockIdx.y * chunk_len;
  const int s1 = min(n_seg, s0 + chunk_len);
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int s = s0 + wave; s < s1; s += FJ_BWD_PART_WAVES) {
    const double* sp = seg + (size_t)s * FJ_SEG_STRIDE;
    const double g = sp[10];
    double G1[3], G2[3], wK;
    fj_bs_vjp<KIND>(x - sp[0], y - sp[1], z - sp[2], x - sp[3], y - sp[4], z - sp[5], wx, wy, wz,
                    kp, G1, G2, wK);
    ax += g * (G1[0] + G2[0]);
    ay += g * (G1[1] + G2[1]);
    az += g * (G1[2] + G2[2]);
  }
  red[wave][0][lane] = ax;
  red[wave][1][lane] = ay;
  red[wave][2][lane] = az;
  __syncthreads();
  if (wave == 0 && t < n_tgt) {
    for (int k = 0; k < 3; ++k) {
      double a = red[0][k][lane];
      for (int q = 1; q < FJ_BWD_PART_WAVES; ++q) a += red[q][k][lane];
      part[((size_t)blockIdx.y * 3 + k) * n_tgt + t] = a;
    }
  }
}

__global__ void k_bwd_targets_reduce(const double* __restrict__ part, int n_chunk, int n_tgt,
                                     double* __restrict__ gx, int accumulate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (target, component)
  if (i >= 3 * n_tgt) return;
  const int t = i / 3, k = i - 3 * t;
  double a = 0.0;
  for (int c = 0; c < n_chunk; ++c) a += part[((size_t)c * 3 + k) * n_tgt + t];
  a *= FJ_INV_4PI;
  gx[i] = accumulate ? gx[i] + a : a;
}

static int fj_bwd_chunks(int n_tgt, int n_seg) {
  const int tiles = fj_cdiv(n_tgt, 64);
  int chunks = fj_cdiv(FJ_BWD_TARGET_BLOCKS, tiles);
  chunks = min(chunks, max(1, fj_cdiv(n_seg, FJ_BWD_MIN_CHUNK)));
  return max(1, min(chunks, 65535));
}

extern "C" size_t fj_pair_sum_backward_workspace_bytes(int n_tgt, int n_seg) {
  if (n_tgt <= 0 || n_seg <= 0) return 256;
  return ((size_t)fj_bwd_chunks(n_tgt, n_seg) * 3 * n_tgt * sizeof(double) + 255) / 256 * 256;
}

// dL/dy0_s, dL/dy1_s (3 each) and dL/dg_s, one filament per lane
template <int KIND>
__global__ __launch_bounds__(FJ_BWD_THREADS) void k_bwd_segments(
    const double* __restrict__ seg, int n_seg, const double* __restrict__ tgt,
    const double* __restrict__ w, int n_tgt, FjKernelParams kp, double* __restrict__ gseg) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int sc = s < n_seg ? s : n_seg - 1;
  const double* sp = seg + (size_t)sc * FJ_SEG_STRIDE;
  const double y0x = sp[0], y0y = sp[1], y0z = sp[2], y1x = sp[3], y1y = sp[4], y1z = sp[5];
  const double g = sp[10];
  double a0[3] = {0, 0, 0}, a1[3] = {0, 0, 0}, ag = 0.0;
  for (int t = 0; t < n_tgt; ++t) {
    const double x = tgt[3 * t], y = tgt[3 * t + 1], z = tgt[3 * t + 2];
    double G1[3], G2[3], wK;
    fj_bs_vjp<KIND>(x - y0x, y - y0y, z - y0z, x - y1x, y - y1y, z - y1z, w[3 * t], w[3 * t + 1],
                    w[3 * t + 2], kp, G1, G2, wK);
    for (int k = 0; k < 3; ++k) { a0[k] -= G1[k]; a1[k] -= G2[k]; }
    ag += wK;
  }
  if (s < n_seg) {
    double* o = gseg + 7 * (size_t)s;
    for (int k = 0; k < 3; ++k) {
      o[k] = g * a0[k] * FJ_INV_4PI;
      o[3 + k] = g * a1[k] * FJ_INV_4PI;
    }
    o[6] = ag * FJ_INV_4PI;
  }
}

extern "C" int fj_pair_sum_backward(const double* seg, int n_seg, const double* targets,
                                    int n_tgt, const double* grad_v, int kernel, double eps,
                                    double rcut, double* grad_targets, int accumulate,
                                    double* grad_seg, void* workspace, void* stream) {
  if (n_seg < 0 || n_tgt < 0) return FJ_ERR_ARG;
  if (kernel < FJ_KERNEL_EPSILON || kernel > FJ_KERNEL_PLAIN) return FJ_ERR_ARG;
  if (n_seg == 0 || n_tgt == 0) return FJ_OK;
  if (!seg || !targets || !grad_v) return FJ_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  FjKernelParams kp;
  kp.eps = eps; kp.eps2 = eps * eps; kp.rcut = rcut;
  const int n_chunk = fj_bwd_chunks(n_tgt, n_seg);
  const int chunk_len = fj_cdiv(n_seg, n_chunk);
#define FJ_BWD(K)                                                                                  \
  do {                                                                                             \
    if (grad_targets && workspace) {                                                               \
      hipLaunchKernelGGL(k_bwd_targets_part<K>, dim3(fj_cdiv(n_tgt, 64), fj_cdiv(n_seg, chunk_len)),\
                         dim3(64 * FJ_BWD_PART_WAVES), 0, st, seg, n_seg, targets, grad_v, n_tgt,  \
                         kp, chunk_len, (double*)workspace);                                       \
      hipLaunchKernelGGL(k_bwd_targets_reduce, dim3(fj_cdiv(3 * n_tgt, 256)), dim3(256), 0, st,    \
                         (const double*)workspace, fj_cdiv(n_seg, chunk_len), n_tgt, grad_targets, \
                         accumulate);                                                              \
    } else if (grad_targets)                                                                       \
      hipLaunchKernelGGL(k_bwd_targets<K>, dim3(fj_cdiv(n_tgt, FJ_BWD_THREADS)),                   \
                         dim3(FJ_BWD_THREADS), 0, st, seg, n_seg, targets, grad_v, n_tgt, kp,      \
                         grad_targets, accumulate);                                                \
    if (grad_seg)                                                                                  \
      hipLaunchKernelGGL(k_bwd_segments<K>, dim3(fj_cdiv(n_seg, FJ_BWD_THREADS)),                  \
                         dim3(FJ_BWD_THREADS), 0, st, seg, n_seg, targets, grad_v, n_tgt, kp,      \
                         grad_seg);                                                                \
  } while (0)
  if (kernel == FJ_KERNEL_EPSILON) FJ_BWD(FJ_KERNEL_EPSILON);
  else if (kernel == FJ_KERNEL_CUTOFF) FJ_BWD(FJ_KERNEL_CUTOFF);
  else FJ_BWD(FJ_KERNEL_PLAIN);
#undef FJ_BWD
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}
