// K1, mixed form: FP32 far field, FP64 near field, FP64 accumulation.
//
// r1 = x - y0 and r2 = x - y1 are formed in FP64 (so the subtraction of two
// nearby absolute coordinates loses nothing) and rounded to FP32; the rest of
// the far-field kernel runs in FP32 with the hardware rsqrt / rcp.  A pair is
// "near" when the target lies within FJ_NEAR_FACTOR filament lengths of an
// end point; near pairs (where 1/r^2 cancellation and the regularisation
// matter) take the FP64 evaluation of the plain path.  FP32 contributions are
// summed in FP32 for FJ_MIXED_FLUSH filaments at a time and then folded into
// FP64 accumulators.  Launch shape and reduction are those of the FP64 path.
#include "fj_common.hpp"

template <int KIND, int TPL>
__global__ __launch_bounds__(FJ_BS_THREADS) void k_pair_partial_mixed(
    const double* __restrict__ seg, int n_seg, int chunk_len, const double* __restrict__ tgt,
    int n_tgt, int ld, double* __restrict__ part, FjKernelParams kp) {
  __shared__ double red[FJ_BS_WAVES - 1][3][64 * TPL];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tile, chunk;
  fj_tile_chunk(tile, chunk);
  const int t0 = tile * 64 * TPL;
  const float eps2f = (float)kp.eps2;
  const float rcut2f = (float)(kp.rcut * kp.rcut);

  double x[TPL], y[TPL], z[TPL], ax[TPL], ay[TPL], az[TPL];
  float bx[TPL], by[TPL], bz[TPL];
#pragma unroll
  for (int p = 0; p < TPL; ++p) {
    int t = t0 + p * 64 + lane;
    t = (t < n_tgt) ? t : (n_tgt - 1);
    x[p] = tgt[3 * t + 0];
    y[p] = tgt[3 * t + 1];
    z[p] = tgt[3 * t + 2];
    ax[p] = 0.0; ay[p] = 0.0; az[p] = 0.0;
    bx[p] = 0.0f; by[p] = 0.0f; bz[p] = 0.0f;
  }

  const int s_begin = chunk * chunk_len;
  const int s_end = min(n_seg, s_begin + chunk_len);
  int since_flush = 0;
  for (int s = s_begin + wave; s < s_end; s += FJ_BS_WAVES) {
    const double* sp = seg + (size_t)s * FJ_SEG_STRIDE;
    double rec[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) rec[k] = sp[k];
    const float* spf = (const float*)(sp + 12);
    float sf[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) sf[k] = spf[k];
    const float r0sq_f = (float)rec[11];
    // near: within FJ_NEAR_FACTOR filament lengths of the start point, which
    // (|r0| apart) also covers the end point's neighbourhood
    const float near2 = FJ_NEAR_FACTOR2 * r0sq_f;
#pragma unroll
    for (int p = 0; p < TPL; ++p) {
      const float r1x = (float)(x[p] - rec[0]), r1y = (float)(y[p] - rec[1]),
                  r1z = (float)(z[p] - rec[2]);
      const float s1 = r1x * r1x + r1y * r1y + r1z * r1z;
      if (s1 < near2) {
        double vx, vy, vz;
        fj_bs_eval<KIND>(x[p], y[p], z[p], rec, kp, vx, vy, vz);
        ax[p] += vx; ay[p] += vy; az[p] += vz;
      } else {
        float vx, vy, vz;
        fj_bs_eval_f32<KIND>(r1x, r1y, r1z, s1, sf, r0sq_f, eps2f, rcut2f, vx, vy, vz);
        bx[p] += vx; by[p] += vy; bz[p] += vz;
      }
    }
    if (++since_flush == FJ_MIXED_FLUSH) {
      since_flush = 0;
#pragma unroll
      for (int p = 0; p < TPL; ++p) {
        ax[p] += (double)bx[p]; ay[p] += (double)by[p]; az[p] += (double)bz[p];
        bx[p] = 0.0f; by[p] = 0.0f; bz[p] = 0.0f;
      }
    }
  }
#pragma unroll
  for (int p = 0; p < TPL; ++p) {
    ax[p] += (double)bx[p]; ay[p] += (double)by[p]; az[p] += (double)bz[p];
  }

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

template <int KIND, int TPL>
static void launch_mixed(const FjPairPlan* p, const double* seg, int n_seg, const double* tgt,
                         int n_tgt, double* part, const FjKernelParams* kp, hipStream_t st) {
  hipLaunchKernelGGL((k_pair_partial_mixed<KIND, TPL>), dim3(p->tiles, p->chunks),
                     dim3(FJ_BS_THREADS), 0, st, seg, n_seg, p->chunk_len, tgt, n_tgt, p->ld,
                     part, *kp);
}

int fj_pair_partial_mixed(const FjPairPlan* p, const double* seg, int n_seg, const double* tgt,
                          int n_tgt, double* part, int kernel, const FjKernelParams* kp,
                          hipStream_t st) {
#define FJ_DISPATCH_TPL(K)                                           \
  if (p->tpl == 2) launch_mixed<K, 2>(p, seg, n_seg, tgt, n_tgt, part, kp, st); \
  else launch_mixed<K, 1>(p, seg, n_seg, tgt, n_tgt, part, kp, st);
  if (kernel == FJ_KERNEL_EPSILON) { FJ_DISPATCH_TPL(FJ_KERNEL_EPSILON) }
  else if (kernel == FJ_KERNEL_CUTOFF) { FJ_DISPATCH_TPL(FJ_KERNEL_CUTOFF) }
  else { FJ_DISPATCH_TPL(FJ_KERNEL_PLAIN) }
#undef FJ_DISPATCH_TPL
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}
