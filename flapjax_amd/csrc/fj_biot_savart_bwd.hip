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
// With a workspace: one fused pass over the pairs (k_bwd_fused, below) and
// two fixed-order reductions.  Without one: a target-major pass (dL/dx, a
// thread per target) and a filament-major pass (dL/dy0, dL/dy1, dL/dg, a
// thread per filament), which leave most of the chip idle on a few thousand
// targets or filaments.
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

// ---------------------------------------------------------------------------
// One pass for both sides of the pair sum (the form used with a workspace).
//
// Lane = filament: each lane keeps its filament's constants (end points,
// r0 = y1 - y0, |r0|^2, g) and its three gradients (dL/dy0, dL/dy1, dL/dg)
// in registers over a chunk of targets that every lane of the wave reads at
// once (scalar loads), so the filament side needs no reduction.  The target
// side, g (G1 + G2), is summed over the wave's 64 filaments by a transposed
// butterfly over FJ_BWDF_TG targets at a time: two exchange steps that each
// halve the values a lane holds (lanes with the mask bit set keep the upper
// half), then four plain steps over the remaining three components; lane l
// then holds target (l >> 4) of the group.  A workgroup of FJ_BWDF_WAVES
// waves puts its waves' target sums in LDS per sub-chunk of FJ_BWDF_TS
// targets and writes their sum in wave order, one partial per (filament
// tile, target); the filament partials are one per (target chunk,
// filament).  Both partial sets are summed in a fixed order by the reduce
// kernels, so results do not depend on scheduling.
//
// Per pair: the same derivatives as fj_bs_vjp, with the reciprocal square
// roots and the reciprocal of D from the hardware estimates refined by one
// third-order step (a few ulp), where fj_bs_vjp divides.
#define FJ_BWDF_WAVES 4
#define FJ_BWDF_TG 4
#define FJ_BWDF_TS 64
#define FJ_BWDF_BLOCKS 2048

__device__ __forceinline__ double fj_rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  const double e = fma(-x, y, 1.0);
  return fma(y, fma(e, e, e), y);
}

template <int KIND>
__global__ __launch_bounds__(64 * FJ_BWDF_WAVES, 2) void k_bwd_fused(
    const double* __restrict__ seg, int n_seg, const double* __restrict__ tgt,
    const double* __restrict__ w, int n_tgt, FjKernelParams kp, int tchunk,
    double* __restrict__ part_seg, double* __restrict__ part_tgt) {
  __shared__ double st[FJ_BWDF_WAVES][FJ_BWDF_TS][3];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int s = (blockIdx.x * FJ_BWDF_WAVES + wave) * 64 + lane;
  const bool live = s < n_seg;
  const double* sp = seg + (size_t)(live ? s : n_seg - 1) * FJ_SEG_STRIDE;
  const double y0x = sp[0], y0y = sp[1], y0z = sp[2];
  const double y1x = sp[3], y1y = sp[4], y1z = sp[5];
  const double r0x = sp[6], r0y = sp[7], r0z = sp[8];
  const double aux = sp[9], r0sq = sp[11];
  const double g = live ? sp[10] : 0.0;
  const double e2 = (KIND == FJ_KERNEL_PLAIN) ? 0.0 : kp.eps2;
  const double e4 = (KIND == FJ_KERNEL_EPSILON) ? 4.0 * kp.eps * r0sq : 0.0;
  const double rcut2 = kp.rcut * kp.rcut;
  double a0x = 0.0, a0y = 0.0, a0z = 0.0, a1x = 0.0, a1y = 0.0, a1z = 0.0, ag = 0.0;
  const int t0 = blockIdx.y * tchunk;
  const int t1 = min(n_tgt, t0 + tchunk);
  for (int tb = t0; tb < t1; tb += FJ_BWDF_TS) {
    const int te = min(t1, tb + FJ_BWDF_TS);
    for (int tg = tb; tg < te; tg += FJ_BWDF_TG) {
      double v[3 * FJ_BWDF_TG];
#pragma unroll
      for (int j = 0; j < FJ_BWDF_TG; ++j) {
        const int t = tg + j;
        if (t < te) {
          const double x = tgt[3 * t], y = tgt[3 * t + 1], z = tgt[3 * t + 2];
          const double wx = w[3 * t], wy = w[3 * t + 1], wz = w[3 * t + 2];
          const double r1x = x - y0x, r1y = y - y0y, r1z = z - y0z;
          const double r2x = x - y1x, r2y = y - y1y, r2z = z - y1z;
          const double cx = r1y * r2z - r1z * r2y;
          const double cy = r1z * r2x - r1x * r2z;
          const double cz = r1x * r2y - r1y * r2x;
          const double s1 = fma(r1z, r1z, fma(r1y, r1y, fma(r1x, r1x, e2)));
          const double s2 = fma(r2z, r2z, fma(r2y, r2y, fma(r2x, r2x, e2)));
          const double ia1 = fj_rsqrt(s1), ia2 = fj_rsqrt(s2);
          const double dx = ia1 * r1x - ia2 * r2x, dy = ia1 * r1y - ia2 * r2y, dz = ia1 * r1z - ia2 * r2z;
          const double N = r0x * dx + r0y * dy + r0z * dz;
          const double D = fma(cz, cz, fma(cy, cy, fma(cx, cx, KIND == FJ_KERNEL_EPSILON ? aux : 0.0)));
          const double invD = fj_rcp(D);
          const double phi = N * invD;
          const double P = wx * cx + wy * cy + wz * cz;
          const double q1 = r0x * r1x + r0y * r1y + r0z * r1z;
          const double q2 = q1 - r0sq;
          const double b1 = ia1 * ia1 * ia1 * q1, b2 = ia2 * ia2 * ia2 * q2;
          const double f = P * invD;
          // G1 = phi (r2 x w) + f (dN1 - phi dD1), G2 = phi (w x r1) + f (dN2 - phi dD2)
          // dN1 = d + a1 r0 - b1 r1, dN2 = -d - a2 r0 + b2 r2,
          // dD1 = 2 r2 x c + e4 r0, dD2 = 2 c x r1 - e4 r0
          const double dD1x = 2.0 * (r2y * cz - r2z * cy) + e4 * r0x;
          const double dD1y = 2.0 * (r2z * cx - r2x * cz) + e4 * r0y;
          const double dD1z = 2.0 * (r2x * cy - r2y * cx) + e4 * r0z;
          const double dD2x = 2.0 * (cy * r1z - cz * r1y) - e4 * r0x;
          const double dD2y = 2.0 * (cz * r1x - cx * r1z) - e4 * r0y;
          const double dD2z = 2.0 * (cx * r1y - cy * r1x) - e4 * r0z;
          double G1x = phi * (r2y * wz - r2z * wy) + f * ((dx + ia1 * r0x - b1 * r1x) - phi * dD1x);
          double G1y = phi * (r2z * wx - r2x * wz) + f * ((dy + ia1 * r0y - b1 * r1y) - phi * dD1y);
          double G1z = phi * (r2x * wy - r2y * wx) + f * ((dz + ia1 * r0z - b1 * r1z) - phi * dD1z);
          double G2x = phi * (wy * r1z - wz * r1y) + f * ((-dx - ia2 * r0x + b2 * r2x) - phi * dD2x);
          double G2y = phi * (wz * r1x - wx * r1z) + f * ((-dy - ia2 * r0y + b2 * r2y) - phi * dD2y);
          double G2z = phi * (wx * r1y - wy * r1x) + f * ((-dz - ia2 * r0z + b2 * r2z) - phi * dD2z);
          double wK = P * phi;
          if (KIND == FJ_KERNEL_CUTOFF) {
            // squared distance from x to the filament line, as the forward pass
            const double rad2 = (s1 - e2) - q1 * (q1 * aux);
            const bool on = rad2 > rcut2;
            G1x = on ? G1x : 0.0; G1y = on ? G1y : 0.0; G1z = on ? G1z : 0.0;
            G2x = on ? G2x : 0.0; G2y = on ? G2y : 0.0; G2z = on ? G2z : 0.0;
            wK = on ? wK : 0.0;
          }
          a0x -= G1x; a0y -= G1y; a0z -= G1z;
          a1x -= G2x; a1y -= G2y; a1z -= G2z;
          ag += wK;
          v[3 * j] = g * (G1x + G2x);
          v[3 * j + 1] = g * (G1y + G2y);
          v[3 * j + 2] = g * (G1z + G2z);
        } else {
          v[3 * j] = 0.0; v[3 * j + 1] = 0.0; v[3 * j + 2] = 0.0;
        }
      }
      // transposed steps: 12 -> 6 (mask 32), 6 -> 3 (mask 16)
      static_assert(FJ_BWDF_TG == 4, "butterfly written for four targets");
      {
        const bool up = lane & 32;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
          const double lo = v[i], hi = v[6 + i];
          const double send = up ? lo : hi;
          v[i] = (up ? hi : lo) + __shfl_xor(send, 32);
        }
      }
      {
        const bool up = lane & 16;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double lo = v[i], hi = v[3 + i];
          const double send = up ? lo : hi;
          v[i] = (up ? hi : lo) + __shfl_xor(send, 16);
        }
      }
#pragma unroll
      for (int m = 8; m >= 1; m >>= 1) {
#pragma unroll
        for (int i = 0; i < 3; ++i) v[i] += __shfl_xor(v[i], m);
      }
      const int j = lane >> 4, k = lane & 15;
      if (k < 3 && tg + j < te) st[wave][tg - tb + j][k] = k == 0 ? v[0] : (k == 1 ? v[1] : v[2]);
    }
    __syncthreads();
    if (part_tgt) {
      for (int q = threadIdx.x; q < 3 * (te - tb); q += 64 * FJ_BWDF_WAVES) {
        const int tt = q / 3, k = q - 3 * tt;
        double a = st[0][tt][k];
#pragma unroll
        for (int ww = 1; ww < FJ_BWDF_WAVES; ++ww) a += st[ww][tt][k];
        part_tgt[((size_t)blockIdx.x * n_tgt + tb) * 3 + q] = a;
      }
    }
    __syncthreads();
  }
  if (part_seg && live) {
    double* o = part_seg + (size_t)blockIdx.y * 7 * n_seg + s;
    o[0] = a0x; o[(size_t)n_seg] = a0y; o[2 * (size_t)n_seg] = a0z;
    o[3 * (size_t)n_seg] = a1x; o[4 * (size_t)n_seg] = a1y; o[5 * (size_t)n_seg] = a1z;
    o[6 * (size_t)n_seg] = ag;
  }
}

// dL/dx from the (filament tile, target) partials: a workgroup per 64
// outputs, its FJ_BWDR_WAVES waves summing interleaved tiles (wave w: tiles
// w, w + W, ...) in order, then the wave sums added in wave order (a fixed
// order, so the result does not depend on scheduling)
#define FJ_BWDR_WAVES 8
__global__ __launch_bounds__(64 * FJ_BWDR_WAVES) void k_bwd_fused_reduce_tgt(
    const double* __restrict__ part, int n_tile, int n_tgt, double* __restrict__ gx, int accumulate) {
  __shared__ double red[FJ_BWDR_WAVES][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;  // (target, component)
  const int n_out = 3 * n_tgt;
  double a = 0.0;
  if (i < n_out)
    for (int b = wave; b < n_tile; b += FJ_BWDR_WAVES) a += part[(size_t)b * n_out + i];
  red[wave][lane] = a;
  __syncthreads();
  if (wave == 0 && i < n_out) {
    double t = red[0][lane];
#pragma unroll
    for (int w = 1; w < FJ_BWDR_WAVES; ++w) t += red[w][lane];
    t *= FJ_INV_4PI;
    gx[i] = accumulate ? gx[i] + t : t;
  }
}

// [dL/dy0, dL/dy1, dL/dg] from the (target chunk, filament) partials
__global__ void k_bwd_fused_reduce_seg(const double* __restrict__ part, int n_chunk, int n_seg,
                                       const double* __restrict__ seg, double* __restrict__ gseg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (component, filament)
  if (i >= 7 * n_seg) return;
  const int k = i / n_seg, s = i - k * n_seg;
  double a = 0.0;
  for (int c = 0; c < n_chunk; ++c) a += part[(size_t)c * 7 * n_seg + i];
  const double g = seg[(size_t)s * FJ_SEG_STRIDE + 10];
  gseg[7 * (size_t)s + k] = (k < 6 ? g * a : a) * FJ_INV_4PI;
}

struct FjBwdPlan {
  int tiles;    // filament tiles of 64 * FJ_BWDF_WAVES (grid x)
  int chunks;   // target chunks (grid y)
  int tchunk;   // targets per chunk (a multiple of FJ_BWDF_TG)
  size_t tgt_bytes, seg_bytes;
};

static FjBwdPlan fj_bwd_plan(int n_tgt, int n_seg) {
  FjBwdPlan p;
  p.tiles = fj_cdiv(n_seg, 64 * FJ_BWDF_WAVES);
  int chunks = fj_cdiv(FJ_BWDF_BLOCKS, p.tiles);
  chunks = max(1, min(chunks, fj_cdiv(n_tgt, FJ_BWDF_TS)));
  chunks = min(chunks, 65535);
  p.tchunk = fj_cdiv(fj_cdiv(n_tgt, chunks), FJ_BWDF_TG) * FJ_BWDF_TG;
  p.chunks = fj_cdiv(n_tgt, p.tchunk);
  p.tgt_bytes = ((size_t)p.tiles * 3 * n_tgt * sizeof(double) + 255) / 256 * 256;
  p.seg_bytes = ((size_t)p.chunks * 7 * n_seg * sizeof(double) + 255) / 256 * 256;
  return p;
}

extern "C" size_t fj_pair_sum_backward_workspace_bytes(int n_tgt, int n_seg) {
  if (n_tgt <= 0 || n_seg <= 0) return 256;
  const FjBwdPlan p = fj_bwd_plan(n_tgt, n_seg);
  return p.tgt_bytes + p.seg_bytes;
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
  const FjBwdPlan p = fj_bwd_plan(n_tgt, n_seg);
  double* part_tgt = (double*)workspace;
  double* part_seg = workspace ? (double*)((char*)workspace + p.tgt_bytes) : nullptr;
#define FJ_BWD(K)                                                                                  \
  do {                                                                                             \
    if (workspace) {                                                                               \
      hipLaunchKernelGGL(k_bwd_fused<K>, dim3(p.tiles, p.chunks), dim3(64 * FJ_BWDF_WAVES), 0, st, \
                         seg, n_seg, targets, grad_v, n_tgt, kp, p.tchunk,                          \
                         grad_seg ? part_seg : nullptr, grad_targets ? part_tgt : nullptr);         \
      if (grad_targets)                                                                            \
        hipLaunchKernelGGL(k_bwd_fused_reduce_tgt, dim3(fj_cdiv(3 * n_tgt, 64)),                   \
                           dim3(64 * FJ_BWDR_WAVES), 0, st,                                        \
                           (const double*)part_tgt, p.tiles, n_tgt, grad_targets, accumulate);     \
      if (grad_seg)                                                                                \
        hipLaunchKernelGGL(k_bwd_fused_reduce_seg, dim3(fj_cdiv(7 * (long long)n_seg, 256)),       \
                           dim3(256), 0, st, (const double*)part_seg, p.chunks, n_seg, seg,         \
                           grad_seg);                                                              \
    } else {                                                                                       \
      if (grad_targets)                                                                            \
        hipLaunchKernelGGL(k_bwd_targets<K>, dim3(fj_cdiv(n_tgt, FJ_BWD_THREADS)),                 \
                           dim3(FJ_BWD_THREADS), 0, st, seg, n_seg, targets, grad_v, n_tgt, kp,    \
                           grad_targets, accumulate);                                              \
      if (grad_seg)                                                                                \
        hipLaunchKernelGGL(k_bwd_segments<K>, dim3(fj_cdiv(n_seg, FJ_BWD_THREADS)),                \
                           dim3(FJ_BWD_THREADS), 0, st, seg, n_seg, targets, grad_v, n_tgt, kp,    \
                           grad_seg);                                                              \
    }                                                                                              \
  } while (0)
  if (kernel == FJ_KERNEL_EPSILON) FJ_BWD(FJ_KERNEL_EPSILON);
  else if (kernel == FJ_KERNEL_CUTOFF) FJ_BWD(FJ_KERNEL_CUTOFF);
  else FJ_BWD(FJ_KERNEL_PLAIN);
#undef FJ_BWD
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}
