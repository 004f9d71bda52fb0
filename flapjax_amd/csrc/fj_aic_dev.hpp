// Ring-influence evaluation of the AIC kernels (fj_aic.hip), shared with the
// batched serving tick (fj_serving.hip).
#pragma once
#include "fj_common.hpp"

template <int KIND>
__device__ __forceinline__ double fj_proj(double x, double y, double z, const double* a,
                                          const double* b, double nx, double ny, double nz,
                                          const FjKernelParams& kp, int kernel) {
  double s[FJ_SEG_STRIDE];
  s[0] = a[0]; s[1] = a[1]; s[2] = a[2];
  s[3] = b[0]; s[4] = b[1]; s[5] = b[2];
  s[6] = b[0] - a[0]; s[7] = b[1] - a[1]; s[8] = b[2] - a[2];
  const double r0sq = s[6] * s[6] + s[7] * s[7] + s[8] * s[8];
  s[9] = (KIND == FJ_KERNEL_CUTOFF) ? 1.0 / r0sq : kp.eps * r0sq * r0sq;
  s[10] = FJ_INV_4PI;
  s[11] = r0sq;
  double vx, vy, vz;
  fj_bs_eval<KIND>(x, y, z, s, kp, vx, vy, vz);
  return vx * nx + vy * ny + vz * nz;
}

__device__ __forceinline__ void fj_reflect(const double* p, const double* mp, const double* mn,
                                           double* o) {
  const double d = (p[0] - mp[0]) * mn[0] + (p[1] - mp[1]) * mn[1] + (p[2] - mp[2]) * mn[2];
  o[0] = p[0] - 2.0 * d * mn[0];
  o[1] = p[1] - 2.0 * d * mn[1];
  o[2] = p[2] - 2.0 * d * mn[2];
}

// ring (i, j) of grid g with n+1 vertex columns
template <int KIND>
__device__ __forceinline__ double fj_ring(double x, double y, double z, double nx, double ny,
                                          double nz, const double* g, int n, int i, int j,
                                          int mirror, const double* mp, const double* mn,
                                          const FjKernelParams& kp) {
  double v00[3], v01[3], v10[3], v11[3];
  const double* p00 = g + 3 * (i * (n + 1) + j);
  const double* p01 = p00 + 3;
  const double* p10 = g + 3 * ((i + 1) * (n + 1) + j);
  const double* p11 = p10 + 3;
  if (mirror) {
    fj_reflect(p00, mp, mn, v00); fj_reflect(p01, mp, mn, v01);
    fj_reflect(p10, mp, mn, v10); fj_reflect(p11, mp, mn, v11);
  } else {
    for (int k = 0; k < 3; ++k) { v00[k] = p00[k]; v01[k] = p01[k]; v10[k] = p10[k]; v11[k] = p11[k]; }
  }
  // chordwise filaments at columns j and j+1, spanwise at rows i and i+1
  const double m_j = fj_proj<KIND>(x, y, z, v00, v10, nx, ny, nz, kp, KIND);
  const double m_j1 = fj_proj<KIND>(x, y, z, v01, v11, nx, ny, nz, kp, KIND);
  const double n_i = fj_proj<KIND>(x, y, z, v00, v01, nx, ny, nz, kp, KIND);
  const double n_i1 = fj_proj<KIND>(x, y, z, v10, v11, nx, ny, nz, kp, KIND);
  return -(m_j1 - m_j) + (n_i1 - n_i);
}

