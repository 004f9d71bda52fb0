// Filament-record construction shared by the table kernels (fj_biot_savart.hip)
// and the batched serving tick (fj_serving.hip).
#pragma once
#include "fj_common.hpp"

// record of filament s of grid (m+1, n+1) with ring strengths gamma (or unit
// strength when gamma is NULL), optionally reflected in the mirror plane
__device__ __forceinline__ void fj_segment_record(const double* __restrict__ zeta,
                                                  const double* __restrict__ gamma, int m,
                                                  int n, int s, int mirror, const double* mp,
                                                  const double* mn, int kernel, double eps,
                                                  double* __restrict__ o) {
  const int n_ch = m * (n + 1);
  int ia, ja, ib, jb;
  double g = 0.0;
  if (s < n_ch) {
    // chordwise filament (i, j) -> (i+1, j); g = G[i,j] - G[i,j-1] (aero/aic.py:284)
    const int i = s / (n + 1), j = s % (n + 1);
    ia = i; ja = j; ib = i + 1; jb = j;
    if (gamma) g = (j < n ? gamma[i * n + j] : 0.0) - (j > 0 ? gamma[i * n + j - 1] : 0.0);
  } else {
    // spanwise filament (i, j) -> (i, j+1); g = G[i-1,j] - G[i,j] (aero/aic.py:285-287)
    const int r = s - n_ch;
    const int i = r / n, j = r % n;
    ia = i; ja = j; ib = i; jb = j + 1;
    if (gamma) g = (i > 0 ? gamma[(i - 1) * n + j] : 0.0) - (i < m ? gamma[i * n + j] : 0.0);
  }
  const double* a = zeta + 3 * (ia * (n + 1) + ja);
  const double* b = zeta + 3 * (ib * (n + 1) + jb);
  double y0[3] = {a[0], a[1], a[2]};
  double y1[3] = {b[0], b[1], b[2]};
  if (mirror) {
    // aero/utils.py:651-663, then the image is subtracted (aero/aic.py:306-308)
    const double d0 = (y0[0] - mp[0]) * mn[0] + (y0[1] - mp[1]) * mn[1] + (y0[2] - mp[2]) * mn[2];
    const double d1 = (y1[0] - mp[0]) * mn[0] + (y1[1] - mp[1]) * mn[1] + (y1[2] - mp[2]) * mn[2];
    y0[0] -= 2.0 * d0 * mn[0]; y0[1] -= 2.0 * d0 * mn[1]; y0[2] -= 2.0 * d0 * mn[2];
    y1[0] -= 2.0 * d1 * mn[0]; y1[1] -= 2.0 * d1 * mn[1]; y1[2] -= 2.0 * d1 * mn[2];
    g = -g;
  }
  const double r0x = y1[0] - y0[0], r0y = y1[1] - y0[1], r0z = y1[2] - y0[2];
  const double r0sq = r0x * r0x + r0y * r0y + r0z * r0z;
  o[0] = y0[0]; o[1] = y0[1]; o[2] = y0[2];
  o[3] = y1[0]; o[4] = y1[1]; o[5] = y1[2];
  o[6] = r0x; o[7] = r0y; o[8] = r0z;
  const double aux = (kernel == FJ_KERNEL_CUTOFF) ? 1.0 / r0sq : eps * r0sq * r0sq;
  o[9] = aux;
  o[10] = g;
  o[11] = r0sq;
  float* f = (float*)(o + 12);
  f[0] = (float)r0x; f[1] = (float)r0y; f[2] = (float)r0z;
  f[3] = (float)aux; f[4] = (float)g;
  f[5] = 0.0f; f[6] = 0.0f; f[7] = 0.0f;
}

