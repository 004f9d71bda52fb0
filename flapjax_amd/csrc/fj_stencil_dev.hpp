// Device helpers of the stencil, convection and force kernels, shared by
// the per-surface entry points (fj_stencil.hip) and the batched serving tick
// (fj_serving.hip), so both evaluate every quantity with the same arithmetic.
#pragma once
#include "fj_common.hpp"

// vertex v of the grid mapped by the (4, 4) pose of its beam node
__device__ __forceinline__ void fj_map_vertex(const double* __restrict__ hg,
                                              const int32_t* __restrict__ dof,
                                              const double* __restrict__ z0, int n, int v,
                                              double* o) {
  const double* p = z0 + 3 * v;
  const double* h = hg + 16 * (size_t)dof[v % (n + 1)];
  for (int a = 0; a < 3; ++a)
    o[a] = (h[4 * a] * p[0] + h[4 * a + 1] * p[1] + h[4 * a + 2] * p[2]) + h[4 * a + 3];
}

// c = neighbour_average(zeta, (-3, -2)) (aero/utils.py:87-93,
// algebra/array_utils.py:106-135): average along the chordwise axis first.
__device__ __forceinline__ double fj_quad_avg(double a00, double a10, double a01, double a11) {
  return 0.5 * (0.5 * (a00 + a10) + 0.5 * (a01 + a11));
}

__device__ __forceinline__ void fj_filament_ends(int m, int n, int s, int& a, int& b) {
  const int n_ch = m * (n + 1);
  if (s < n_ch) {
    const int i = s / (n + 1), j = s % (n + 1);
    a = i * (n + 1) + j; b = a + (n + 1);
  } else {
    const int r = s - n_ch, i = r / n, j = r % n;
    a = i * (n + 1) + j; b = a + 1;
  }
}

// filament circulations: chordwise Gc[i,j] = G[i,j] - G[i,j-1] (zero padded),
// spanwise Gs[i,j] = -G[i,j] + G[i-1,j], TE spanwise row minus the first wake row.
__device__ __forceinline__ void fj_fil_force(const double* __restrict__ z,
                                             const double* __restrict__ zd, double g,
                                             const double* __restrict__ v, int a, int b,
                                             const FjFlow& flow, double time, double rho,
                                             double* f) {
  double mp[3], u[3], vr[3], r[3];
  for (int k = 0; k < 3; ++k) mp[k] = 0.5 * (z[3 * a + k] + z[3 * b + k]);
  if (flow.kind >= 0) {
    fj_flow_eval(flow, mp[0], mp[1], mp[2], time, u);
  } else {
    u[0] = 0.0; u[1] = 0.0; u[2] = 0.0;
  }
  for (int k = 0; k < 3; ++k) {
    vr[k] = u[k] + v[k];
    if (zd) vr[k] -= 0.5 * (zd[3 * a + k] + zd[3 * b + k]);
    r[k] = z[3 * b + k] - z[3 * a + k];
  }
  f[0] = rho * (g * (vr[1] * r[2] - vr[2] * r[1]));
  f[1] = rho * (g * (vr[2] * r[0] - vr[0] * r[2]));
  f[2] = rho * (g * (vr[0] * r[1] - vr[1] * r[0]));
}

__device__ __forceinline__ void fj_vertex_steady_force(
    const double* __restrict__ z, const double* __restrict__ zd, const double* __restrict__ G,
    const double* __restrict__ gw0, const double* __restrict__ vmid, int m, int n,
    const FjFlow& flow, double time, double rho, int v, double* __restrict__ fo) {
  const int i = v / (n + 1), j = v % (n + 1);
  const int n_ch = m * (n + 1);
  double fc_lo[3] = {0, 0, 0}, fc_hi[3] = {0, 0, 0}, fs_lo[3] = {0, 0, 0}, fs_hi[3] = {0, 0, 0};
  // chordwise filaments (i-1, j) and (i, j) in column j
  if (i < m) {
    const double g = (j < n ? G[i * n + j] : 0.0) - (j > 0 ? G[i * n + j - 1] : 0.0);
    const int s = i * (n + 1) + j;
    fj_fil_force(z, zd, g, vmid + 3 * s, i * (n + 1) + j, (i + 1) * (n + 1) + j, flow, time, rho, fc_hi);
  }
  if (i > 0) {
    const int ii = i - 1;
    const double g = (j < n ? G[ii * n + j] : 0.0) - (j > 0 ? G[ii * n + j - 1] : 0.0);
    const int s = ii * (n + 1) + j;
    fj_fil_force(z, zd, g, vmid + 3 * s, ii * (n + 1) + j, i * (n + 1) + j, flow, time, rho, fc_lo);
  }
  // spanwise filaments (i, j-1) and (i, j) in row i
  if (j < n) {
    double g = (i < m ? -G[i * n + j] : 0.0) + (i > 0 ? G[(i - 1) * n + j] : 0.0);
    if (i == m && gw0) g += -gw0[j];
    const int s = n_ch + i * n + j;
    fj_fil_force(z, zd, g, vmid + 3 * s, i * (n + 1) + j, i * (n + 1) + j + 1, flow, time, rho, fs_hi);
  }
  if (j > 0) {
    const int jj = j - 1;
    double g = (i < m ? -G[i * n + jj] : 0.0) + (i > 0 ? G[(i - 1) * n + jj] : 0.0);
    if (i == m && gw0) g += -gw0[jj];
    const int s = n_ch + i * n + jj;
    fj_fil_force(z, zd, g, vmid + 3 * s, i * (n + 1) + jj, i * (n + 1) + j, flow, time, rho, fs_lo);
  }
  // split_to_vertex (algebra/array_utils.py:331-366): 0.5 * own + 0.5 * previous
  for (int k = 0; k < 3; ++k) {
    const double sc = 0.5 * fc_hi[k] + 0.5 * fc_lo[k];
    const double ss = 0.5 * fs_hi[k] + 0.5 * fs_lo[k];
    fo[3 * v + k] = sc + ss;
  }
}

__device__ __forceinline__ double fj_gamma_rate(const double* __restrict__ gn,
                                                const double* __restrict__ gp,
                                                const double* __restrict__ gdp, int p, double dt,
                                                double relax) {
  return (gn[p] - gp[p]) / dt * relax + gdp[p] * (1.0 - relax);
}

// f_unsteady at vertex v = split_to_vertex(rho Gamma_dot n) (aero/uvlm.py:1093-1103);
// gd(p) gives the circulation rate of panel p
template <typename GD>
__device__ __forceinline__ void fj_vertex_unsteady_force(GD gd, const double* __restrict__ nc,
                                                         int m, int n, double rho, int v,
                                                         double* __restrict__ fo) {
  const int i = v / (n + 1), j = v % (n + 1);
  double r[2][2];  // rho * Gamma_dot of panels (i-1 or i, j-1 or j)
  for (int a = 0; a < 2; ++a)
    for (int q = 0; q < 2; ++q) {
      const int ii = i - 1 + a, jj = j - 1 + q;
      r[a][q] = (ii >= 0 && ii < m && jj >= 0 && jj < n) ? rho * gd(ii * n + jj) : 0.0;
    }
  // split along the chordwise axis, then along the spanwise axis
  for (int k = 0; k < 3; ++k) {
    double s0[2];  // chordwise-split value at panel columns j-1 (0) and j (1)
    for (int q = 0; q < 2; ++q) {
      const int jj = j - 1 + q;
      double hi = 0.0, lo = 0.0;
      if (jj >= 0 && jj < n) {
        if (i < m) hi = 0.5 * (r[1][q] * nc[3 * (i * n + jj) + k]);
        if (i > 0) lo = 0.5 * (r[0][q] * nc[3 * ((i - 1) * n + jj) + k]);
      }
      s0[q] = hi + lo;
    }
    fo[3 * v + k] = 0.5 * s0[1] + 0.5 * s0[0];
  }
}

