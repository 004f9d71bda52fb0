// K4-K8: grid mapping, collocation stencils, right-hand side, wake shift and
// prescribed convection, steady / unsteady vertex forces and strip alpha.
// All are O(panels) element-wise kernels (HBM- or launch-bound); every output
// element is computed by one thread that gathers its inputs (no atomics), and
// the arithmetic follows the reference's operation order.
#include "fj_common.hpp"
#include "fj_stencil_dev.hpp"

#define FJ_EW 256

// ---- K8: zeta = R zeta0 + x, zeta_dot = W zeta0 + v (aero/uvlm.py:537-621) --
__global__ void k_grid_from_hg(const double* __restrict__ hg, const double* __restrict__ hg_dot,
                               const int32_t* __restrict__ dof, const double* __restrict__ z0,
                               int m, int n, double* __restrict__ zeta,
                               double* __restrict__ zeta_dot) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (m + 1) * (n + 1)) return;
  fj_map_vertex(hg, dof, z0, n, v, zeta + 3 * v);
  if (hg_dot && zeta_dot) fj_map_vertex(hg_dot, dof, z0, n, v, zeta_dot + 3 * v);
}

extern "C" int fj_grid_from_hg(const double* hg, const double* hg_dot, const int32_t* dof_map,
                               const double* zeta0, int m, int n, double* zeta, double* zeta_dot,
                               void* stream) {
  if (m < 0 || n < 0 || !hg || !dof_map || !zeta0 || !zeta) return FJ_ERR_ARG;
  const int nv = (m + 1) * (n + 1);
  hipLaunchKernelGGL(k_grid_from_hg, dim3(fj_cdiv(nv, FJ_EW)), dim3(FJ_EW), 0,
                     (hipStream_t)stream, hg, hg_dot, dof_map, zeta0, m, n, zeta, zeta_dot);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---- K5: collocation points, normals, collocation velocities ---------------
__global__ void k_colloc_normals(const double* __restrict__ z, const double* __restrict__ zd,
                                 int m, int n, double* __restrict__ c, double* __restrict__ nc,
                                 double* __restrict__ cd) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m * n) return;
  const int i = p / n, j = p % n;
  const int v00 = i * (n + 1) + j, v01 = v00 + 1, v10 = v00 + (n + 1), v11 = v10 + 1;
  if (c)
    for (int k = 0; k < 3; ++k)
      c[3 * p + k] = fj_quad_avg(z[3 * v00 + k], z[3 * v10 + k], z[3 * v01 + k], z[3 * v11 + k]);
  if (nc) {
    // diag1 = z[i+1,j+1] - z[i,j], diag2 = z[i+1,j] - z[i,j+1] (aero/utils.py:103-105)
    double d1[3], d2[3];
    for (int k = 0; k < 3; ++k) {
      d1[k] = z[3 * v11 + k] - z[3 * v00 + k];
      d2[k] = z[3 * v10 + k] - z[3 * v01 + k];
    }
    nc[3 * p + 0] = d1[1] * d2[2] - d1[2] * d2[1];
    nc[3 * p + 1] = d1[2] * d2[0] - d1[0] * d2[2];
    nc[3 * p + 2] = d1[0] * d2[1] - d1[1] * d2[0];
  }
  if (cd && zd)
    for (int k = 0; k < 3; ++k)
      cd[3 * p + k] =
          fj_quad_avg(zd[3 * v00 + k], zd[3 * v10 + k], zd[3 * v01 + k], zd[3 * v11 + k]);
}

extern "C" int fj_colloc_normals(const double* zeta, const double* zeta_dot, int m, int n,
                                 double* c, double* nc, double* c_dot, void* stream) {
  if (m < 0 || n < 0 || !zeta) return FJ_ERR_ARG;
  if (m * n == 0) return FJ_OK;
  hipLaunchKernelGGL(k_colloc_normals, dim3(fj_cdiv(m * n, FJ_EW)), dim3(FJ_EW), 0,
                     (hipStream_t)stream, zeta, zeta_dot, m, n, c, nc, c_dot);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---- fused step geometry ------------------------------------------------------
// One launch for everything a time step derives from the beam poses
// (aero/uvlm.py:537-621, aero/utils.py:87-105 and 162-164, uvlm.py:888-893):
// threads [0, V) map the vertices at n (and n-1, and the vertex velocities);
// threads [V, V+P) recompute their panel's four corners from the poses and
// write collocation point, normal and collocation velocity; threads
// [V+P, V+P+S) write filament midpoints.  Every value is computed with the
// same expressions as the separate kernels above, so results are identical.
__global__ void k_step_geometry(const double* __restrict__ hg, const double* __restrict__ hg_nm1,
                                const double* __restrict__ hg_dot,
                                const int32_t* __restrict__ dof, const double* __restrict__ z0,
                                int m, int n, double* __restrict__ zeta,
                                double* __restrict__ zeta_nm1, double* __restrict__ zeta_dot,
                                double* __restrict__ c, double* __restrict__ nc,
                                double* __restrict__ cd, double* __restrict__ mid) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int V = (m + 1) * (n + 1), P = m * n, S = m * (n + 1) + (m + 1) * n;
  if (e < V) {
    if (zeta) fj_map_vertex(hg, dof, z0, n, e, zeta + 3 * e);
    if (hg_nm1 && zeta_nm1) fj_map_vertex(hg_nm1, dof, z0, n, e, zeta_nm1 + 3 * e);
    if (hg_dot && zeta_dot) fj_map_vertex(hg_dot, dof, z0, n, e, zeta_dot + 3 * e);
  } else if (e < V + P) {
    const int p = e - V;
    const int i = p / n, j = p % n;
    const int v00 = i * (n + 1) + j, v01 = v00 + 1, v10 = v00 + (n + 1), v11 = v10 + 1;
    double z00[3], z01[3], z10[3], z11[3];
    fj_map_vertex(hg, dof, z0, n, v00, z00);
    fj_map_vertex(hg, dof, z0, n, v01, z01);
    fj_map_vertex(hg, dof, z0, n, v10, z10);
    fj_map_vertex(hg, dof, z0, n, v11, z11);
    if (c)
      for (int k = 0; k < 3; ++k) c[3 * p + k] = fj_quad_avg(z00[k], z10[k], z01[k], z11[k]);
    if (nc) {
      double d1[3], d2[3];
      for (int k = 0; k < 3; ++k) {
        d1[k] = z11[k] - z00[k];
        d2[k] = z10[k] - z01[k];
      }
      nc[3 * p + 0] = d1[1] * d2[2] - d1[2] * d2[1];
      nc[3 * p + 1] = d1[2] * d2[0] - d1[0] * d2[2];
      nc[3 * p + 2] = d1[0] * d2[1] - d1[1] * d2[0];
    }
    if (cd && hg_dot) {
      fj_map_vertex(hg_dot, dof, z0, n, v00, z00);
      fj_map_vertex(hg_dot, dof, z0, n, v01, z01);
      fj_map_vertex(hg_dot, dof, z0, n, v10, z10);
      fj_map_vertex(hg_dot, dof, z0, n, v11, z11);
      for (int k = 0; k < 3; ++k) cd[3 * p + k] = fj_quad_avg(z00[k], z10[k], z01[k], z11[k]);
    }
  } else if (e < V + P + S && mid) {
    const int s = e - V - P;
    int a, b;
    fj_filament_ends(m, n, s, a, b);
    double za[3], zb[3];
    fj_map_vertex(hg, dof, z0, n, a, za);
    fj_map_vertex(hg, dof, z0, n, b, zb);
    for (int k = 0; k < 3; ++k) mid[3 * s + k] = 0.5 * (za[k] + zb[k]);
  }
}

extern "C" int fj_step_geometry(const double* hg_n, const double* hg_nm1, const double* hg_dot_n,
                                const int32_t* dof_map, const double* zeta0, int m, int n,
                                double* zeta_n, double* zeta_nm1, double* zeta_dot_n, double* c,
                                double* nc, double* c_dot, double* mid, void* stream) {
  if (m < 0 || n < 0 || !hg_n || !dof_map || !zeta0) return FJ_ERR_ARG;
  const int total = (m + 1) * (n + 1) + m * n + m * (n + 1) + (m + 1) * n;
  hipLaunchKernelGGL(k_step_geometry, dim3(fj_cdiv(total, FJ_EW)), dim3(FJ_EW), 0,
                     (hipStream_t)stream, hg_n, hg_nm1, hg_dot_n, dof_map, zeta0, m, n, zeta_n,
                     zeta_nm1, zeta_dot_n, c, nc, c_dot, mid);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---- filament midpoints (aero/utils.py:162-164) ------------------------------
__global__ void k_midpoints(const double* __restrict__ z, int m, int n, double* __restrict__ mid) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m * (n + 1) + (m + 1) * n) return;
  int a, b;
  fj_filament_ends(m, n, s, a, b);
  for (int k = 0; k < 3; ++k) mid[3 * s + k] = 0.5 * (z[3 * a + k] + z[3 * b + k]);
}

extern "C" int fj_filament_midpoints(const double* zeta, int m, int n, double* mid,
                                     void* stream) {
  if (m < 0 || n < 0 || !zeta || !mid) return FJ_ERR_ARG;
  const int ns = m * (n + 1) + (m + 1) * n;
  if (ns == 0) return FJ_OK;
  hipLaunchKernelGGL(k_midpoints, dim3(fj_cdiv(ns, FJ_EW)), dim3(FJ_EW), 0, (hipStream_t)stream,
                     zeta, m, n, mid);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---- right-hand side (aero/uvlm.py:961-985) --------------------------------
__global__ void k_rhs(const double* __restrict__ c, const double* __restrict__ nc,
                      const double* __restrict__ cd, const double* __restrict__ vw, int nt,
                      FjFlow flow, double time, double* __restrict__ rhs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  double u[3];
  fj_flow_eval(flow, c[3 * t], c[3 * t + 1], c[3 * t + 2], time, u);
  if (cd) { u[0] -= cd[3 * t]; u[1] -= cd[3 * t + 1]; u[2] -= cd[3 * t + 2]; }
  if (vw) { u[0] += vw[3 * t]; u[1] += vw[3 * t + 1]; u[2] += vw[3 * t + 2]; }
  rhs[t] = -(u[0] * nc[3 * t] + u[1] * nc[3 * t + 1] + u[2] * nc[3 * t + 2]);
}

extern "C" int fj_rhs(const double* c, const double* nc, const double* c_dot, const double* v_w,
                      int n_tgt, const fj_flowfield* flow, double time, double* rhs,
                      void* stream) {
  if (n_tgt < 0 || (n_tgt > 0 && (!c || !nc || !rhs))) return FJ_ERR_ARG;
  if (n_tgt == 0) return FJ_OK;
  hipLaunchKernelGGL(k_rhs, dim3(fj_cdiv(n_tgt, FJ_EW)), dim3(FJ_EW), 0, (hipStream_t)stream, c,
                     nc, c_dot, v_w, n_tgt, fj_flow_from_desc(flow), time, rhs);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---- K4: wake shift and prescribed convection (aero/utils.py:427-450) --------
__global__ void k_wake_shift(const double* __restrict__ gb, const double* __restrict__ gw,
                             const double* __restrict__ zb, int m, int n, int m_star,
                             double* __restrict__ gw_n, double* __restrict__ zw_n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < m_star * n) {
    const int i = e / n, j = e % n;
    gw_n[e] = (i == 0) ? gb[(m - 1) * n + j] : gw[(i - 1) * n + j];
  }
  if (e < n + 1 && zw_n) {
    for (int k = 0; k < 3; ++k) zw_n[3 * e + k] = zb[3 * (m * (n + 1) + e) + k];
  }
}

extern "C" int fj_wake_shift(const double* gamma_b_nm1, const double* gamma_w_nm1,
                             const double* zeta_b_n, int m, int n, int m_star, double* gamma_w_n,
                             double* zeta_w_n, void* stream) {
  if (m < 1 || n < 0 || m_star < 0) return FJ_ERR_ARG;
  if (m_star > 0 && (!gamma_b_nm1 || !gamma_w_n || !gamma_w_nm1)) return FJ_ERR_ARG;
  if (zeta_w_n && !zeta_b_n) return FJ_ERR_ARG;
  const int ne = (m_star * n > n + 1) ? m_star * n : n + 1;
  hipLaunchKernelGGL(k_wake_shift, dim3(fj_cdiv(ne, FJ_EW)), dim3(FJ_EW), 0, (hipStream_t)stream,
                     gamma_b_nm1, gamma_w_nm1, zeta_b_n, m, n, m_star, gamma_w_n, zeta_w_n);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

__global__ void k_convect(const double* __restrict__ zw, const double* __restrict__ vel, int nv,
                          FjFlow flow, double time, double dt, double* __restrict__ out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const double x = zw[3 * v], y = zw[3 * v + 1], z = zw[3 * v + 2];
  double u[3];
  if (vel) {
    u[0] = vel[3 * v]; u[1] = vel[3 * v + 1]; u[2] = vel[3 * v + 2];
  } else {
    fj_flow_eval(flow, x, y, z, time, u);
  }
  out[3 * v] = x + dt * u[0];
  out[3 * v + 1] = y + dt * u[1];
  out[3 * v + 2] = z + dt * u[2];
}

extern "C" int fj_wake_convect(const double* zeta_w_nm1, const double* vel, int m_star, int n,
                               const fj_flowfield* flow, double time, double dt, double* zeta_w_n,
                               void* stream) {
  if (m_star < 0 || n < 0) return FJ_ERR_ARG;
  const int nv = m_star * (n + 1);
  if (nv == 0) return FJ_OK;
  if (!zeta_w_nm1 || !zeta_w_n) return FJ_ERR_ARG;
  hipLaunchKernelGGL(k_convect, dim3(fj_cdiv(nv, FJ_EW)), dim3(FJ_EW), 0, (hipStream_t)stream,
                     zeta_w_nm1, vel, nv, fj_flow_from_desc(flow), time, dt,
                     zeta_w_n + 3 * (n + 1));
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---- initial wake (aero/uvlm.py:662-689) and static wake circulation ---------
__global__ void k_wake_init(const double* __restrict__ zb, int m, int n,
                            const double* __restrict__ s, int m_star, double dx, double dy,
                            double dz, double* __restrict__ zw) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (m_star + 1) * (n + 1)) return;
  const int i = v / (n + 1), j = v % (n + 1);
  const double* te = zb + 3 * (m * (n + 1) + j);
  zw[3 * v] = te[0] + s[i] * dx;
  zw[3 * v + 1] = te[1] + s[i] * dy;
  zw[3 * v + 2] = te[2] + s[i] * dz;
}

extern "C" int fj_wake_initialise(const double* zeta_b, int m, int n, const double* s, int m_star,
                                  const double* dir, double* zeta_w, void* stream) {
  if (m < 0 || n < 0 || m_star < 0 || !zeta_b || !s || !dir || !zeta_w) return FJ_ERR_ARG;
  const int nv = (m_star + 1) * (n + 1);
  hipLaunchKernelGGL(k_wake_init, dim3(fj_cdiv(nv, FJ_EW)), dim3(FJ_EW), 0, (hipStream_t)stream,
                     zeta_b, m, n, s, m_star, dir[0], dir[1], dir[2], zeta_w);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

__global__ void k_te_broadcast(const double* __restrict__ gb, int m, int n, int rows,
                               double* __restrict__ gw) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * n) return;
  gw[e] = gb[(m - 1) * n + e % n];
}

extern "C" int fj_static_wake_gamma(const double* gamma_b, int m, int n, int rows, double* gamma_w,
                                    void* stream) {
  if (m < 1 || n < 0 || rows < 0 || !gamma_b || (rows * n > 0 && !gamma_w)) return FJ_ERR_ARG;
  if (rows * n == 0) return FJ_OK;
  hipLaunchKernelGGL(k_te_broadcast, dim3(fj_cdiv(rows * n, FJ_EW)), dim3(FJ_EW), 0,
                     (hipStream_t)stream, gamma_b, m, n, rows, gamma_w);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---- background flow at points (FlowField.vmap_call, aero/flowfields.py:49-60)
__global__ void k_flow_eval(FjFlow flow, const double* __restrict__ x, int nt, double time,
                            double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  double u[3];
  fj_flow_eval(flow, x[3 * t], x[3 * t + 1], x[3 * t + 2], time, u);
  out[3 * t] = u[0]; out[3 * t + 1] = u[1]; out[3 * t + 2] = u[2];
}

extern "C" int fj_flow_eval(const fj_flowfield* flow, const double* x, int n_pts, double time,
                            double* out, void* stream) {
  if (!flow || n_pts < 0 || (n_pts > 0 && (!x || !out))) return FJ_ERR_ARG;
  if (n_pts == 0) return FJ_OK;
  hipLaunchKernelGGL(k_flow_eval, dim3(fj_cdiv(n_pts, FJ_EW)), dim3(FJ_EW), 0,
                     (hipStream_t)stream, fj_flow_from_desc(flow), x, n_pts, time, out);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---- K6: steady vertex forces (aero/utils.py:126-223) ------------------------
__global__ void k_steady_forces(const double* __restrict__ z, const double* __restrict__ zd,
                                const double* __restrict__ G, const double* __restrict__ gw0,
                                const double* __restrict__ vmid, int m, int n, FjFlow flow,
                                double time, double rho, double* __restrict__ fo) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (m + 1) * (n + 1)) return;
  fj_vertex_steady_force(z, zd, G, gw0, vmid, m, n, flow, time, rho, v, fo);
}

extern "C" int fj_steady_forces(const double* zeta, const double* zeta_dot, const double* gamma_b,
                                const double* gamma_w0, const double* v_mid, int m, int n,
                                const fj_flowfield* flow, double time, double rho, double* f_out,
                                void* stream) {
  if (m < 0 || n < 0 || !zeta || !gamma_b || !v_mid || !f_out) return FJ_ERR_ARG;
  const int nv = (m + 1) * (n + 1);
  FjFlow fl = fj_flow_from_desc(flow);
  if (!flow) fl.kind = -1;  // v_mid already holds the total velocity
  hipLaunchKernelGGL(k_steady_forces, dim3(fj_cdiv(nv, FJ_EW)), dim3(FJ_EW), 0,
                     (hipStream_t)stream, zeta, zeta_dot, gamma_b, gamma_w0, v_mid, m, n, fl,
                     time, rho, f_out);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---- K7: circulation rate and unsteady force (aero/uvlm.py:691-715, 1086-1103)
__global__ void k_gamma_dot(const double* __restrict__ gn, const double* __restrict__ gp,
                            const double* __restrict__ gdp, int np, double dt, double relax,
                            double* __restrict__ gd) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  gd[p] = fj_gamma_rate(gn, gp, gdp, p, dt, relax);
}

__global__ void k_unsteady_force(const double* __restrict__ gd, const double* __restrict__ nc,
                                 int m, int n, double rho, double* __restrict__ fo) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (m + 1) * (n + 1)) return;
  fj_vertex_unsteady_force([=](int p) { return gd[p]; }, nc, m, n, rho, v, fo);
}

extern "C" int fj_gamma_dot_unsteady(const double* gamma_n, const double* gamma_nm1,
                                     const double* gamma_dot_nm1, const double* nc, int m, int n,
                                     double dt, double relax, double rho, double* gamma_dot_n,
                                     double* f_unsteady, void* stream) {
  if (m < 0 || n < 0 || !gamma_n || !gamma_nm1 || !gamma_dot_nm1 || !gamma_dot_n) return FJ_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (m * n > 0)
    hipLaunchKernelGGL(k_gamma_dot, dim3(fj_cdiv(m * n, FJ_EW)), dim3(FJ_EW), 0, st, gamma_n,
                       gamma_nm1, gamma_dot_nm1, m * n, dt, relax, gamma_dot_n);
  FJ_CHECK_LAUNCH();
  if (f_unsteady) {
    if (!nc) return FJ_ERR_ARG;
    hipLaunchKernelGGL(k_unsteady_force, dim3(fj_cdiv((m + 1) * (n + 1), FJ_EW)), dim3(FJ_EW), 0,
                       st, gamma_dot_n, nc, m, n, rho, f_unsteady);
    FJ_CHECK_LAUNCH();
  }
  return FJ_OK;
}

// ---- fused step forces -------------------------------------------------------
// Steady vertex forces, circulation rate and unsteady vertex forces of a
// dynamic step in one launch: threads [0, V) write both forces of vertex v
// (the rates of the up to four adjacent panels are recomputed with the rate
// expression), threads [V, V+P) write the panel rates.
__global__ void k_step_forces(const double* __restrict__ z, const double* __restrict__ zd,
                              const double* __restrict__ G, const double* __restrict__ Gp,
                              const double* __restrict__ Gdp, const double* __restrict__ gw0,
                              const double* __restrict__ vmid, const double* __restrict__ nc,
                              int m, int n, FjFlow flow, double time, double rho, double dt,
                              double relax, double* __restrict__ f_steady,
                              double* __restrict__ gdot, double* __restrict__ f_uns) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int V = (m + 1) * (n + 1), P = m * n;
  if (e < V) {
    fj_vertex_steady_force(z, zd, G, gw0, vmid, m, n, flow, time, rho, e, f_steady);
    if (f_uns && Gp)
      fj_vertex_unsteady_force([=](int p) { return fj_gamma_rate(G, Gp, Gdp, p, dt, relax); }, nc,
                               m, n, rho, e, f_uns);
  } else if (e < V + P && Gp && gdot) {
    gdot[e - V] = fj_gamma_rate(G, Gp, Gdp, e - V, dt, relax);
  }
}

extern "C" int fj_step_forces(const double* zeta, const double* zeta_dot, const double* gamma_b,
                              const double* gamma_b_nm1, const double* gamma_dot_nm1,
                              const double* gamma_w0, const double* v_mid, const double* nc,
                              int m, int n, const fj_flowfield* flow, double time, double rho,
                              double dt, double relax, double* f_steady, double* gamma_dot_n,
                              double* f_unsteady, void* stream) {
  if (m < 0 || n < 0 || !zeta || !gamma_b || !v_mid || !f_steady) return FJ_ERR_ARG;
  if (gamma_b_nm1 && (!gamma_dot_nm1 || !gamma_dot_n || !(dt > 0.0))) return FJ_ERR_ARG;
  if (f_unsteady && (!gamma_b_nm1 || !nc)) return FJ_ERR_ARG;
  const int total = (m + 1) * (n + 1) + m * n;
  FjFlow fl = fj_flow_from_desc(flow);
  if (!flow) fl.kind = -1;
  hipLaunchKernelGGL(k_step_forces, dim3(fj_cdiv(total, FJ_EW)), dim3(FJ_EW), 0,
                     (hipStream_t)stream, zeta, zeta_dot, gamma_b, gamma_b_nm1, gamma_dot_nm1,
                     gamma_w0, v_mid, nc, m, n, fl, time, rho, dt, relax, f_steady, gamma_dot_n,
                     f_unsteady);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---- strip alpha (aero/utils.py:226-273) -------------------------------------
__global__ void k_strip_alpha(const double* __restrict__ z, const double* __restrict__ f, int m,
                              int n, FjFlow flow, double time, double rho,
                              double* __restrict__ alpha) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double* le0 = z + 3 * j;
  const double* le1 = z + 3 * (j + 1);
  const double* te0 = z + 3 * (m * (n + 1) + j);
  const double* te1 = z + 3 * (m * (n + 1) + j + 1);
  double lej[3], tej[3], ch[3], sp[3], mid[3];
  for (int k = 0; k < 3; ++k) {
    lej[k] = 0.5 * (le0[k] + le1[k]);
    tej[k] = 0.5 * (te0[k] + te1[k]);
    ch[k] = tej[k] - lej[k];
    sp[k] = 0.5 * ((le1[k] + te1[k]) - (le0[k] + te0[k]));
    mid[k] = 0.5 * (lej[k] + tej[k]);
  }
  const double c_len = sqrt(ch[0] * ch[0] + ch[1] * ch[1] + ch[2] * ch[2]);
  const double b_len = sqrt(sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2]);
  const double es[3] = {sp[0] / b_len, sp[1] / b_len, sp[2] / b_len};
  double u[3];
  fj_flow_eval(flow, mid[0], mid[1], mid[2], time, u);
  const double v2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  double el[3] = {u[1] * es[2] - u[2] * es[1], u[2] * es[0] - u[0] * es[2],
                  u[0] * es[1] - u[1] * es[0]};
  const double eln = sqrt(el[0] * el[0] + el[1] * el[1] + el[2] * el[2]);
  el[0] /= eln; el[1] /= eln; el[2] /= eln;
  double fs[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i <= m; ++i)
    for (int k = 0; k < 3; ++k)
      fs[k] += 0.5 * (f[3 * (i * (n + 1) + j) + k] + f[3 * (i * (n + 1) + j + 1) + k]);
  const double q = 0.5 * rho * v2;
  alpha[j] = (fs[0] * el[0] + fs[1] * el[1] + fs[2] * el[2]) / (q * c_len * b_len) / (2.0 * M_PI);
}

extern "C" int fj_strip_alpha(const double* zeta, const double* f_steady, int m, int n,
                              const fj_flowfield* flow, double time, double rho, double* alpha,
                              void* stream) {
  if (m < 0 || n < 0 || !zeta || !f_steady || !alpha) return FJ_ERR_ARG;
  if (n == 0) return FJ_OK;
  hipLaunchKernelGGL(k_strip_alpha, dim3(fj_cdiv(n, 64)), dim3(64), 0, (hipStream_t)stream, zeta,
                     f_steady, m, n, fj_flow_from_desc(flow), time, rho, alpha);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

extern "C" const char* fj_version(void) { return "flapjax_amd 0.1.0 (gfx950)"; }

// ---- variable wake discretisation (aero/utils.py:452-491) --------------------
// Per spanwise column: arc length along the m_pre = m_star+2 convected points,
// linear interpolation (clamped at the ends, as jnp.interp) onto the target
// stations s_disc = [0, cumsum(delta_w)] for the vertices and onto their
// midpoints for the circulations; circulation stations are the arc length
// averaged over the two neighbouring columns and rows
// (neighbour_average(s, (0, 1)), algebra/array_utils.py:106-135).
__device__ __forceinline__ double fj_interp(double x, const double* xp, const double* fp, int np,
                                            int fstride) {
  if (x <= xp[0]) return fp[0];
  if (x >= xp[np - 1]) return fp[(size_t)(np - 1) * fstride];
  int lo = 0, hi = np - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] <= x) lo = mid; else hi = mid;
  }
  const double t = (x - xp[lo]) / (xp[hi] - xp[lo]);
  const double f0 = fp[(size_t)lo * fstride], f1 = fp[(size_t)hi * fstride];
  return f0 + t * (f1 - f0);
}

// arc length s (m_pre, n+1) of the pre-discretisation wake zeta_pre (m_pre, n+1, 3)
__global__ void k_wake_arclength(const double* __restrict__ zp, int m_pre, int n,
                                 double* __restrict__ s) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > n) return;
  double acc = 0.0;
  s[j] = 0.0;
  for (int i = 1; i < m_pre; ++i) {
    const double* a = zp + 3 * ((size_t)(i - 1) * (n + 1) + j);
    const double* b = zp + 3 * ((size_t)i * (n + 1) + j);
    const double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    acc += sqrt(dx * dx + dy * dy + dz * dz);
    s[(size_t)i * (n + 1) + j] = acc;
  }
}

__global__ void k_wake_rediscretise(const double* __restrict__ zp, const double* __restrict__ gp,
                                    const double* __restrict__ s, int m_pre, int n,
                                    const double* __restrict__ s_disc, int m_star,
                                    double* __restrict__ scratch, double* __restrict__ zw,
                                    double* __restrict__ gw) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > n) return;
  // column j of the arc length, contiguous copy for the search
  double* xs = scratch + (size_t)j * 2 * m_pre;
  for (int i = 0; i < m_pre; ++i) xs[i] = s[(size_t)i * (n + 1) + j];
  for (int i = 0; i <= m_star; ++i)
    for (int k = 0; k < 3; ++k)
      zw[3 * ((size_t)i * (n + 1) + j) + k] =
          fj_interp(s_disc[i], xs, zp + 3 * (size_t)j + k, m_pre, 3 * (n + 1));
  if (j < n && m_star > 0) {
    // circulation stations: s averaged along the rows, then the two columns
    double* xg = xs + m_pre;
    for (int i = 0; i < m_pre - 1; ++i) {
      const double a0 = 0.5 * (s[(size_t)i * (n + 1) + j] + s[(size_t)(i + 1) * (n + 1) + j]);
      const double a1 = 0.5 * (s[(size_t)i * (n + 1) + j + 1] + s[(size_t)(i + 1) * (n + 1) + j + 1]);
      xg[i] = 0.5 * (a0 + a1);
    }
    for (int i = 0; i < m_star; ++i) {
      const double sd = 0.5 * (s_disc[i] + s_disc[i + 1]);
      gw[(size_t)i * n + j] = fj_interp(sd, xg, gp + j, m_pre - 1, n);
    }
  }
}

extern "C" size_t fj_wake_rediscretise_workspace_bytes(int m_star, int n) {
  const size_t m_pre = (size_t)m_star + 2;
  return (m_pre * (n + 1) + (size_t)(n + 1) * 2 * m_pre) * sizeof(double) + 256;
}

extern "C" int fj_wake_rediscretise(const double* zeta_pre, const double* gamma_pre, int m_star,
                                    int n, const double* s_disc, double* zeta_w, double* gamma_w,
                                    void* workspace, void* stream) {
  if (m_star < 0 || n < 0 || !zeta_pre || !s_disc || !zeta_w || !workspace) return FJ_ERR_ARG;
  if (m_star > 0 && (!gamma_pre || !gamma_w)) return FJ_ERR_ARG;
  const int m_pre = m_star + 2;
  double* s = (double*)workspace;
  double* scratch = s + (size_t)m_pre * (n + 1);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_wake_arclength, dim3(fj_cdiv(n + 1, 64)), dim3(64), 0, st, zeta_pre, m_pre, n,
                     s);
  hipLaunchKernelGGL(k_wake_rediscretise, dim3(fj_cdiv(n + 1, 64)), dim3(64), 0, st, zeta_pre,
                     gamma_pre, s, m_pre, n, s_disc, m_star, scratch, zeta_w, gamma_w);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}
