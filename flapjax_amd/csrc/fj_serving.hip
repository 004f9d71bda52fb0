// Serving: one UVLM time step of many independent cases in one set of
// launches (fj_step_batched).
//
// The per-step work of base_solve_from_grid (aero/uvlm.py:824-1117) for a
// uniform wake and no mirror, for every active case at once: the host builds
// a table of case records (sizes, the case's buffers, and each case's first
// work item in every batched kernel), copies it to the device once, and
// every stage is ONE launch over all cases -- a work item finds its case by
// a binary search over the records' prefix offsets.  Stages:
//   geometry (grids at n, grid velocities, collocation points, normals,
//     filament midpoints, wake circulation shift and TE row, or the
//     prescribed convection), the rollup filament tables, the fused rollup
//     pair sum with the convection epilogue, the wake tables, the fused
//     upwash pair sum, the AIC of every case, the right-hand sides, the
//     kept-inverse refinement of every case (10 launches for all), the bound
//     tables, the fused force pair sum, steady forces with circulation rate,
//     and the lift of every case.
// Every quantity is computed with the device functions of the per-case entry
// points, so a served case equals the same case stepped alone.
#include <cstring>
#include <vector>

#include "fj_common.hpp"
#include "fj_lattice_dev.hpp"
#include "fj_refine_dev.hpp"
#include "fj_segments_dev.hpp"
#include "fj_stencil_dev.hpp"

struct FjCaseDev {
  fj_step_case c;
  int nb, s_b, s_w;     // panels, bound filaments, wake filaments (0 without wake rows)
  int p_off, mid_off;   // first panel / filament midpoint in the packed arrays
  int seg_roll, seg_frc;  // first row of the case's rollup / force filament tables
  // first work item (or workgroup) of the case in each batched kernel
  int geo_off, roll_off, wk_off, bnd_off, aic_off, rhs_off, rblk_off, frc_off;
};

template <int FjCaseDev::*F>
__device__ __forceinline__ int fj_case_of(const FjCaseDev* __restrict__ cs, int n, int item) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cs[mid].*F <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

#define FJ_SV_EW 256

// ---- geometry, wake shift and TE row (or prescribed convection) ------------
__global__ void k_bt_geometry(const FjCaseDev* __restrict__ cs, int ncase, int total, FjFlow flow,
                              int free_wake, double* __restrict__ c, double* __restrict__ nc,
                              double* __restrict__ cd, double* __restrict__ mids) {
  for (int item = blockIdx.x * blockDim.x + threadIdx.x; item < total;
       item += gridDim.x * blockDim.x) {
    const FjCaseDev& C = cs[fj_case_of<&FjCaseDev::geo_off>(cs, ncase, item)];
    const fj_step_case& K = C.c;
    const int m = K.m, n = K.n, ms = K.m_star;
    const int V = (m + 1) * (n + 1), P = m * n, S = C.s_b;
    int e = item - C.geo_off;
    if (e < V) {
      fj_map_vertex(K.hg, K.dof, K.zeta0, n, e, K.zeta_b + 3 * e);
      fj_map_vertex(K.hg_dot, K.dof, K.zeta0, n, e, K.zeta_b_dot + 3 * e);
      continue;
    }
    e -= V;
    if (e < P) {
      const int i = e / n, j = e % n;
      const int v00 = i * (n + 1) + j, v01 = v00 + 1, v10 = v00 + (n + 1), v11 = v10 + 1;
      double z00[3], z01[3], z10[3], z11[3];
      fj_map_vertex(K.hg, K.dof, K.zeta0, n, v00, z00);
      fj_map_vertex(K.hg, K.dof, K.zeta0, n, v01, z01);
      fj_map_vertex(K.hg, K.dof, K.zeta0, n, v10, z10);
      fj_map_vertex(K.hg, K.dof, K.zeta0, n, v11, z11);
      const size_t o = 3 * (size_t)(C.p_off + e);
      for (int k = 0; k < 3; ++k) c[o + k] = fj_quad_avg(z00[k], z10[k], z01[k], z11[k]);
      double d1[3], d2[3];
      for (int k = 0; k < 3; ++k) {
        d1[k] = z11[k] - z00[k];
        d2[k] = z10[k] - z01[k];
      }
      nc[o + 0] = d1[1] * d2[2] - d1[2] * d2[1];
      nc[o + 1] = d1[2] * d2[0] - d1[0] * d2[2];
      nc[o + 2] = d1[0] * d2[1] - d1[1] * d2[0];
      fj_map_vertex(K.hg_dot, K.dof, K.zeta0, n, v00, z00);
      fj_map_vertex(K.hg_dot, K.dof, K.zeta0, n, v01, z01);
      fj_map_vertex(K.hg_dot, K.dof, K.zeta0, n, v10, z10);
      fj_map_vertex(K.hg_dot, K.dof, K.zeta0, n, v11, z11);
      for (int k = 0; k < 3; ++k) cd[o + k] = fj_quad_avg(z00[k], z10[k], z01[k], z11[k]);
      continue;
    }
    e -= P;
    if (e < S) {
      int a, b;
      fj_filament_ends(m, n, e, a, b);
      double za[3], zb[3];
      fj_map_vertex(K.hg, K.dof, K.zeta0, n, a, za);
      fj_map_vertex(K.hg, K.dof, K.zeta0, n, b, zb);
      const size_t o = 3 * (size_t)(C.mid_off + e);
      for (int k = 0; k < 3; ++k) mids[o + k] = 0.5 * (za[k] + zb[k]);
      continue;
    }
    e -= S;
    if (e < ms * n) {
      // circulation shift (aero/utils.py:440-442)
      const int i = e / n, j = e % n;
      K.gamma_w[e] = (i == 0) ? K.gamma_b_nm1[(m - 1) * n + j] : K.gamma_w_nm1[(i - 1) * n + j];
      continue;
    }
    e -= ms * n;
    if (e < n + 1) {
      // first wake row = trailing edge at n (aero/utils.py:436-439)
      fj_map_vertex(K.hg, K.dof, K.zeta0, n, m * (n + 1) + e, K.zeta_w + 3 * e);
      continue;
    }
    e -= n + 1;
    if (!free_wake && e < ms * (n + 1)) {
      // prescribed convection x + dt U(x, t) (aero/utils.py:444-450)
      const double* xp = K.zeta_w_nm1 + 3 * e;
      double u[3];
      fj_flow_eval(flow, xp[0], xp[1], xp[2], K.time, u);
      double* o = K.zeta_w + 3 * (n + 1) + 3 * e;
      o[0] = xp[0] + K.dt * u[0];
      o[1] = xp[1] + K.dt * u[1];
      o[2] = xp[2] + K.dt * u[2];
    }
  }
}

// ---- filament tables ------------------------------------------------------------
// which 0: rollup table [bound n-1 | wake n-1]; 1: wake n into the force
// table after the bound part; 2: bound n at the head of the force table
__global__ void k_bt_segments(const FjCaseDev* __restrict__ cs, int ncase, int total, int which,
                              int kernel, double eps, double* __restrict__ table) {
  for (int item = blockIdx.x * blockDim.x + threadIdx.x; item < total;
       item += gridDim.x * blockDim.x) {
    int ci, local;
    if (which == 0) {
      ci = fj_case_of<&FjCaseDev::roll_off>(cs, ncase, item);
      local = item - cs[ci].roll_off;
    } else if (which == 1) {
      ci = fj_case_of<&FjCaseDev::wk_off>(cs, ncase, item);
      local = item - cs[ci].wk_off;
    } else {
      ci = fj_case_of<&FjCaseDev::bnd_off>(cs, ncase, item);
      local = item - cs[ci].bnd_off;
    }
    const FjCaseDev& C = cs[ci];
    const fj_step_case& K = C.c;
    const double* zeta;
    const double* gamma;
    int m, s, row;
    if (which == 0) {
      if (local < C.s_b) {
        zeta = K.zeta_b_nm1; gamma = K.gamma_b_nm1; m = K.m; s = local;
      } else {
        zeta = K.zeta_w_nm1; gamma = K.gamma_w_nm1; m = K.m_star; s = local - C.s_b;
      }
      row = C.seg_roll + local;
    } else if (which == 1) {
      zeta = K.zeta_w; gamma = K.gamma_w; m = K.m_star; s = local;
      row = C.seg_frc + C.s_b + local;
    } else {
      zeta = K.zeta_b; gamma = K.gamma_b; m = K.m; s = local;
      row = C.seg_frc + local;
    }
    fj_segment_record(zeta, gamma, m, K.n, s, 0, nullptr, nullptr, kernel, eps,
                      table + (size_t)row * FJ_SEG_STRIDE);
  }
}

// ---- AIC of every case (bound rings, no mirror, no wake lumping), by column
// strips: one thread per (case, collocation point, block of FJ_AIC_W ring
// columns) walks the bound lattice row by row, keeping the per-vertex terms
// (r, 1/|r|) of its FJ_AIC_W + 1 vertices of the current row and the
// spanwise filaments between them in registers; per row it evaluates the
// vertices below, the chordwise filaments between the rows and the spanwise
// filaments of the new row, each once, and combines them into the ring
// columns as the reference does,
// A[t, (i, j)] = -(M[i, j+1] - M[i, j]) + (N[i+1, j] - N[i, j])
// (aero/aic.py:55-57): (W+1) vertex and 2W+1 filament evaluations per W
// rings and row instead of 4W filament evaluations; no barriers, no shuffles.
#define FJ_AIC_W 2
#define FJ_AIC_THREADS 256

template <int KIND>
__device__ __forceinline__ double fj_aic_fil(const FjVtx& A, const FjVtx& B,
                                             const double* __restrict__ ya,
                                             const double* __restrict__ yb, double nx, double ny,
                                             double nz, double eps, double e2, double rcut2) {
  const double r0x = yb[0] - ya[0], r0y = yb[1] - ya[1], r0z = yb[2] - ya[2];
  const double cx = A.ry * B.rz - A.rz * B.ry;
  const double cy = A.rz * B.rx - A.rx * B.rz;
  const double cz = A.rx * B.ry - A.ry * B.rx;
  const double dx = fma(A.rx, A.inv, -(B.rx * B.inv));
  const double dy = fma(A.ry, A.inv, -(B.ry * B.inv));
  const double dz = fma(A.rz, A.inv, -(B.rz * B.inv));
  const double a = r0x * dx + r0y * dy + r0z * dz;
  const double r0sq = r0x * r0x + r0y * r0y + r0z * r0z;
  const double D = fma(cz, cz, fma(cy, cy, fma(cx, cx, KIND == FJ_KERNEL_EPSILON ? eps * r0sq * r0sq : 0.0)));
  double w = __builtin_amdgcn_rcp(D);
  const double e = fma(-D, w, 1.0);
  w = fma(w, fma(e, e, e), w);
  double q = a * w * FJ_INV_4PI;
  if (KIND == FJ_KERNEL_CUTOFF) {
    const double a1 = r0x * A.rx + r0y * A.ry + r0z * A.rz;
    const double rad2 = (A.s - e2) - a1 * a1 / r0sq;
    q = (rad2 > rcut2) ? q : 0.0;
  }
  return (cx * nx + cy * ny + cz * nz) * q;
}

template <int KIND>
__global__ __launch_bounds__(FJ_AIC_THREADS) void k_bt_aic_strips(
    const FjCaseDev* __restrict__ cs, int ncase, int total, const double* __restrict__ c,
    const double* __restrict__ nc, FjKernelParams kp) {
  const int item = blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= total) return;
  const FjCaseDev& C = cs[fj_case_of<&FjCaseDev::aic_off>(cs, ncase, item)];
  const fj_step_case& K = C.c;
  const int m = K.m, n = K.n;
  const int blocks = fj_cdiv(n, FJ_AIC_W);
  const int local = item - C.aic_off;
  const int t = local / blocks, j0 = (local - t * blocks) * FJ_AIC_W;
  const size_t o = 3 * (size_t)(C.p_off + t);
  const double x = c[o], y = c[o + 1], zt = c[o + 2];
  const double nx = nc[o], ny = nc[o + 1], nz = nc[o + 2];
  const double* zb = K.zeta_b;
  const double e2 = (KIND == FJ_KERNEL_PLAIN) ? 0.0 : kp.eps2;
  const double rcut2 = kp.rcut * kp.rcut;
  // vertex columns j0 .. j0 + W (clamped to n: columns past the lattice
  // repeat the last one and their rings are not written)
  int col[FJ_AIC_W + 1];
#pragma unroll
  for (int k = 0; k <= FJ_AIC_W; ++k) col[k] = min(j0 + k, n);
  FjVtx T[FJ_AIC_W + 1];
  double ntop[FJ_AIC_W];
  const double* ya = zb;  // vertex row 0
#pragma unroll
  for (int k = 0; k <= FJ_AIC_W; ++k) T[k] = fj_lat_vertex(ya + 3 * col[k], x, y, zt, e2);
#pragma unroll
  for (int k = 0; k < FJ_AIC_W; ++k)
    ntop[k] = fj_aic_fil<KIND>(T[k], T[k + 1], ya + 3 * col[k], ya + 3 * col[k + 1], nx, ny, nz,
                               kp.eps, e2, rcut2);
  double* arow = K.aic + (size_t)t * C.nb;
  for (int i = 0; i < m; ++i) {
    const double* yb = ya + 3 * (size_t)(n + 1);  // vertex row i + 1
    FjVtx B[FJ_AIC_W + 1];
    double mc[FJ_AIC_W + 1];
#pragma unroll
    for (int k = 0; k <= FJ_AIC_W; ++k) {
      B[k] = fj_lat_vertex(yb + 3 * col[k], x, y, zt, e2);
      mc[k] = fj_aic_fil<KIND>(T[k], B[k], ya + 3 * col[k], yb + 3 * col[k], nx, ny, nz, kp.eps, e2,
                               rcut2);
    }
#pragma unroll
    for (int k = 0; k < FJ_AIC_W; ++k) {
      const double nbot = fj_aic_fil<KIND>(B[k], B[k + 1], yb + 3 * col[k], yb + 3 * col[k + 1], nx,
                                           ny, nz, kp.eps, e2, rcut2);
      if (j0 + k < n) arow[i * n + j0 + k] = -(mc[k + 1] - mc[k]) + (nbot - ntop[k]);
      ntop[k] = nbot;
    }
#pragma unroll
    for (int k = 0; k <= FJ_AIC_W; ++k) T[k] = B[k];
    ya = yb;
  }
}

// ---- right-hand sides (aero/uvlm.py:961-985) -------------------------------------
__global__ void k_bt_rhs(const FjCaseDev* __restrict__ cs, int ncase, int total, FjFlow flow,
                         const double* __restrict__ c, const double* __restrict__ nc,
                         const double* __restrict__ cd, const double* __restrict__ vw,
                         double* __restrict__ rhs) {
  for (int item = blockIdx.x * blockDim.x + threadIdx.x; item < total;
       item += gridDim.x * blockDim.x) {
    const FjCaseDev& C = cs[fj_case_of<&FjCaseDev::rhs_off>(cs, ncase, item)];
    const size_t t = (size_t)C.p_off + (item - C.rhs_off);
    double u[3];
    fj_flow_eval(flow, c[3 * t], c[3 * t + 1], c[3 * t + 2], C.c.time, u);
    u[0] -= cd[3 * t]; u[1] -= cd[3 * t + 1]; u[2] -= cd[3 * t + 2];
    u[0] += vw[3 * t]; u[1] += vw[3 * t + 1]; u[2] += vw[3 * t + 2];
    rhs[t] = -(u[0] * nc[3 * t] + u[1] * nc[3 * t + 1] + u[2] * nc[3 * t + 2]);
  }
}

// ---- kept-inverse refinement of every case (the scheme of fj_lu.hip) -----------
__global__ void k_bt_ref_init(const FjCaseDev* __restrict__ cs, int max_iter) {
  const FjCaseDev& C = cs[blockIdx.x];
  const FjRefineWs w = fj_refine_ws(C.c.refine_ws, C.nb);
  for (int i = threadIdx.x; i < 3 * (max_iter + 1); i += blockDim.x) w.slots[i] = 0ull;
  if (threadIdx.x == 0) { w.flags[0] = 0; w.flags[1] = 0; }
}

__global__ __launch_bounds__(64 * FJ_RV_WAVES) void k_bt_ref_residual(
    const FjCaseDev* __restrict__ cs, int ncase, const double* __restrict__ rhs, int k) {
  __shared__ double sred[2][FJ_RV_WAVES];
  const FjCaseDev& C = cs[fj_case_of<&FjCaseDev::rblk_off>(cs, ncase, blockIdx.x)];
  const FjRefineWs w = fj_refine_ws(C.c.refine_ws, C.nb);
  if (w.flags[0]) return;
  const int n = C.nb;
  const int lane = threadIdx.x & 63;
  const int row = (blockIdx.x - C.rblk_off) * FJ_RV_WAVES + (threadIdx.x >> 6);
  const double* x = C.c.gamma_b;
  const double* b = rhs + C.p_off;
  double s = 0.0, a = 0.0;
  if (row < n) {
    const double* ar = C.c.aic + (size_t)row * n;
#pragma unroll 4
    for (int q = lane; q < n; q += 64) {
      const double v = ar[q];
      s += v * x[q];
      a += fabs(v);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off);
    a += __shfl_xor(a, off);
  }
  double rr = 0.0;
  if (row < n) {
    rr = b[row] - s;
    if (lane == 0) w.r[row] = rr;
  }
  fj_group_max(a, &w.slots[3 * k + 0], sred[0]);
  fj_group_max(fabs(rr), &w.slots[3 * k + 1], sred[1]);
}

template <bool CORRECT>
__global__ __launch_bounds__(64 * FJ_RV_WAVES) void k_bt_ref_apply(
    const FjCaseDev* __restrict__ cs, int ncase, const double* __restrict__ rhs, int k,
    double tol_arg) {
  __shared__ double sred[FJ_RV_WAVES];
  const FjCaseDev& C = cs[fj_case_of<&FjCaseDev::rblk_off>(cs, ncase, blockIdx.x)];
  const int n = C.nb;
  const FjRefineWs w = fj_refine_ws(C.c.refine_ws, n);
  if (w.flags[0]) return;
  const double tol = fj_refine_tol(tol_arg, n);
  const bool lead = (blockIdx.x == C.rblk_off) && threadIdx.x == 0;
  if (CORRECT && fj_refine_converged(w.slots + 3 * k, tol)) {
    if (lead) { w.flags[0] = 1; w.flags[1] = k; }
    return;
  }
  const int lane = threadIdx.x & 63;
  const int row = (blockIdx.x - C.rblk_off) * FJ_RV_WAVES + (threadIdx.x >> 6);
  const double* v = CORRECT ? w.r : rhs + C.p_off;
  double* x = C.c.gamma_b;
  double s = 0.0;
  if (row < n) {
    const double* mr = C.c.aic_inv + (size_t)row * n;
#pragma unroll 4
    for (int q = lane; q < n; q += 64) s += mr[q] * v[q];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  double xn = 0.0;
  if (row < n) {
    xn = CORRECT ? x[row] + s : s;
    if (lane == 0) x[row] = xn;
  }
  fj_group_max(fabs(xn), &w.slots[3 * (CORRECT ? k + 1 : 0) + 2], sred);
  if (CORRECT && lead) w.flags[1] = k + 1;
}

__global__ void k_bt_ref_finish(const FjCaseDev* __restrict__ cs, int ncase, int max_iter,
                                double tol_arg) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncase) return;
  const FjCaseDev& C = cs[ci];
  const FjRefineWs w = fj_refine_ws(C.c.refine_ws, C.nb);
  const double tol = fj_refine_tol(tol_arg, C.nb);
  C.c.iters[0] = (w.flags[0] == 1) ? w.flags[1]
                 : (fj_refine_converged(w.slots + 3 * max_iter, tol) ? max_iter : -max_iter);
}

// ---- steady forces, circulation rate, lift --------------------------------------
__global__ void k_bt_forces(const FjCaseDev* __restrict__ cs, int ncase, int total, FjFlow flow,
                            double rho, double relax, const double* __restrict__ vmid) {
  for (int item = blockIdx.x * blockDim.x + threadIdx.x; item < total;
       item += gridDim.x * blockDim.x) {
    const FjCaseDev& C = cs[fj_case_of<&FjCaseDev::frc_off>(cs, ncase, item)];
    const fj_step_case& K = C.c;
    const int V = (K.m + 1) * (K.n + 1);
    const int e = item - C.frc_off;
    if (e < V) {
      fj_vertex_steady_force(K.zeta_b, K.zeta_b_dot, K.gamma_b,
                             K.m_star > 0 ? K.gamma_w : nullptr, vmid + 3 * (size_t)C.mid_off,
                             K.m, K.n, flow, K.time, rho, e, K.f_steady);
    } else {
      const int p = e - V;
      K.gamma_b_dot[p] = fj_gamma_rate(K.gamma_b, K.gamma_b_nm1, K.gamma_b_dot, p, K.dt, relax);
    }
  }
}

// lift = sum of the vertical vertex forces, one workgroup per case, fixed order
__global__ __launch_bounds__(256) void k_bt_lift(const FjCaseDev* __restrict__ cs) {
  __shared__ double red[256];
  const fj_step_case& K = cs[blockIdx.x].c;
  const int V = (K.m + 1) * (K.n + 1);
  double s = 0.0;
  for (int v = threadIdx.x; v < V; v += 256) s += K.f_steady[3 * v + 2];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) K.lift[0] = red[0];
}

// ---- host side ---------------------------------------------------------------------
static size_t fj_align256(size_t b) { return (b + 255) / 256 * 256; }

struct FjTickLayout {
  int nb_tot = 0, mid_tot = 0, roll_rows = 0, frc_rows = 0;
  int geo_tot = 0, roll_tot = 0, wk_tot = 0, bnd_tot = 0, aic_tot = 0, rblk_tot = 0, frc_tot = 0;
  size_t ws_roll = 0, ws_up = 0, ws_frc = 0;
  size_t off_c, off_nc, off_cd, off_rhs, off_vw, off_mid, off_vmid, off_seg_roll, off_seg_frc;
  size_t off_ws_roll, off_ws_up, off_ws_frc, total;
};

static int fj_case_counts(const fj_step_case& K, int& nb, int& s_b, int& s_w) {
  if (K.m < 1 || K.n < 1 || K.m_star < 0) return FJ_ERR_ARG;
  nb = K.m * K.n;
  s_b = K.m * (K.n + 1) + (K.m + 1) * K.n;
  s_w = K.m_star > 0 ? K.m_star * (K.n + 1) + (K.m_star + 1) * K.n : 0;
  return FJ_OK;
}

// fills the case records (when recs is non-NULL) and the workspace layout
static int fj_tick_layout(int n_cases, const fj_step_case* cases, int free_wake, FjCaseDev* recs,
                          FjTaskDesc* roll, FjTaskDesc* up, FjTaskDesc* frc, FjTickLayout& L) {
  for (int i = 0; i < n_cases; ++i) {
    const fj_step_case& K = cases[i];
    int nb, s_b, s_w;
    if (fj_case_counts(K, nb, s_b, s_w) != FJ_OK) return FJ_ERR_ARG;
    const int V = (K.m + 1) * (K.n + 1);
    if (recs) {
      FjCaseDev& C = recs[i];
      C.c = K;
      C.nb = nb; C.s_b = s_b; C.s_w = s_w;
      C.p_off = L.nb_tot; C.mid_off = L.mid_tot;
      C.seg_roll = L.roll_rows; C.seg_frc = L.frc_rows;
      C.geo_off = L.geo_tot; C.roll_off = L.roll_tot; C.wk_off = L.wk_tot; C.bnd_off = L.bnd_tot;
      C.aic_off = L.aic_tot; C.rhs_off = L.nb_tot; C.rblk_off = L.rblk_tot; C.frc_off = L.frc_tot;
    }
    if (roll) {
      FjTaskDesc& d = roll[i];
      d.tgt = K.zeta_w_nm1;
      d.out = K.zeta_w + 3 * (size_t)(K.n + 1);
      d.n_tgt = (free_wake && K.m_star > 0) ? K.m_star * (K.n + 1) : 0;
      d.seg_off = L.roll_rows; d.n_seg = s_b + s_w;
      d.mode = FJ_OUT_CONVECT; d.dt = K.dt; d.time = K.time;
      // sources: the bound and the wake lattice at n-1
      d.n_lat = 2;
      d.lat_off[0] = L.roll_rows; d.lat_m[0] = K.m; d.lat_n[0] = K.n;
      d.lat_off[1] = L.roll_rows + s_b; d.lat_m[1] = K.m_star; d.lat_n[1] = K.n;
    }
    L.geo_tot += V + nb + s_b + K.m_star * K.n + (K.n + 1) + (free_wake ? 0 : K.m_star * (K.n + 1));
    L.roll_tot += s_b + s_w;
    L.wk_tot += s_w;
    L.bnd_tot += s_b;
    L.aic_tot += nb * fj_cdiv(K.n, FJ_AIC_W);
    L.rblk_tot += fj_cdiv(nb, FJ_RV_WAVES);
    L.frc_tot += V + nb;
    L.nb_tot += nb;
    L.mid_tot += s_b;
    L.roll_rows += s_b + s_w;
    L.frc_rows += s_b + s_w;
  }
  return FJ_OK;
}

static void fj_tick_tasks(int n_cases, const fj_step_case* cases, const FjCaseDev* recs,
                          double* c, double* vw, double* mids, double* vmid, FjTaskDesc* up,
                          FjTaskDesc* frc) {
  for (int i = 0; i < n_cases; ++i) {
    const FjCaseDev& C = recs[i];
    FjTaskDesc& u = up[i];
    u.tgt = c + 3 * (size_t)C.p_off; u.out = vw + 3 * (size_t)C.p_off;
    u.n_tgt = C.nb; u.seg_off = C.seg_frc + C.s_b; u.n_seg = C.s_w;
    u.mode = FJ_OUT_STORE; u.dt = 0.0; u.time = cases[i].time;
    // sources: the wake lattice at n
    u.n_lat = 1;
    u.lat_off[0] = C.seg_frc + C.s_b; u.lat_m[0] = cases[i].m_star; u.lat_n[0] = cases[i].n;
    FjTaskDesc& f = frc[i];
    f.tgt = mids + 3 * (size_t)C.mid_off; f.out = vmid + 3 * (size_t)C.mid_off;
    f.n_tgt = C.s_b; f.seg_off = C.seg_frc; f.n_seg = C.s_b + C.s_w;
    f.mode = FJ_OUT_STORE; f.dt = 0.0; f.time = cases[i].time;
    // sources: the bound and the wake lattice at n
    f.n_lat = 2;
    f.lat_off[0] = C.seg_frc; f.lat_m[0] = cases[i].m; f.lat_n[0] = cases[i].n;
    f.lat_off[1] = C.seg_frc + C.s_b; f.lat_m[1] = cases[i].m_star; f.lat_n[1] = cases[i].n;
  }
}

static void fj_tick_offsets(FjTickLayout& L, const FjTaskDesc* roll, const FjTaskDesc* up,
                            const FjTaskDesc* frc, int n_cases) {
  size_t o = 0;
  auto take = [&](size_t bytes) { const size_t at = o; o += fj_align256(bytes); return at; };
  L.off_c = take(3 * sizeof(double) * (size_t)L.nb_tot);
  L.off_nc = take(3 * sizeof(double) * (size_t)L.nb_tot);
  L.off_cd = take(3 * sizeof(double) * (size_t)L.nb_tot);
  L.off_rhs = take(sizeof(double) * (size_t)L.nb_tot);
  L.off_vw = take(3 * sizeof(double) * (size_t)L.nb_tot);
  L.off_mid = take(3 * sizeof(double) * (size_t)L.mid_tot);
  L.off_vmid = take(3 * sizeof(double) * (size_t)L.mid_tot);
  L.off_seg_roll = take(FJ_SEG_STRIDE * sizeof(double) * (size_t)L.roll_rows);
  L.off_seg_frc = take(FJ_SEG_STRIDE * sizeof(double) * (size_t)L.frc_rows);
  // the rollup and force pair sums run in either form (FP64 lattice or
  // mixed table), the upwash always in the FP64 lattice form
  auto either = [&](const FjTaskDesc* d) {
    const size_t a = fj_batched_ws_bytes(d, n_cases, FJ_PREC_FP64);
    const size_t b = fj_batched_ws_bytes(d, n_cases, FJ_PREC_MIXED);
    return a > b ? a : b;
  };
  L.off_ws_roll = take(either(roll));
  L.off_ws_up = take(fj_batched_ws_bytes(up, n_cases, FJ_PREC_FP64));
  L.off_ws_frc = take(either(frc));
  L.total = o;
}

// the pair-sum sizes depend only on the case shapes, so the workspace query
// builds the task descriptors with placeholder pointers
static int fj_tick_plan(int n_cases, const fj_step_case* cases, FjTickLayout& L,
                        std::vector<FjCaseDev>& recs, std::vector<FjTaskDesc>& roll,
                        std::vector<FjTaskDesc>& up, std::vector<FjTaskDesc>& frc, int free_wake) {
  recs.resize(n_cases); roll.resize(n_cases); up.resize(n_cases); frc.resize(n_cases);
  const int rc = fj_tick_layout(n_cases, cases, free_wake, recs.data(), roll.data(), up.data(),
                                frc.data(), L);
  if (rc != FJ_OK) return rc;
  fj_tick_tasks(n_cases, cases, recs.data(), nullptr, nullptr, nullptr, nullptr, up.data(),
                frc.data());
  fj_tick_offsets(L, roll.data(), up.data(), frc.data(), n_cases);
  return FJ_OK;
}

extern "C" size_t fj_step_batched_workspace_bytes(int n_cases, const fj_step_case* cases) {
  if (n_cases <= 0 || !cases) return 256;
  FjTickLayout L;
  std::vector<FjCaseDev> recs;
  std::vector<FjTaskDesc> roll, up, frc;
  if (fj_tick_plan(n_cases, cases, L, recs, roll, up, frc, 1) != FJ_OK) return 0;
  return L.total + 256;
}

extern "C" size_t fj_step_batched_table_bytes(int n_cases) {
  if (n_cases < 1) n_cases = 1;
  return fj_align256((size_t)n_cases * sizeof(FjCaseDev)) +
         3 * fj_align256(fj_batched_task_bytes(n_cases));
}

static int fj_step_batched_run(int n_cases, const fj_step_case* cases, const fj_flowfield* flow,
                               double rho, double relax, int kernel, double eps, double rcut,
                               int precision, int free_wake, int max_iter, double tol,
                               void* device_table, void* host_staging, void* workspace,
                               void* stream, bool launch) {
  if (n_cases <= 0) return FJ_OK;
  if (!cases || !flow || !device_table || !host_staging || !workspace) return FJ_ERR_ARG;
  if (kernel < FJ_KERNEL_EPSILON || kernel > FJ_KERNEL_PLAIN) return FJ_ERR_ARG;
  if (precision != FJ_PREC_FP64 && precision != FJ_PREC_MIXED) return FJ_ERR_ARG;
  if (max_iter < 1 || max_iter > FJ_REFINE_MAX_ITER) return FJ_ERR_ARG;
  for (int i = 0; i < n_cases; ++i) {
    const fj_step_case& K = cases[i];
    if (!K.hg || !K.hg_dot || !K.dof || !K.zeta0 || !K.zeta_b_nm1 || !K.zeta_b ||
        !K.zeta_b_dot || !K.gamma_b_nm1 || !K.zeta_w_nm1 || !K.zeta_w || !K.aic || !K.aic_inv ||
        !K.gamma_b || !K.gamma_b_dot || !K.f_steady || !K.lift || !K.refine_ws || !K.iters ||
        !(K.dt > 0.0))
      return FJ_ERR_ARG;
    if (K.m_star > 0 && (!K.gamma_w_nm1 || !K.gamma_w)) return FJ_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  FjTickLayout L;
  std::vector<FjCaseDev> recs;
  std::vector<FjTaskDesc> roll, up, frc;
  int rc = fj_tick_plan(n_cases, cases, L, recs, roll, up, frc, free_wake);
  if (rc != FJ_OK) return rc;
  char* ws = (char*)workspace;
  double* c = (double*)(ws + L.off_c);
  double* nc = (double*)(ws + L.off_nc);
  double* cd = (double*)(ws + L.off_cd);
  double* rhs = (double*)(ws + L.off_rhs);
  double* vw = (double*)(ws + L.off_vw);
  double* mids = (double*)(ws + L.off_mid);
  double* vmid = (double*)(ws + L.off_vmid);
  double* seg_roll = (double*)(ws + L.off_seg_roll);
  double* seg_frc = (double*)(ws + L.off_seg_frc);
  fj_tick_tasks(n_cases, cases, recs.data(), c, vw, mids, vmid, up.data(), frc.data());
  const FjFlow fl = fj_flow_from_desc(flow);
  for (int i = 0; i < n_cases; ++i) {
    roll[i].flow = fl; up[i].flow = fl; frc[i].flow = fl;
  }
  FjKernelParams kp;
  kp.eps = eps; kp.eps2 = eps * eps; kp.rcut = rcut;

  // tables: case records, then the three pair-sum task tables (device and
  // host staging share the layout)
  const size_t rec_bytes = fj_align256((size_t)n_cases * sizeof(FjCaseDev));
  const size_t task_bytes = fj_align256(fj_batched_task_bytes(n_cases));
  char* dt_base = (char*)device_table;
  char* hs_base = (char*)host_staging;
  memcpy(hs_base, recs.data(), (size_t)n_cases * sizeof(FjCaseDev));
  if (!launch) {
    // host staging only: the three task tables as the launches below write them
    if (free_wake) {
      rc = fj_batched_launch(roll.data(), n_cases, seg_roll, kernel, precision, kp, dt_base + rec_bytes,
                             hs_base + rec_bytes, ws + L.off_ws_roll, st, false);
      if (rc != FJ_OK) return rc;
    }
    rc = fj_batched_launch(up.data(), n_cases, seg_frc, kernel, FJ_PREC_FP64, kp,
                           dt_base + rec_bytes + task_bytes, hs_base + rec_bytes + task_bytes,
                           ws + L.off_ws_up, st, false);
    if (rc != FJ_OK) return rc;
    return fj_batched_launch(frc.data(), n_cases, seg_frc, kernel, precision, kp,
                             dt_base + rec_bytes + 2 * task_bytes, hs_base + rec_bytes + 2 * task_bytes,
                             ws + L.off_ws_frc, st, false);
  }
  if (hipMemcpyAsync(dt_base, hs_base, (size_t)n_cases * sizeof(FjCaseDev),
                     hipMemcpyHostToDevice, st) != hipSuccess)
    return FJ_ERR_LAUNCH;
  const FjCaseDev* dcs = (const FjCaseDev*)dt_base;
  auto grid_of = [](int total) { return dim3(fj_cdiv(total > 0 ? total : 1, FJ_SV_EW) < 65536
                                                 ? fj_cdiv(total > 0 ? total : 1, FJ_SV_EW)
                                                 : 65536); };

  // 1. geometry, wake shift and TE row (or prescribed convection)
  hipLaunchKernelGGL(k_bt_geometry, grid_of(L.geo_tot), dim3(FJ_SV_EW), 0, st, dcs, n_cases,
                     L.geo_tot, fl, free_wake, c, nc, cd, mids);
  FJ_CHECK_LAUNCH();
  // 2. rollup tables and the fused rollup with the convection epilogue
  if (free_wake) {
    hipLaunchKernelGGL(k_bt_segments, grid_of(L.roll_tot), dim3(FJ_SV_EW), 0, st, dcs, n_cases,
                       L.roll_tot, 0, kernel, eps, seg_roll);
    FJ_CHECK_LAUNCH();
    rc = fj_batched_launch(roll.data(), n_cases, seg_roll, kernel, precision, kp,
                           dt_base + rec_bytes, hs_base + rec_bytes, ws + L.off_ws_roll, st);
    if (rc != FJ_OK) return rc;
  }
  // 3. wake tables at n and the fused upwash at the collocation points
  if (L.wk_tot > 0) {
    hipLaunchKernelGGL(k_bt_segments, grid_of(L.wk_tot), dim3(FJ_SV_EW), 0, st, dcs, n_cases,
                       L.wk_tot, 1, kernel, eps, seg_frc);
    FJ_CHECK_LAUNCH();
  }
  // the upwash stays FP64 in both forms (it enters the right-hand side)
  rc = fj_batched_launch(up.data(), n_cases, seg_frc, kernel, FJ_PREC_FP64, kp,
                         dt_base + rec_bytes + task_bytes, hs_base + rec_bytes + task_bytes,
                         ws + L.off_ws_up, st);
  if (rc != FJ_OK) return rc;
  // 4. AIC, right-hand sides, refinement
  {
    const dim3 grid(fj_cdiv(L.aic_tot, FJ_AIC_THREADS)), blk(FJ_AIC_THREADS);
    if (kernel == FJ_KERNEL_EPSILON)
      hipLaunchKernelGGL(k_bt_aic_strips<FJ_KERNEL_EPSILON>, grid, blk, 0, st, dcs, n_cases,
                         L.aic_tot, c, nc, kp);
    else if (kernel == FJ_KERNEL_CUTOFF)
      hipLaunchKernelGGL(k_bt_aic_strips<FJ_KERNEL_CUTOFF>, grid, blk, 0, st, dcs, n_cases,
                         L.aic_tot, c, nc, kp);
    else
      hipLaunchKernelGGL(k_bt_aic_strips<FJ_KERNEL_PLAIN>, grid, blk, 0, st, dcs, n_cases,
                         L.aic_tot, c, nc, kp);
  }
  FJ_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bt_rhs, grid_of(L.nb_tot), dim3(FJ_SV_EW), 0, st, dcs, n_cases, L.nb_tot,
                     fl, c, nc, cd, vw, rhs);
  FJ_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bt_ref_init, dim3(n_cases), dim3(64), 0, st, dcs, max_iter);
  hipLaunchKernelGGL(k_bt_ref_apply<false>, dim3(L.rblk_tot), dim3(64 * FJ_RV_WAVES), 0, st, dcs,
                     n_cases, rhs, 0, tol);
  for (int k = 0; k < max_iter; ++k) {
    hipLaunchKernelGGL(k_bt_ref_residual, dim3(L.rblk_tot), dim3(64 * FJ_RV_WAVES), 0, st, dcs,
                       n_cases, rhs, k);
    hipLaunchKernelGGL(k_bt_ref_apply<true>, dim3(L.rblk_tot), dim3(64 * FJ_RV_WAVES), 0, st, dcs,
                       n_cases, rhs, k, tol);
  }
  hipLaunchKernelGGL(k_bt_ref_residual, dim3(L.rblk_tot), dim3(64 * FJ_RV_WAVES), 0, st, dcs,
                     n_cases, rhs, max_iter);
  hipLaunchKernelGGL(k_bt_ref_finish, dim3(fj_cdiv(n_cases, 64)), dim3(64), 0, st, dcs, n_cases,
                     max_iter, tol);
  FJ_CHECK_LAUNCH();
  // 5. bound tables at n, fused force pair sum, forces, rates, lift
  hipLaunchKernelGGL(k_bt_segments, grid_of(L.bnd_tot), dim3(FJ_SV_EW), 0, st, dcs, n_cases,
                     L.bnd_tot, 2, kernel, eps, seg_frc);
  FJ_CHECK_LAUNCH();
  rc = fj_batched_launch(frc.data(), n_cases, seg_frc, kernel, precision, kp,
                         dt_base + rec_bytes + 2 * task_bytes, hs_base + rec_bytes + 2 * task_bytes,
                         ws + L.off_ws_frc, st);
  if (rc != FJ_OK) return rc;
  hipLaunchKernelGGL(k_bt_forces, grid_of(L.frc_tot), dim3(FJ_SV_EW), 0, st, dcs, n_cases,
                     L.frc_tot, fl, rho, relax, vmid);
  FJ_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bt_lift, dim3(n_cases), dim3(256), 0, st, dcs);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

extern "C" int fj_step_batched(int n_cases, const fj_step_case* cases, const fj_flowfield* flow,
                               double rho, double relax, int kernel, double eps, double rcut,
                               int precision, int free_wake, int max_iter, double tol,
                               void* device_table, void* host_staging, void* workspace,
                               void* stream) {
  return fj_step_batched_run(n_cases, cases, flow, rho, relax, kernel, eps, rcut, precision, free_wake,
                             max_iter, tol, device_table, host_staging, workspace, stream, true);
}

extern "C" int fj_step_batched_restage(int n_cases, const fj_step_case* cases, const fj_flowfield* flow,
                                       double rho, double relax, int kernel, double eps, double rcut,
                                       int precision, int free_wake, int max_iter, double tol,
                                       void* device_table, void* host_staging, void* workspace,
                                       void* stream) {
  return fj_step_batched_run(n_cases, cases, flow, rho, relax, kernel, eps, rcut, precision, free_wake,
                             max_iter, tol, device_table, host_staging, workspace, stream, false);
}
