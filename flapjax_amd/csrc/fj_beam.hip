// Geometrically exact SE(3) beam on the device: the Newton system of the
// static and the generalised-alpha dynamic solve (structure/beam.py:1878-2861
// with structure/utils.py:_k_t_entry, _integrate_m_l, _integrate_c_t), the
// node-frame updates, and the aero -> beam force projection
// (aero/utils.py:666-697).
//
// One Newton iteration is four launches plus the dense LU of fj_lu.hip:
//   k_beam_points    a thread per (element, quadrature slot): the relative
//                    configuration d = log(H_A^-1 H_B) of the element frames,
//                    P(d), strain and internal force (slot 0), Q at the Lobatto
//                    midpoint (slot 1), Q and its JVP along d_dot at the three
//                    Gauss-Legendre points (slots 2-4, dual numbers)
//   k_beam_matrices  a workgroup per element, a thread per entry of the
//                    12 x 12 blocks: K_t = P^T K P / l, M_t, C_L, C_T, the
//                    element's share of the system matrix and of the residual
//   k_beam_assemble  a thread per entry of the reduced (solve-DOF) system and
//                    per reduced residual entry, summing the element blocks
//                    that meet at the two nodes (node -> element CSR)
//   k_beam_update    one workgroup: increments, state update and the four
//                    convergence norms the host loop tests
// The reference's optional Jacobians (geometric and dead-load stiffness, q_dot
// in C_T) are off by default there and are not part of this path.
#include "fj_common.hpp"
#include "fj_beam_dev.hpp"

// per-element workspace (doubles)
#define FJ_BW_D 0
#define FJ_BW_EPS 6
#define FJ_BW_P 12
#define FJ_BW_FINT 84
#define FJ_BW_QM 96
#define FJ_BW_QL 168  // 3 x {Q 72, Q_dot 72, u = Q v 6, w = M u 6}
#define FJ_BW_QL_STRIDE 156
#define FJ_BW_S 636
#define FJ_BW_FRES 780
#define FJ_BW_FABS 792
#define FJ_BW_STRIDE 804

#define FJ_BEAM_MAT_THREADS 256

// Gauss-Legendre order 3 on [0, 1] (algebra/integration.py:gauss_legendre)
__constant__ double fj_leg_s[3] = {0.11270166537925831148207346002176, 0.5,
                                   0.88729833462074168851792653997824};
__constant__ double fj_leg_w[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};

struct FjBeamDev {
  int n_nodes, n_elem;
  const int32_t* conn;
  const double* o0;
  const double* l0;
  const double* d0;
  const double* k_cs;
  const double* m_cs;
  const int32_t* node_off;
  const int32_t* node_elem;
};

// A guarded launch returns at once when the Newton status word says the
// loop has finished (status[1] != 0): the device-driven Newton loop enqueues
// iterations ahead and lets the ones after convergence fall through.
__device__ __forceinline__ bool fj_guard_done(const int32_t* guard) { return guard && guard[1] != 0; }

static FjBeamDev fj_beam_dev(const fj_beam_model* m) {
  FjBeamDev d;
  d.n_nodes = m->n_nodes; d.n_elem = m->n_elem;
  d.conn = m->conn; d.o0 = m->o0; d.l0 = m->l0; d.d0 = m->d0;
  d.k_cs = m->k_cs; d.m_cs = m->m_cs; d.node_off = m->node_off; d.node_elem = m->node_elem;
  return d;
}

// element frames H = hg_node . [[O, 0], [0, 1]] and d = log(H_A^-1 H_B)
// (structure/beam.py:1482-1500)
__device__ void fj_beam_d(const FjBeamDev& bm, const double* hg, int e, double* d) {
  const int a = bm.conn[2 * e], b = bm.conn[2 * e + 1];
  const double* o = bm.o0 + 9 * e;
  double ha[16], hb[16], r[16];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      ha[4 * i + j] = hg[16 * a + 4 * i] * o[j] + hg[16 * a + 4 * i + 1] * o[3 + j] + hg[16 * a + 4 * i + 2] * o[6 + j];
      hb[4 * i + j] = hg[16 * b + 4 * i] * o[j] + hg[16 * b + 4 * i + 1] * o[3 + j] + hg[16 * b + 4 * i + 2] * o[6 + j];
    }
    ha[4 * i + 3] = hg[16 * a + 4 * i + 3];
    hb[4 * i + 3] = hg[16 * b + 4 * i + 3];
  }
  ha[12] = ha[13] = ha[14] = 0.0; ha[15] = 1.0;
  hb[12] = hb[13] = hb[14] = 0.0; hb[15] = 1.0;
  fj_hg_inv_mul(ha, hb, r);
  fj_log_se3(r, d);
}

__device__ __forceinline__ void fj_beam_elem_vec(const FjBeamDev& bm, const double* x, int e, double* out) {
  const int a = bm.conn[2 * e], b = bm.conn[2 * e + 1];
  for (int k = 0; k < 6; ++k) {
    out[k] = x[6 * a + k];
    out[6 + k] = x[6 * b + k];
  }
}

template <int DYN>
__global__ __launch_bounds__(64) void k_beam_points(FjBeamDev bm, const double* __restrict__ hg, const double* __restrict__ v,
                              double* __restrict__ ws, const int32_t* __restrict__ guard) {
  if (fj_guard_done(guard)) return;
  const int slots = DYN ? 5 : 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= bm.n_elem * slots) return;
  const int e = t / slots, slot = t - e * slots;
  double* w = ws + (size_t)FJ_BW_STRIDE * e;
  const double* o = bm.o0 + 9 * e;
  double d[6];
  fj_beam_d(bm, hg, e, d);
  if (slot == 0) {
    // P(d), eps = (d - d0) / l, f_int = -P^T K eps (structure/beam.py:1329-1345, 1463-1480)
    double p[6][12];
    fj_p_mat(d, o, p);
    const double l = bm.l0[e];
    const double* k = bm.k_cs + 36 * e;
    double eps[6], ke[6];
    for (int i = 0; i < 6; ++i) eps[i] = (d[i] - bm.d0[6 * e + i]) / l;
    for (int i = 0; i < 6; ++i) {
      double acc = 0.0;
      for (int j = 0; j < 6; ++j) acc += k[6 * i + j] * eps[j];
      ke[i] = acc;
    }
    for (int i = 0; i < 6; ++i) {
      w[FJ_BW_D + i] = d[i];
      w[FJ_BW_EPS + i] = eps[i];
    }
    for (int i = 0; i < 6; ++i)
      for (int c = 0; c < 12; ++c) w[FJ_BW_P + 12 * i + c] = p[i][c];
    for (int c = 0; c < 12; ++c) {
      double acc = 0.0;
      for (int i = 0; i < 6; ++i) acc += p[i][c] * ke[i];
      w[FJ_BW_FINT + c] = -acc;
    }
  } else if (slot == 1) {
    double q[6][12];
    fj_q_mat<double>(0.5, d, o, q);
    for (int i = 0; i < 6; ++i)
      for (int c = 0; c < 12; ++c) w[FJ_BW_QM + 12 * i + c] = q[i][c];
  } else {
    // Q and Q_dot at a Legendre point: d_dot = P(d) v_AB, one dual pass
    const int k = slot - 2;
    double p[6][12], ve[12];
    fj_p_mat(d, o, p);
    fj_beam_elem_vec(bm, v, e, ve);
    FjDual dd[6];
    for (int i = 0; i < 6; ++i) {
      double acc = 0.0;
      for (int c = 0; c < 12; ++c) acc += p[i][c] * ve[c];
      dd[i] = FjDual(d[i], acc);
    }
    FjDual q[6][12];
    fj_q_mat<FjDual>(fj_leg_s[k], dd, o, q);
    double* wq = w + FJ_BW_QL + FJ_BW_QL_STRIDE * k;
    double u[6];
    for (int i = 0; i < 6; ++i) {
      double acc = 0.0;
      for (int c = 0; c < 12; ++c) {
        wq[12 * i + c] = q[i][c].v;
        wq[72 + 12 * i + c] = q[i][c].d;
        acc += q[i][c].v * ve[c];
      }
      u[i] = acc;
    }
    const double* m = bm.m_cs + 36 * e;
    for (int i = 0; i < 6; ++i) {
      double acc = 0.0;
      for (int j = 0; j < 6; ++j) acc += m[6 * i + j] * u[j];
      wq[144 + i] = u[i];
      wq[150 + i] = acc;
    }
  }
}

template <int DYN>
__global__ __launch_bounds__(FJ_BEAM_MAT_THREADS) void k_beam_matrices(
    FjBeamDev bm, const double* __restrict__ v, const double* __restrict__ v_dot,
    const double* __restrict__ t_phi, double beta_p, double gamma_p, double* __restrict__ ws,
    const int32_t* __restrict__ guard) {
  if (fj_guard_done(guard)) return;
  const int e = blockIdx.x;
  const int t = threadIdx.x;
  double* w = ws + (size_t)FJ_BW_STRIDE * e;
  const double l = bm.l0[e];
  __shared__ double sP[72], sK[36], sF[12];
  __shared__ double sQm[72], sQ[3][72], sQd[3][72], sU[3][6], sW[3][6], sM[36];
  __shared__ double sH[3][36], sC[3][36], sA0[36], sV[12], sVd[12], sT[2][36];
  __shared__ double sKe[144], sMe[144], sCL[144];
  for (int i = t; i < 72; i += blockDim.x) sP[i] = w[FJ_BW_P + i];
  for (int i = t; i < 36; i += blockDim.x) sK[i] = bm.k_cs[36 * e + i];
  for (int i = t; i < 12; i += blockDim.x) sF[i] = w[FJ_BW_FINT + i];
  const int a = bm.conn[2 * e], b = bm.conn[2 * e + 1];
  if (DYN) {
    for (int i = t; i < 72; i += blockDim.x) sQm[i] = w[FJ_BW_QM + i];
    for (int i = t; i < 3 * 72; i += blockDim.x) {
      const int k = i / 72, j = i - 72 * k;
      sQ[k][j] = w[FJ_BW_QL + FJ_BW_QL_STRIDE * k + j];
      sQd[k][j] = w[FJ_BW_QL + FJ_BW_QL_STRIDE * k + 72 + j];
    }
    for (int i = t; i < 18; i += blockDim.x) {
      const int k = i / 6, j = i - 6 * k;
      sU[k][j] = w[FJ_BW_QL + FJ_BW_QL_STRIDE * k + 144 + j];
      sW[k][j] = w[FJ_BW_QL + FJ_BW_QL_STRIDE * k + 150 + j];
    }
    for (int i = t; i < 36; i += blockDim.x) {
      sM[i] = bm.m_cs[36 * e + i];
      sT[0][i] = t_phi[36 * a + i];
      sT[1][i] = t_phi[36 * b + i];
    }
    for (int i = t; i < 6; i += blockDim.x) {
      sV[i] = v[6 * a + i]; sV[6 + i] = v[6 * b + i];
      sVd[i] = v_dot[6 * a + i]; sVd[6 + i] = v_dot[6 * b + i];
    }
  }
  __syncthreads();
  if (DYN) {
    // H_k = hat(u_k)^T M, C_k = check(w_k)^T, A0 = Ad^T M Ad
    if (t < 108) {
      const int k = t / 36, i = (t % 36) / 6, j = t % 6;
      double hat[6][6];
      fj_hat6(sU[k], hat);
      double acc = 0.0;
      for (int m = 0; m < 6; ++m) acc += hat[m][i] * sM[6 * m + j];
      sH[k][6 * i + j] = acc;
    } else if (t < 216) {
      const int k = (t - 108) / 36, i = ((t - 108) % 36) / 6, j = (t - 108) % 6;
      // check(h) = -[[0, S(h_u)], [S(h_u), S(h_w)]]; entry (j, i) of it
      const double* h = sW[k];
      double s_u[3][3], s_w[3][3];
      fj_skew(h, s_u);
      fj_skew(h + 3, s_w);
      double c;
      if (j < 3) c = (i < 3) ? 0.0 : -s_u[j][i - 3];
      else c = (i < 3) ? -s_u[j - 3][i] : -s_w[j - 3][i - 3];
      sC[k][6 * i + j] = c;
    } else if (t < 252) {
      const int i = (t - 216) / 6, j = (t - 216) % 6;
      // Ad = blockdiag(O^T, O^T): Ad[3B + p][3B + q] = O[q][p]
      const double* o = bm.o0 + 9 * e;
      const int bi = (i / 3) * 3, bj = (j / 3) * 3, ii = i % 3, jj = j % 3;
      double acc = 0.0;
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) acc += o[3 * ii + p] * sM[6 * (bi + p) + bj + q] * o[3 * jj + q];
      sA0[6 * i + j] = acc;
    }
    __syncthreads();
  }
  double s_rc = 0.0;
  const int r = t / 12, c = t % 12;
  if (t < 144) {
    double ke = 0.0;
    for (int i = 0; i < 6; ++i) {
      double kp = 0.0;
      for (int j = 0; j < 6; ++j) kp += sK[6 * i + j] * sP[12 * j + c];
      ke += sP[12 * i + r] * kp;
    }
    ke = ke / l;
    if (DYN) {
      double mq = 0.0;
      for (int i = 0; i < 6; ++i) {
        double x = 0.0;
        for (int j = 0; j < 6; ++j) x += sM[6 * i + j] * sQm[12 * j + c];
        mq += sQm[12 * i + r] * x;
      }
      double f01 = 0.0;
      if (r < 6 && c < 6) f01 = sA0[6 * r + c];
      else if (r >= 6 && c >= 6) f01 = sA0[6 * (r - 6) + c - 6];
      const double me = l * (0.5 * (2.0 / 6.0 * f01 + 4.0 / 3.0 * mq));
      double cl = 0.0, ct = 0.0;
      for (int k = 0; k < 3; ++k) {
        double x = 0.0, y = 0.0;
        for (int i = 0; i < 6; ++i) {
          double xi = 0.0, yi = 0.0;
          for (int j = 0; j < 6; ++j) {
            xi += sM[6 * i + j] * sQd[k][12 * j + c] - sH[k][6 * i + j] * sQ[k][12 * j + c];
            yi += sC[k][6 * i + j] * sQ[k][12 * j + c];
          }
          x += sQ[k][12 * i + r] * xi;
          y += sQ[k][12 * i + r] * yi;
        }
        cl += fj_leg_w[k] * x;
        ct += fj_leg_w[k] * (x - y);
      }
      cl = 0.5 * l * cl;
      ct = 0.5 * l * ct;
      s_rc = beta_p * me + gamma_p * ct;
      sKe[t] = ke;
      sMe[t] = me;
      sCL[t] = cl;
    } else {
      s_rc = ke;
    }
  }
  if (DYN) __syncthreads();
  if (t < 144) {
    if (DYN) {
      // K_t T(phi): the node block of column c takes T of that node
      const int bc = c / 6, cc = c % 6;
      double kt = 0.0;
      for (int m = 0; m < 6; ++m) kt += sKe[12 * r + 6 * bc + m] * sT[bc][6 * m + cc];
      s_rc += kt;
    }
    w[FJ_BW_S + t] = s_rc;
  }
  if (t < 12) {
    double fr = sF[t], fa = fabs(sF[t]);
    if (DYN) {
      double ig = 0.0;
      for (int j = 0; j < 12; ++j) ig -= sMe[12 * t + j] * sVd[j];
      double gy = 0.0;
      for (int j = 0; j < 12; ++j) gy -= sCL[12 * t + j] * sV[j];
      fr = fr + (ig + gy);
      fa = fa + fabs(ig + gy);
    }
    w[FJ_BW_FRES + t] = fr;
    w[FJ_BW_FABS + t] = fa;
  }
}

// local component c of a node's global dead load: R^T f (structure/beam.py:1387-1395)
__device__ __forceinline__ double fj_dead_local(const double* hg, const double* f, int node, int comp) {
  const double* g = hg + 16 * node;
  const int base = comp < 3 ? 0 : 3, cc = comp % 3;
  return g[cc] * f[6 * node + base] + g[4 + cc] * f[6 * node + base + 1] + g[8 + cc] * f[6 * node + base + 2];
}

__global__ void k_beam_assemble(FjBeamDev bm, const int32_t* __restrict__ sd, int n_s,
                                const double* __restrict__ hg, const double* __restrict__ f_fol,
                                const double* __restrict__ f_dead, const double* __restrict__ ws,
                                double* __restrict__ A, double* __restrict__ rv, double* __restrict__ fav,
                                const int32_t* __restrict__ guard) {
  if (fj_guard_done(guard)) return;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nn = (long long)n_s * n_s;
  if (t < nn) {
    const int i = (int)(t / n_s), j = (int)(t - (long long)i * n_s);
    const int di = sd[i], dj = sd[j];
    const int na = di / 6, ca = di % 6, nb = dj / 6, cb = dj % 6;
    double acc = 0.0;
    for (int q = bm.node_off[na]; q < bm.node_off[na + 1]; ++q) {
      const int code = bm.node_elem[q];
      const int e = code >> 1, ea = code & 1;
      int eb = -1;
      if (bm.conn[2 * e] == nb) eb = 0;
      else if (bm.conn[2 * e + 1] == nb) eb = 1;
      if (eb < 0) continue;
      acc += ws[(size_t)FJ_BW_STRIDE * e + FJ_BW_S + 12 * (6 * ea + ca) + 6 * eb + cb];
    }
    A[t] = acc;
  } else if (t < nn + n_s) {
    const int i = (int)(t - nn);
    const int di = sd[i], na = di / 6, ca = di % 6;
    double fr = 0.0, fa = 0.0;
    for (int q = bm.node_off[na]; q < bm.node_off[na + 1]; ++q) {
      const int code = bm.node_elem[q];
      const int e = code >> 1, ea = code & 1;
      fr += ws[(size_t)FJ_BW_STRIDE * e + FJ_BW_FRES + 6 * ea + ca];
      fa += ws[(size_t)FJ_BW_STRIDE * e + FJ_BW_FABS + 6 * ea + ca];
    }
    if (f_fol) {
      fr += f_fol[di];
      fa += fabs(f_fol[di]);
    }
    if (f_dead) {
      const double x = fj_dead_local(hg, f_dead, na, ca);
      fr += x;
      fa += fabs(x);
    }
    rv[i] = fr;
    fav[i] = fa;
  }
}

#define FJ_BEAM_UPD_THREADS 256

__device__ __forceinline__ void fj_block_sum_max(double& s0, double& s1, double& s2, double& mx, int& nan,
                                                 double* red, int* redn) {
  const int t = threadIdx.x;
  red[t] = s0; red[FJ_BEAM_UPD_THREADS + t] = s1; red[2 * FJ_BEAM_UPD_THREADS + t] = s2;
  red[3 * FJ_BEAM_UPD_THREADS + t] = mx;
  redn[t] = nan;
  __syncthreads();
  for (int s = FJ_BEAM_UPD_THREADS / 2; s > 0; s >>= 1) {
    if (t < s) {
      red[t] += red[t + s];
      red[FJ_BEAM_UPD_THREADS + t] += red[FJ_BEAM_UPD_THREADS + t + s];
      red[2 * FJ_BEAM_UPD_THREADS + t] += red[2 * FJ_BEAM_UPD_THREADS + t + s];
      red[3 * FJ_BEAM_UPD_THREADS + t] = fmax(red[3 * FJ_BEAM_UPD_THREADS + t], red[3 * FJ_BEAM_UPD_THREADS + t + s]);
      redn[t] |= redn[t + s];
    }
    __syncthreads();
  }
  s0 = red[0]; s1 = red[FJ_BEAM_UPD_THREADS]; s2 = red[2 * FJ_BEAM_UPD_THREADS];
  mx = red[3 * FJ_BEAM_UPD_THREADS];
  nan = redn[0];
}

// Newton update and the convergence norms (utils/data_structures.py:81-151):
// norms = [|dx|, |total|, |r|, max |f_abs|, any NaN in dx]
__global__ __launch_bounds__(FJ_BEAM_UPD_THREADS) void k_beam_update(
    int dynamic, int n_dof, const int32_t* __restrict__ sd, int n_s, const double* __restrict__ x,
    double relax, double gamma_p, double beta_p, double* __restrict__ phi, double* __restrict__ v,
    double* __restrict__ v_dot, const double* __restrict__ rv, const double* __restrict__ fav,
    double* __restrict__ norms, const int32_t* __restrict__ guard) {
  if (fj_guard_done(guard)) return;
  __shared__ double red[4 * FJ_BEAM_UPD_THREADS];
  __shared__ int redn[FJ_BEAM_UPD_THREADS];
  const int t = threadIdx.x;
  if (!dynamic)
    for (int i = t; i < n_dof; i += blockDim.x) phi[i] = 0.0;
  __syncthreads();
  double sdx = 0.0, sr = 0.0, mx = 0.0;
  int nan = 0;
  for (int i = t; i < n_s; i += blockDim.x) {
    const double dx = x[i] * relax;
    const int k = sd[i];
    if (dynamic) {
      phi[k] += dx;
      v[k] += gamma_p * dx;
      v_dot[k] += beta_p * dx;
    } else {
      phi[k] = dx;
    }
    sdx += dx * dx;
    nan |= (dx != dx) ? 1 : 0;
    sr += rv[i] * rv[i];
    mx = fmax(mx, fabs(fav[i]));
  }
  __syncthreads();
  double st = 0.0;
  if (dynamic)
    for (int i = t; i < n_dof; i += blockDim.x) st += phi[i] * phi[i];
  fj_block_sum_max(sdx, st, sr, mx, nan, red, redn);
  if (t == 0) {
    norms[0] = sqrt(sdx);
    norms[1] = sqrt(st);
    norms[2] = sqrt(sr);
    norms[3] = mx;
    norms[4] = nan ? 1.0 : 0.0;
  }
}

__global__ __launch_bounds__(FJ_BEAM_UPD_THREADS) void k_beam_subnorm(const double* __restrict__ x,
                                                                       const int32_t* __restrict__ idx, int n,
                                                                       double* __restrict__ out) {
  __shared__ double red[4 * FJ_BEAM_UPD_THREADS];
  __shared__ int redn[FJ_BEAM_UPD_THREADS];
  double s = 0.0, z1 = 0.0, z2 = 0.0, z3 = 0.0;
  int zn = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += x[idx[i]] * x[idx[i]];
  fj_block_sum_max(s, z1, z2, z3, zn, red, redn);
  if (threadIdx.x == 0) *out = sqrt(s);
}

// node frames: update H exp(phi) (+ T(phi), + log(H_ref^-1 H)), relative
// configuration log(H^-1 H_ref), or rate H tilde(v)
// H tilde(v) = [[R S(w), R u], [0, 0]] for the frame g = [[R, p], [0, 1]]
// and v = (u, w) (structure/beam.py:1521-1531)
__device__ __forceinline__ void fj_hg_rate(const double* g, const double* vv, double* out) {
  double s[3][3];
  fj_skew(vv + 3, s);
  for (int p = 0; p < 3; ++p) {
    for (int q = 0; q < 3; ++q) out[4 * p + q] = g[4 * p] * s[0][q] + g[4 * p + 1] * s[1][q] + g[4 * p + 2] * s[2][q];
    out[4 * p + 3] = g[4 * p] * vv[0] + g[4 * p + 1] * vv[1] + g[4 * p + 2] * vv[2];
  }
  for (int q = 12; q < 16; ++q) out[q] = 0.0;
}

__global__ void k_beam_nodes(int op, int n, const double* __restrict__ hg, const double* __restrict__ x,
                             const double* __restrict__ hg_ref, double* __restrict__ hg_out,
                             double* __restrict__ t_out, double* __restrict__ h_out,
                             const int32_t* __restrict__ guard) {
  if (fj_guard_done(guard)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* g = hg + 16 * i;
  if (op == FJ_BEAM_NODE_UPDATE) {
    double ex[16], gn[16];
    fj_exp_se3(x + 6 * i, ex);
    fj_hg_mul(g, ex, gn);
    for (int k = 0; k < 16; ++k) hg_out[16 * i + k] = gn[k];
    if (t_out) {
      double tt[6][6];
      fj_t_se3(x + 6 * i, tt);
      for (int p = 0; p < 6; ++p)
        for (int q = 0; q < 6; ++q) t_out[36 * i + 6 * p + q] = tt[p][q];
    }
    if (h_out) {
      double r[16];
      fj_hg_inv_mul(hg_ref + 16 * i, gn, r);
      fj_log_se3(r, h_out + 6 * i);
    }
  } else if (op == FJ_BEAM_NODE_REL) {
    double r[16];
    fj_hg_inv_mul(g, hg_ref + 16 * i, r);
    fj_log_se3(r, h_out + 6 * i);
  } else if (op == FJ_BEAM_NODE_LOCAL || op == FJ_BEAM_NODE_GLOBAL) {
    // R^T f or R f on the force and the moment of a node's 6-vector
    const double* f = x + 6 * i;
    const bool tr = op == FJ_BEAM_NODE_LOCAL;
    for (int h = 0; h < 6; h += 3)
      for (int p = 0; p < 3; ++p) {
        const double r0 = tr ? g[p] : g[4 * p], r1 = tr ? g[4 + p] : g[4 * p + 1], r2 = tr ? g[8 + p] : g[4 * p + 2];
        h_out[6 * i + h + p] = r0 * f[h] + r1 * f[h + 1] + r2 * f[h + 2];
      }
  } else {
    fj_hg_rate(g, x + 6 * i, hg_out + 16 * i);
  }
}

// aero/utils.py:666-697 for one surface: node dof_mapping[j] receives the
// strip sum of f and of (R x0) x f over the m+1 chordwise vertices
__global__ void k_beam_project(int m1, int n1, const double* __restrict__ rmat, const double* __restrict__ fa,
                               const double* __restrict__ fb, const double* __restrict__ x0,
                               const int32_t* __restrict__ dof, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n1) return;
  const int node = dof[j];
  const double* g = rmat + 9 * node;
  double f[3] = {0.0, 0.0, 0.0}, mo[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < m1; ++i) {
    const size_t o = 3 * ((size_t)i * n1 + j);
    double fx = fa[o], fy = fa[o + 1], fz = fa[o + 2];
    if (fb) { fx += fb[o]; fy += fb[o + 1]; fz += fb[o + 2]; }
    const double ax = g[0] * x0[o] + g[1] * x0[o + 1] + g[2] * x0[o + 2];
    const double ay = g[3] * x0[o] + g[4] * x0[o + 1] + g[5] * x0[o + 2];
    const double az = g[6] * x0[o] + g[7] * x0[o + 1] + g[8] * x0[o + 2];
    f[0] += fx; f[1] += fy; f[2] += fz;
    mo[0] += ay * fz - az * fy;
    mo[1] += az * fx - ax * fz;
    mo[2] += ax * fy - ay * fx;
  }
  for (int k = 0; k < 3; ++k) {
    out[6 * node + k] = f[k];
    out[6 * node + 3 + k] = mo[k];
  }
}

__global__ void k_beam_zero(double* __restrict__ x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// ---- C-ABI ----------------------------------------------------------------------
// ---- exact element derivatives for the implicit adjoint -----------------
// The element force f_e(d, v_e, v_dot_e) of the residual (internal
// -P^T K eps, inertial -M_t v_dot, gyroscopic -C_L v; structure/beam.py:
// 1749-1843 with structure/utils.py:_integrate_m_l/_integrate_c_t), written
// once for any scalar S.  The gyroscopic term needs Q_dot = dQ/dd along
// d_dot = P(d) v_e (the reference's JVP), taken on FjDualT<S>; with S a dual
// number this differentiates Q_dot itself.  C_L v is formed from vectors:
// 0.5 l sum_k w_k Q_k^T (M (Q_dot_k v) - hat(u_k)^T M u_k), u_k = Q_k v.
template <typename S>
__device__ void fj_elem_force(const FjBeamDev& bm, int e, const S* d, const S* ve, const double* vde,
                              const double* kmat, int dyn, S fe[12]) {
  const double* o = bm.o0 + 9 * e;
  const double l = bm.l0[e];
  S p[6][12], ke[6];
  fj_p_mat(d, o, p);
  {
    S eps[6];
    for (int i = 0; i < 6; ++i) eps[i] = (d[i] - bm.d0[6 * e + i]) / l;
    for (int i = 0; i < 6; ++i) {
      S acc(0.0);
      for (int j = 0; j < 6; ++j) acc = acc + kmat[6 * i + j] * eps[j];
      ke[i] = acc;
    }
  }
  for (int c = 0; c < 12; ++c) {
    S acc(0.0);
    for (int i = 0; i < 6; ++i) acc = acc + p[i][c] * ke[i];
    fe[c] = -acc;
  }
  if (!dyn) return;
  const double* m = bm.m_cs + 36 * e;
  // inertial: M_t v_dot = l (1/6 blockdiag(A0, A0) v_dot + 2/3 Qm^T M Qm v_dot),
  // A0 = Ad^T M Ad with Ad = blockdiag(O^T, O^T): (Ad x)[3B + p] = sum_j O[j][p] x[3B + j]
  {
    double a0v[12];
    for (int blk = 0; blk < 2; ++blk) {
      double x[6], y[6];
      for (int bb = 0; bb < 2; ++bb)
        for (int pp = 0; pp < 3; ++pp)
          x[3 * bb + pp] = o[pp] * vde[6 * blk + 3 * bb] + o[3 + pp] * vde[6 * blk + 3 * bb + 1] +
                           o[6 + pp] * vde[6 * blk + 3 * bb + 2];
      for (int i = 0; i < 6; ++i) {
        double acc = 0.0;
        for (int j = 0; j < 6; ++j) acc += m[6 * i + j] * x[j];
        y[i] = acc;
      }
      for (int bb = 0; bb < 2; ++bb)
        for (int q = 0; q < 3; ++q)
          a0v[6 * blk + 3 * bb + q] = o[3 * q] * y[3 * bb] + o[3 * q + 1] * y[3 * bb + 1] + o[3 * q + 2] * y[3 * bb + 2];
    }
    S qm[6][12], z[6], mz[6];
    fj_q_mat<S>(0.5, d, o, qm);
    for (int i = 0; i < 6; ++i) {
      S acc(0.0);
      for (int c = 0; c < 12; ++c) acc = acc + qm[i][c] * vde[c];
      z[i] = acc;
    }
    for (int i = 0; i < 6; ++i) {
      S acc(0.0);
      for (int j = 0; j < 6; ++j) acc = acc + m[6 * i + j] * z[j];
      mz[i] = acc;
    }
    for (int c = 0; c < 12; ++c) {
      S acc(0.0);
      for (int i = 0; i < 6; ++i) acc = acc + qm[i][c] * mz[i];
      fe[c] = fe[c] - l * (a0v[c] / 6.0 + (2.0 / 3.0) * acc);
    }
  }
  // gyroscopic
  S dd[6];
  for (int i = 0; i < 6; ++i) {
    S acc(0.0);
    for (int c = 0; c < 12; ++c) acc = acc + p[i][c] * ve[c];
    dd[i] = acc;
  }
  for (int k = 0; k < 3; ++k) {
    FjDualT<S> dk[6], q2[6][12];
    for (int i = 0; i < 6; ++i) dk[i] = FjDualT<S>(d[i], dd[i]);
    fj_q_mat<FjDualT<S>>(fj_leg_s[k], dk, o, q2);
    S u[6], a[6], mu[6], b[6], hh[6][6];
    for (int i = 0; i < 6; ++i) {
      S su(0.0), sa(0.0);
      for (int c = 0; c < 12; ++c) {
        su = su + q2[i][c].v * ve[c];
        sa = sa + q2[i][c].d * ve[c];
      }
      u[i] = su;
      a[i] = sa;
    }
    for (int i = 0; i < 6; ++i) {
      S acc(0.0);
      for (int j = 0; j < 6; ++j) acc = acc + m[6 * i + j] * u[j];
      mu[i] = acc;
    }
    fj_hat6(u, hh);
    for (int i = 0; i < 6; ++i) {
      S acc(0.0);
      for (int j = 0; j < 6; ++j) acc = acc + m[6 * i + j] * a[j] - hh[j][i] * mu[j];
      b[i] = acc;
    }
    const double wl = 0.5 * l * fj_leg_w[k];
    for (int c = 0; c < 12; ++c) {
      S acc(0.0);
      for (int i = 0; i < 6; ++i) acc = acc + q2[i][c].v * b[i];
      fe[c] = fe[c] - wl * acc;
    }
  }
}

// A thread per (element, direction): directions 0-5 seed d, 6-17 the element
// velocities (dynamic only).  Outputs per element: f_e (12), df/dd (12 x 6),
// df/dv_e (12 x 12), M_t (12 x 12; df/dv_dot_e = -M_t), P(d) (6 x 12) and the
// strain eps (6); direction 0 also writes the values.
__global__ __launch_bounds__(64) void k_beam_elem_jac(FjBeamDev bm, int dyn, const double* __restrict__ hg,
                                                      const double* __restrict__ v, const double* __restrict__ v_dot,
                                                      const double* __restrict__ k_elem, double* __restrict__ fe_out,
                                                      double* __restrict__ jd, double* __restrict__ jv,
                                                      double* __restrict__ mt, double* __restrict__ p_out,
                                                      double* __restrict__ eps_out) {
  const int nd = dyn ? 18 : 6;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= bm.n_elem * nd) return;
  const int e = t / nd, k = t - e * nd;
  double dv[6], vev[12], vdev[12];
  fj_beam_d(bm, hg, e, dv);
  if (dyn) {
    fj_beam_elem_vec(bm, v, e, vev);
    fj_beam_elem_vec(bm, v_dot, e, vdev);
  } else {
    for (int c = 0; c < 12; ++c) vev[c] = vdev[c] = 0.0;
  }
  FjDual d[6], ve[12], fe[12];
  for (int i = 0; i < 6; ++i) d[i] = FjDual(dv[i], k == i ? 1.0 : 0.0);
  for (int c = 0; c < 12; ++c) ve[c] = FjDual(vev[c], k == 6 + c ? 1.0 : 0.0);
  const double* kmat = k_elem + 36 * e;
  fj_elem_force<FjDual>(bm, e, d, ve, vdev, kmat, dyn, fe);
  if (k < 6) {
    for (int r = 0; r < 12; ++r) jd[(12 * e + r) * 6 + k] = fe[r].d;
  } else {
    for (int r = 0; r < 12; ++r) jv[(12 * e + r) * 12 + k - 6] = fe[r].d;
  }
  if (k != 0) return;
  const double* o = bm.o0 + 9 * e;
  const double l = bm.l0[e];
  for (int r = 0; r < 12; ++r) fe_out[12 * e + r] = fe[r].v;
  double p[6][12];
  fj_p_mat(dv, o, p);
  for (int i = 0; i < 6; ++i) {
    for (int c = 0; c < 12; ++c) p_out[72 * e + 12 * i + c] = p[i][c];
    eps_out[6 * e + i] = (dv[i] - bm.d0[6 * e + i]) / l;
  }
  if (!dyn) return;
  // M_t = l (1/6 blockdiag(A0, A0) + 2/3 Qm^T M Qm)
  const double* m = bm.m_cs + 36 * e;
  double qm[6][12], a0[6][6];
  fj_q_mat<double>(0.5, dv, o, qm);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      const int bi = (i / 3) * 3, bj = (j / 3) * 3, ii = i % 3, jj = j % 3;
      double acc = 0.0;
      for (int pp = 0; pp < 3; ++pp)
        for (int q = 0; q < 3; ++q) acc += o[3 * ii + pp] * m[6 * (bi + pp) + bj + q] * o[3 * jj + q];
      a0[i][j] = acc;
    }
  for (int r = 0; r < 12; ++r)
    for (int c = 0; c < 12; ++c) {
      double mq = 0.0;
      for (int i = 0; i < 6; ++i) {
        double x = 0.0;
        for (int j = 0; j < 6; ++j) x += m[6 * i + j] * qm[j][c];
        mq += qm[i][r] * x;
      }
      double f01 = 0.0;
      if (r < 6 && c < 6) f01 = a0[r][c];
      else if (r >= 6 && c >= 6) f01 = a0[r - 6][c - 6];
      mt[144 * e + 12 * r + c] = l * (f01 / 6.0 + (2.0 / 3.0) * mq);
    }
}

extern "C" int fj_beam_element_jacobians(const fj_beam_model* model, int dynamic, const double* hg, const double* v,
                                         const double* v_dot, const double* k_elem, double* fe, double* jd,
                                         double* jv, double* mt, double* p, double* eps, void* stream) {
  if (!model || !hg || !k_elem || !fe || !jd || !p || !eps) return FJ_ERR_ARG;
  if (dynamic && (!v || !v_dot || !jv || !mt)) return FJ_ERR_ARG;
  if (model->n_elem <= 0) return FJ_ERR_ARG;
  const FjBeamDev bm = fj_beam_dev(model);
  const int nt = model->n_elem * (dynamic ? 18 : 6);
  hipLaunchKernelGGL(k_beam_elem_jac, dim3(fj_cdiv(nt, 64)), dim3(64), 0, (hipStream_t)stream, bm, dynamic ? 1 : 0,
                     hg, v, v_dot, k_elem, fe, jd, jv, mt, p, eps);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

extern "C" size_t fj_beam_workspace_bytes(int n_elem) {
  return n_elem <= 0 ? 0 : (size_t)n_elem * FJ_BW_STRIDE * sizeof(double);
}

static int fj_beam_model_ok(const fj_beam_model* m) {
  if (!m || m->n_nodes < 1 || m->n_elem < 0) return 0;
  if (m->n_elem > 0 && (!m->conn || !m->o0 || !m->l0 || !m->d0 || !m->k_cs || !m->m_cs || !m->node_off ||
                        !m->node_elem))
    return 0;
  return 1;
}

static int fj_beam_system_g(const fj_beam_model* model, int dynamic, const double* hg, const double* v,
                            const double* v_dot, const double* t_phi, const double* f_follower,
                            const double* f_dead, double beta_p, double gamma_p, const int32_t* solve_dofs,
                            int n_solve, double* A, double* r, double* f_abs, void* workspace, void* stream,
                            const int32_t* guard) {
  if (!fj_beam_model_ok(model) || n_solve < 0 || n_solve > 6 * model->n_nodes) return FJ_ERR_ARG;
  if (n_solve == 0) return FJ_OK;
  if (!hg || !solve_dofs || !A || !r || !f_abs || (model->n_elem > 0 && !workspace)) return FJ_ERR_ARG;
  if (dynamic && (!v || !v_dot || !t_phi)) return FJ_ERR_ARG;
  if ((long long)n_solve * n_solve + n_solve > (1LL << 31)) return FJ_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const FjBeamDev bm = fj_beam_dev(model);
  double* ws = (double*)workspace;
  if (bm.n_elem > 0) {
    const int pts = bm.n_elem * (dynamic ? 5 : 1);
    if (dynamic) {
      hipLaunchKernelGGL(k_beam_points<1>, dim3(fj_cdiv(pts, 64)), dim3(64), 0, st, bm, hg, v, ws, guard);
      hipLaunchKernelGGL(k_beam_matrices<1>, dim3(bm.n_elem), dim3(FJ_BEAM_MAT_THREADS), 0, st, bm, v, v_dot,
                         t_phi, beta_p, gamma_p, ws, guard);
    } else {
      hipLaunchKernelGGL(k_beam_points<0>, dim3(fj_cdiv(pts, 64)), dim3(64), 0, st, bm, hg, v, ws, guard);
      hipLaunchKernelGGL(k_beam_matrices<0>, dim3(bm.n_elem), dim3(FJ_BEAM_MAT_THREADS), 0, st, bm, v, v_dot,
                         t_phi, beta_p, gamma_p, ws, guard);
    }
    FJ_CHECK_LAUNCH();
  }
  const long long tot = (long long)n_solve * n_solve + n_solve;
  hipLaunchKernelGGL(k_beam_assemble, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, bm, solve_dofs,
                     n_solve, hg, f_follower, f_dead, ws, A, r, f_abs, guard);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

extern "C" int fj_beam_system(const fj_beam_model* model, int dynamic, const double* hg, const double* v,
                              const double* v_dot, const double* t_phi, const double* f_follower,
                              const double* f_dead, double beta_p, double gamma_p, const int32_t* solve_dofs,
                              int n_solve, double* A, double* r, double* f_abs, void* workspace, void* stream) {
  return fj_beam_system_g(model, dynamic, hg, v, v_dot, t_phi, f_follower, f_dead, beta_p, gamma_p, solve_dofs,
                          n_solve, A, r, f_abs, workspace, stream, nullptr);
}

extern "C" int fj_beam_update(int dynamic, int n_nodes, const int32_t* solve_dofs, int n_solve, const double* x,
                              double relax, double gamma_p, double beta_p, double* phi, double* v, double* v_dot,
                              const double* r, const double* f_abs, double* norms, void* stream) {
  if (n_nodes < 1 || n_solve < 0 || n_solve > 6 * n_nodes || !phi || !norms) return FJ_ERR_ARG;
  if (n_solve > 0 && (!solve_dofs || !x || !r || !f_abs)) return FJ_ERR_ARG;
  if (dynamic && (!v || !v_dot)) return FJ_ERR_ARG;
  hipLaunchKernelGGL(k_beam_update, dim3(1), dim3(FJ_BEAM_UPD_THREADS), 0, (hipStream_t)stream, dynamic,
                     6 * n_nodes, solve_dofs, n_solve, x, relax, gamma_p, beta_p, phi, v, v_dot, r, f_abs, norms,
                     nullptr);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

extern "C" int fj_beam_subnorm(const double* x, const int32_t* idx, int n, double* out, void* stream) {
  if (n < 0 || !out || (n > 0 && (!x || !idx))) return FJ_ERR_ARG;
  hipLaunchKernelGGL(k_beam_subnorm, dim3(1), dim3(FJ_BEAM_UPD_THREADS), 0, (hipStream_t)stream, x, idx, n, out);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

extern "C" int fj_beam_nodes(int op, int n_nodes, const double* hg, const double* x, const double* hg_ref,
                             double* hg_out, double* t_out, double* h_out, void* stream) {
  if (n_nodes < 0 || !hg) return FJ_ERR_ARG;
  if (op == FJ_BEAM_NODE_UPDATE) {
    if (!x || !hg_out || (h_out && !hg_ref)) return FJ_ERR_ARG;
  } else if (op == FJ_BEAM_NODE_REL) {
    if (!hg_ref || !h_out) return FJ_ERR_ARG;
  } else if (op == FJ_BEAM_NODE_RATE) {
    if (!x || !hg_out) return FJ_ERR_ARG;
  } else if (op == FJ_BEAM_NODE_LOCAL || op == FJ_BEAM_NODE_GLOBAL) {
    if (!x || !h_out) return FJ_ERR_ARG;
  } else {
    return FJ_ERR_ARG;
  }
  if (n_nodes == 0) return FJ_OK;
  hipLaunchKernelGGL(k_beam_nodes, dim3(fj_cdiv(n_nodes, 64)), dim3(64), 0, (hipStream_t)stream, op, n_nodes,
                     hg, x, hg_ref, hg_out, t_out, h_out, nullptr);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

extern "C" int fj_beam_project_forces(int n_nodes, const double* rmat, int m1, int n1, const double* f_a,
                                      const double* f_b, const double* x0, const int32_t* dof_mapping, int clear,
                                      double* out, void* stream) {
  if (n_nodes < 1 || m1 < 0 || n1 < 0 || !rmat || !out) return FJ_ERR_ARG;
  if (m1 * n1 > 0 && (!f_a || !x0 || !dof_mapping)) return FJ_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (clear) hipLaunchKernelGGL(k_beam_zero, dim3(fj_cdiv(6 * n_nodes, 256)), dim3(256), 0, st, out, 6 * n_nodes);
  if (m1 * n1 > 0)
    hipLaunchKernelGGL(k_beam_project, dim3(fj_cdiv(n1, 64)), dim3(64), 0, st, m1, n1, rmat, f_a, f_b, x0,
                       dof_mapping, out);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---------------------------------------------------------------------------
// Device-driven dynamic Newton loop
// ---------------------------------------------------------------------------
__global__ void k_newton_reset(int32_t* __restrict__ status) {
  if (threadIdx.x < 4) status[threadIdx.x] = 0;
}

// ConvergenceStatus.update and get_status (utils/data_structures.py:81-170)
// on the device: norms = [|dx|, |total disp|, |r|, max |f_abs|, NaN flag];
// a NaN tolerance is an unset one (every comparison with it is false).
__device__ __forceinline__ double fj_ratio(double num, double den) {
  if (den == 0.0) return num != 0.0 ? __builtin_huge_val() : __builtin_nan("");
  return num / den;
}
__device__ __forceinline__ double fj_finite_or_inf(double x) {
  if (x != x) return __builtin_huge_val();
  if (x == __builtin_huge_val()) return 1.7976931348623157e308;
  return x;
}

__global__ void k_newton_status(const double* __restrict__ norms, fj_convergence conv,
                                int32_t* __restrict__ status) {
  if (threadIdx.x != 0 || status[1] != 0) return;
  const int it = status[0] + 1;
  const bool has_nan = norms[4] != 0.0;
  const bool c_ad = norms[0] < conv.abs_disp_tol;
  const bool c_rd = fj_finite_or_inf(fj_ratio(norms[0], norms[1])) < conv.rel_disp_tol;
  const bool c_af = norms[2] < conv.abs_force_tol;
  const bool c_rf = fj_finite_or_inf(fj_ratio(norms[2], norms[3])) < conv.rel_force_tol;
  const bool converged = (c_rd || c_ad || c_rf || c_af) && it > 0;
  const bool final_iter = conv.max_n_iter > 0 && it >= conv.max_n_iter;
  status[0] = it;
  status[2] = converged ? 1 : 0;
  status[3] = has_nan ? 1 : 0;
  status[1] = (converged || has_nan || final_iter) ? 1 : 0;
}

int fj_band_launch(const double* A, int n, int lda, int kl, int ku, const double* b, double* x, int32_t* info,
                   const int32_t* guard, hipStream_t st);

extern "C" int fj_beam_newton_dynamic(const fj_beam_model* model, int n_iter, int reset, const double* hg_prev,
                                      double* phi_a, double* v_a, double* v_dot_a, const double* f_follower,
                                      const double* f_dead, double beta_p, double gamma_p, double relax,
                                      const int32_t* solve_dofs, int n_solve, int kl, int ku,
                                      const fj_convergence* conv, double* hg_u, double* t_phi, double* A,
                                      double* r, double* f_abs, double* x, double* norms, int32_t* status,
                                      int32_t* band_info, void* workspace, void* stream) {
  if (!fj_beam_model_ok(model) || n_iter < 0 || !conv || !status || !hg_prev || !phi_a || !v_a || !v_dot_a ||
      !hg_u || !t_phi || !norms || n_solve < 0 || n_solve > 6 * model->n_nodes)
    return FJ_ERR_ARG;
  if (n_solve > 0 && (!A || !r || !f_abs || !x || !solve_dofs)) return FJ_ERR_ARG;
  if (n_solve > 0 && (kl < 0 || ku < 0 || kl > 63)) return FJ_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (reset) {
    hipLaunchKernelGGL(k_newton_reset, dim3(1), dim3(64), 0, st, status);
    FJ_CHECK_LAUNCH();
  }
  const int nn = model->n_nodes;
  for (int it = 0; it < n_iter; ++it) {
    hipLaunchKernelGGL(k_beam_nodes, dim3(fj_cdiv(nn, 64)), dim3(64), 0, st, FJ_BEAM_NODE_UPDATE, nn, hg_prev,
                       (const double*)phi_a, (const double*)nullptr, hg_u, t_phi, (double*)nullptr,
                       (const int32_t*)status);
    int rc = fj_beam_system_g(model, 1, hg_u, v_a, v_dot_a, t_phi, f_follower, f_dead, beta_p, gamma_p,
                              solve_dofs, n_solve, A, r, f_abs, workspace, stream, status);
    if (rc != FJ_OK) return rc;
    if (n_solve > 0) {
      rc = fj_band_launch(A, n_solve, n_solve, kl, ku, r, x, band_info, status, st);
      if (rc != FJ_OK) return rc;
    }
    hipLaunchKernelGGL(k_beam_update, dim3(1), dim3(FJ_BEAM_UPD_THREADS), 0, st, 1, 6 * nn, solve_dofs, n_solve,
                       (const double*)x, relax, gamma_p, beta_p, phi_a, v_a, v_dot_a, (const double*)r,
                       (const double*)f_abs, norms, (const int32_t*)status);
    hipLaunchKernelGGL(k_newton_status, dim3(1), dim3(64), 0, st, (const double*)norms, *conv, status);
    FJ_CHECK_LAUNCH();
  }
  return FJ_OK;
}

// ---- generalised-alpha bookkeeping of the coupled time step ------------------
// predict_q + calculate_q_alpha (structure/time_integration.py:86-150), the
// alpha-point state of an FSI iteration and calculate_q_n_from_q_alpha +
// calculate_a_n (:53-84), the alpha-point aero force and the FSI convergence
// norms (structure/beam.py:2612-2725): one launch each where the host would
// issue a dozen array operations.  Every product and sum is rounded on its
// own, in the order of the reference's array expressions (no contraction).
#define FJ_GA_THREADS 256

__global__ void k_beam_ga_predict(int n, double am, double af, double g, double b, double dt,
                                  const double* __restrict__ v, const double* __restrict__ vd,
                                  const double* __restrict__ a, double* __restrict__ phi_a,
                                  double* __restrict__ v_a, double* __restrict__ vd_a) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a_n = (af * vd[i] - am * a[i]) / (1.0 - am);
  const double phi_n = dt * v[i] + (dt * dt) * ((0.5 - b) * a[i] + b * a_n);
  const double v_n = v[i] + ((1.0 - g) * dt) * a[i] + (g * dt) * a_n;
  const double vd_n = ((1.0 - am) * a_n + am * a[i] - af * vd[i]) / (1.0 - af);
  phi_a[i] = (1.0 - af) * phi_n;
  v_a[i] = (1.0 - af) * v_n + af * v[i];
  vd_a[i] = (1.0 - af) * vd_n + af * vd[i];
}

// the step-n state from the alpha point: phi_n, v_n (and a copy of phi_a)
__global__ void k_beam_ga_alpha(int n, double af, const double* __restrict__ phi_a,
                                const double* __restrict__ v_a, const double* __restrict__ v_prev,
                                double* __restrict__ phi_n, double* __restrict__ v_n,
                                double* __restrict__ phi_copy) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  phi_n[i] = phi_a[i] / (1.0 - af);
  v_n[i] = (v_a[i] - af * v_prev[i]) / (1.0 - af);
  if (phi_copy) phi_copy[i] = phi_a[i];
}

__global__ void k_beam_ga_finish(int n, double am, double af, const double* __restrict__ phi_a,
                                 const double* __restrict__ v_a, const double* __restrict__ vd_a,
                                 const double* __restrict__ v_prev, const double* __restrict__ vd_prev,
                                 const double* __restrict__ a_prev, double* __restrict__ phi,
                                 double* __restrict__ v, double* __restrict__ vd, double* __restrict__ a) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  phi[i] = phi_a[i] / (1.0 - af);
  const double vi = (v_a[i] - af * v_prev[i]) / (1.0 - af);
  const double vdi = (vd_a[i] - af * vd_prev[i]) / (1.0 - af);
  v[i] = vi;
  vd[i] = vdi;
  a[i] = (1.0 / (1.0 - am)) * ((1.0 - af) * vdi + af * vd_prev[i] - am * a_prev[i]);
}

__global__ void k_beam_fsi_force(int n, double af, const double* __restrict__ f_nm1,
                                 const double* __restrict__ f_n, double* __restrict__ f_al) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f_al[i] = (1.0 - af) * f_n[i] + af * f_nm1[i];
}

// out = {|dphi|, |phi_a|, |df|, max |f_al|, any NaN in dphi} over the solve
// DOFs, dphi = phi_old - phi_a, df = f_al - f_prev (one workgroup, a fixed
// summation order)
__global__ __launch_bounds__(FJ_GA_THREADS) void k_beam_fsi_norms(
    int n_solve, const int32_t* __restrict__ sd, const double* __restrict__ phi_old,
    const double* __restrict__ phi_a, const double* __restrict__ f_al, const double* __restrict__ f_prev,
    double* __restrict__ out) {
  __shared__ double red[4][FJ_GA_THREADS];
  __shared__ int nan_any;
  const int tid = threadIdx.x;
  if (tid == 0) nan_any = 0;
  double s_dd = 0.0, s_phi = 0.0, s_df = 0.0, mx = 0.0;
  int nan_here = 0;
  for (int k = tid; k < n_solve; k += FJ_GA_THREADS) {
    const int j = sd[k];
    const double dd = phi_old[j] - phi_a[j];
    const double df = f_al[j] - f_prev[j];
    s_dd += dd * dd;
    s_phi += phi_a[j] * phi_a[j];
    s_df += df * df;
    mx = fmax(mx, fabs(f_al[j]));
    nan_here |= (dd != dd) ? 1 : 0;
  }
  __syncthreads();
  if (nan_here) atomicOr(&nan_any, 1);
  red[0][tid] = s_dd;
  red[1][tid] = s_phi;
  red[2][tid] = s_df;
  red[3][tid] = mx;
  __syncthreads();
  for (int s = FJ_GA_THREADS / 2; s > 0; s >>= 1) {
    if (tid < s) {
      red[0][tid] += red[0][tid + s];
      red[1][tid] += red[1][tid + s];
      red[2][tid] += red[2][tid + s];
      red[3][tid] = fmax(red[3][tid], red[3][tid + s]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    out[0] = sqrt(red[0][0]);
    out[1] = sqrt(red[1][0]);
    out[2] = sqrt(red[2][0]);
    out[3] = red[3][0];
    out[4] = nan_any ? 1.0 : 0.0;
  }
}

// The per-DOF arithmetic of k_beam_ga_alpha and k_beam_ga_finish (each
// product and sum rounded on its own), for the node-level fused kernel.
__device__ __forceinline__ void fj_ga_alpha_dof(double af, double pa, double va, double vp, double& pn,
                                                double& vn) {
#pragma clang fp contract(off)
  pn = pa / (1.0 - af);
  vn = (va - af * vp) / (1.0 - af);
}

__device__ __forceinline__ void fj_ga_finish_dof(double am, double af, double pa, double va, double vda,
                                                 double vp, double vdp, double ap, double& p, double& v,
                                                 double& vd, double& a) {
#pragma clang fp contract(off)
  p = pa / (1.0 - af);
  v = (va - af * vp) / (1.0 - af);
  vd = (vda - af * vdp) / (1.0 - af);
  a = (1.0 / (1.0 - am)) * ((1.0 - af) * vd + af * vdp - am * ap);
}

// One thread per node: the generalised-alpha state of the node's six DOFs
// (as k_beam_ga_alpha / k_beam_ga_finish), then its frame hg_prev exp(phi)
// and, for FJ_GA_ALPHA, the frame rate hg tilde(v) (as k_beam_nodes), in
// one launch instead of three.
__global__ void k_beam_ga_nodes(int op, int nn, double am, double af, const double* __restrict__ phi_a,
                                const double* __restrict__ v_a, const double* __restrict__ vd_a,
                                const double* __restrict__ v_prev, const double* __restrict__ vd_prev,
                                const double* __restrict__ a_prev, const double* __restrict__ hg_prev,
                                double* __restrict__ phi, double* __restrict__ v, double* __restrict__ vd,
                                double* __restrict__ a, double* __restrict__ phi_copy, double* __restrict__ hg,
                                double* __restrict__ hg_dot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  double ph[6], vv[6];
  for (int d = 0; d < 6; ++d) {
    const int k = 6 * i + d;
    if (op == FJ_GA_ALPHA) {
      fj_ga_alpha_dof(af, phi_a[k], v_a[k], v_prev[k], ph[d], vv[d]);
      if (phi_copy) phi_copy[k] = phi_a[k];
    } else {
      double vdk, ak;
      fj_ga_finish_dof(am, af, phi_a[k], v_a[k], vd_a[k], v_prev[k], vd_prev[k], a_prev[k], ph[d], vv[d], vdk, ak);
      vd[k] = vdk;
      a[k] = ak;
    }
    phi[k] = ph[d];
    v[k] = vv[d];
  }
  double ex[16], gn[16];
  fj_exp_se3(ph, ex);
  fj_hg_mul(hg_prev + 16 * i, ex, gn);
  for (int q = 0; q < 16; ++q) hg[16 * i + q] = gn[q];
  if (hg_dot) fj_hg_rate(gn, vv, hg_dot + 16 * i);
}

extern "C" int fj_beam_ga(int op, int n, double am, double af, double g, double b, double dt,
                          const double* p0, const double* p1, const double* p2, const double* p3,
                          const double* p4, const double* p5, double* o0, double* o1, double* o2,
                          double* o3, void* stream) {
  if (n < 0) return FJ_ERR_ARG;
  if (n == 0) return FJ_OK;
  const dim3 grid(fj_cdiv(n, FJ_GA_THREADS)), block(FJ_GA_THREADS);
  hipStream_t st = (hipStream_t)stream;
  switch (op) {
    case FJ_GA_PREDICT:  // p0 v, p1 v_dot, p2 a -> o0 phi_a, o1 v_a, o2 v_dot_a
      if (!p0 || !p1 || !p2 || !o0 || !o1 || !o2) return FJ_ERR_ARG;
      hipLaunchKernelGGL(k_beam_ga_predict, grid, block, 0, st, n, am, af, g, b, dt, p0, p1, p2, o0, o1, o2);
      break;
    case FJ_GA_ALPHA:  // p0 phi_a, p1 v_a, p2 v_prev -> o0 phi_n, o1 v_n, o2 copy of phi_a (optional)
      if (!p0 || !p1 || !p2 || !o0 || !o1) return FJ_ERR_ARG;
      hipLaunchKernelGGL(k_beam_ga_alpha, grid, block, 0, st, n, af, p0, p1, p2, o0, o1, o2);
      break;
    case FJ_GA_FINISH:  // p0 phi_a, p1 v_a, p2 v_dot_a, p3 v_prev, p4 v_dot_prev, p5 a_prev -> o0 phi, o1 v, o2 v_dot, o3 a
      if (!p0 || !p1 || !p2 || !p3 || !p4 || !p5 || !o0 || !o1 || !o2 || !o3) return FJ_ERR_ARG;
      hipLaunchKernelGGL(k_beam_ga_finish, grid, block, 0, st, n, am, af, p0, p1, p2, p3, p4, p5, o0, o1, o2, o3);
      break;
    case FJ_GA_FORCE:  // p0 f_nm1, p1 f_n -> o0 f_alpha
      if (!p0 || !p1 || !o0) return FJ_ERR_ARG;
      hipLaunchKernelGGL(k_beam_fsi_force, grid, block, 0, st, n, af, p0, p1, o0);
      break;
    default:
      return FJ_ERR_ARG;
  }
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

extern "C" int fj_beam_ga_nodes(int op, int n_nodes, double am, double af, const double* phi_a, const double* v_a,
                                const double* v_dot_a, const double* v_prev, const double* v_dot_prev,
                                const double* a_prev, const double* hg_prev, double* phi, double* v,
                                double* v_dot, double* a, double* phi_copy, double* hg, double* hg_dot,
                                void* stream) {
  if (n_nodes < 0) return FJ_ERR_ARG;
  if (!phi_a || !v_a || !v_prev || !hg_prev || !phi || !v || !hg) return FJ_ERR_ARG;
  if (op == FJ_GA_FINISH) {
    if (!v_dot_a || !v_dot_prev || !a_prev || !v_dot || !a || hg_dot || phi_copy) return FJ_ERR_ARG;
  } else if (op != FJ_GA_ALPHA) {
    return FJ_ERR_ARG;
  }
  if (n_nodes == 0) return FJ_OK;
  hipLaunchKernelGGL(k_beam_ga_nodes, dim3(fj_cdiv(n_nodes, 64)), dim3(64), 0, (hipStream_t)stream, op, n_nodes, am,
                     af, phi_a, v_a, v_dot_a, v_prev, v_dot_prev, a_prev, hg_prev, phi, v, v_dot, a, phi_copy, hg,
                     hg_dot);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

extern "C" int fj_beam_fsi_norms(int n_solve, const int32_t* solve_dofs, const double* phi_old,
                                 const double* phi_a, const double* f_al, const double* f_prev, double* out,
                                 void* stream) {
  if (n_solve < 0 || !out || (n_solve > 0 && (!solve_dofs || !phi_old || !phi_a || !f_al || !f_prev)))
    return FJ_ERR_ARG;
  hipLaunchKernelGGL(k_beam_fsi_norms, dim3(1), dim3(FJ_GA_THREADS), 0, (hipStream_t)stream, n_solve, solve_dofs,
                     phi_old, phi_a, f_al, f_prev, out);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

// ---- the end of a coupled time step in one launch ----------------------------
// calculate_q_n_from_q_alpha + calculate_a_n and the node frames (as
// k_beam_ga_nodes, FJ_GA_FINISH), the aerodynamic vertex forces of one
// surface projected onto the new frames (as k_beam_zero + k_beam_project),
// their local-frame copy (as k_beam_nodes, FJ_BEAM_NODE_LOCAL) and the
// predictor of the next step (as k_beam_ga_predict): one workgroup, the
// phases separated by barriers, the frames' rotations and the projected
// forces kept in LDS.  Every value is the same expression as in the
// separate kernels, so the results are the same bits.
#define FJ_STEP_END_MAX_NODES 512
#define FJ_STEP_END_THREADS 256

__device__ __forceinline__ void fj_ga_predict_dof(double am, double af, double g, double b, double dt, double v,
                                                  double vd, double a, double& phi_a, double& v_a, double& vd_a) {
#pragma clang fp contract(off)
  const double a_n = (af * vd - am * a) / (1.0 - am);
  const double phi_n = dt * v + (dt * dt) * ((0.5 - b) * a + b * a_n);
  const double v_n = v + ((1.0 - g) * dt) * a + (g * dt) * a_n;
  const double vd_n = ((1.0 - am) * a_n + am * a - af * vd) / (1.0 - af);
  phi_a = (1.0 - af) * phi_n;
  v_a = (1.0 - af) * v_n + af * v;
  vd_a = (1.0 - af) * vd_n + af * vd;
}

__global__ __launch_bounds__(FJ_STEP_END_THREADS) void k_beam_step_end(
    int nn, double am, double af, double g, double b, double dt, const double* __restrict__ phi_a,
    const double* __restrict__ v_a, const double* __restrict__ vd_a, const double* __restrict__ v_prev,
    const double* __restrict__ vd_prev, const double* __restrict__ a_prev, const double* __restrict__ hg_prev,
    int m1, int n1, const double* __restrict__ fa, const double* __restrict__ fb, const double* __restrict__ x0,
    const int32_t* __restrict__ dof, double* __restrict__ phi, double* __restrict__ v, double* __restrict__ vd,
    double* __restrict__ a, double* __restrict__ hg, double* __restrict__ f_glob, double* __restrict__ f_loc,
    double* __restrict__ phi_a_next, double* __restrict__ v_a_next, double* __restrict__ vd_a_next) {
  __shared__ double rot[FJ_STEP_END_MAX_NODES][9];
  __shared__ double fg[FJ_STEP_END_MAX_NODES][6];
  const int tid = threadIdx.x;
  for (int i = tid; i < nn; i += FJ_STEP_END_THREADS) {
    double ph[6];
    for (int d = 0; d < 6; ++d) {
      const int k = 6 * i + d;
      double pk, vk, vdk, ak;
      fj_ga_finish_dof(am, af, phi_a[k], v_a[k], vd_a[k], v_prev[k], vd_prev[k], a_prev[k], pk, vk, vdk, ak);
      phi[k] = pk;
      v[k] = vk;
      vd[k] = vdk;
      a[k] = ak;
      ph[d] = pk;
      fg[i][d] = 0.0;
    }
    double ex[16], gn[16];
    fj_exp_se3(ph, ex);
    fj_hg_mul(hg_prev + 16 * i, ex, gn);
    for (int q = 0; q < 16; ++q) hg[16 * i + q] = gn[q];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) rot[i][3 * r + c] = gn[4 * r + c];
  }
  __syncthreads();
  for (int j = tid; j < n1; j += FJ_STEP_END_THREADS) {
    const int node = dof[j];
    if (node < 0 || node >= nn) continue;
    const double* gr = rot[node];
    double f[3] = {0.0, 0.0, 0.0}, mo[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < m1; ++i) {
      const size_t o = 3 * ((size_t)i * n1 + j);
      double fx = fa[o], fy = fa[o + 1], fz = fa[o + 2];
      if (fb) { fx += fb[o]; fy += fb[o + 1]; fz += fb[o + 2]; }
      const double ax = gr[0] * x0[o] + gr[1] * x0[o + 1] + gr[2] * x0[o + 2];
      const double ay = gr[3] * x0[o] + gr[4] * x0[o + 1] + gr[5] * x0[o + 2];
      const double az = gr[6] * x0[o] + gr[7] * x0[o + 1] + gr[8] * x0[o + 2];
      f[0] += fx; f[1] += fy; f[2] += fz;
      mo[0] += ay * fz - az * fy;
      mo[1] += az * fx - ax * fz;
      mo[2] += ax * fy - ay * fx;
    }
    for (int k = 0; k < 3; ++k) {
      fg[node][k] = f[k];
      fg[node][3 + k] = mo[k];
    }
  }
  __syncthreads();
  for (int i = tid; i < nn; i += FJ_STEP_END_THREADS) {
    const double* gg = hg + 16 * i;  // written by this thread above
    const double* f = fg[i];
    for (int k = 0; k < 6; ++k) f_glob[6 * i + k] = f[k];
    for (int h = 0; h < 6; h += 3)
      for (int p = 0; p < 3; ++p) {
        const double r0 = gg[p], r1 = gg[4 + p], r2 = gg[8 + p];
        f_loc[6 * i + h + p] = r0 * f[h] + r1 * f[h + 1] + r2 * f[h + 2];
      }
    if (phi_a_next)
      for (int d = 0; d < 6; ++d) {
        const int k = 6 * i + d;
        fj_ga_predict_dof(am, af, g, b, dt, v[k], vd[k], a[k], phi_a_next[k], v_a_next[k], vd_a_next[k]);
      }
  }
}

extern "C" int fj_beam_step_end(int n_nodes, double am, double af, double g, double b, double dt,
                                const double* phi_a, const double* v_a, const double* v_dot_a, const double* v_prev,
                                const double* v_dot_prev, const double* a_prev, const double* hg_prev, int m1,
                                int n1, const double* f_a, const double* f_b, const double* x0, const int32_t* dof,
                                double* phi, double* v, double* v_dot, double* a, double* hg, double* f_global,
                                double* f_local, double* phi_a_next, double* v_a_next, double* v_dot_a_next,
                                void* stream) {
  if (n_nodes <= 0 || n_nodes > FJ_STEP_END_MAX_NODES || m1 < 1 || n1 < 1) return FJ_ERR_ARG;
  if (!phi_a || !v_a || !v_dot_a || !v_prev || !v_dot_prev || !a_prev || !hg_prev || !f_a || !x0 || !dof) return FJ_ERR_ARG;
  if (!phi || !v || !v_dot || !a || !hg || !f_global || !f_local) return FJ_ERR_ARG;
  if ((phi_a_next || v_a_next || v_dot_a_next) && !(phi_a_next && v_a_next && v_dot_a_next)) return FJ_ERR_ARG;
  hipLaunchKernelGGL(k_beam_step_end, dim3(1), dim3(FJ_STEP_END_THREADS), 0, (hipStream_t)stream, n_nodes, am, af,
                     g, b, dt, phi_a, v_a, v_dot_a, v_prev, v_dot_prev, a_prev, hg_prev, m1, n1, f_a, f_b, x0, dof,
                     phi, v, v_dot, a, hg, f_global, f_local, phi_a_next, v_a_next, v_dot_a_next);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}
