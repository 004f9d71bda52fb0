// K2: AIC assembly (compute_aic_grid / compute_aic_sys / add_wake_influence,
// aero/aic.py:11-205).
//
// One thread per (collocation point, ring).  The ring's four filaments are
// evaluated directly (the shared filament of two neighbouring rings is
// evaluated twice: the bound AIC is <1% of a step's pair count, and this
// keeps the reference's difference-of-projections rounding order,
// A = -(M[i,j+1] - M[i,j]) + (N[i+1,j] - N[i,j]), aero/aic.py:55-57).
// grid.x runs over rings so the row-major AIC rows are written coalesced.
#include "fj_common.hpp"
#include "fj_aic_dev.hpp"

template <int KIND>
__global__ void k_aic(const double* __restrict__ colloc, const double* __restrict__ normals,
                      int n_tgt, const double* __restrict__ zeta, int m, int n,
                      const double* __restrict__ zeta_w, int m_star, int mirror, double mpx,
                      double mpy, double mpz, double mnx, double mny, double mnz,
                      FjKernelParams kp, double* __restrict__ A, int lda, int col0,
                      int accumulate) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (r >= m * n || t >= n_tgt) return;
  const int i = r / n, j = r % n;
  const double x = colloc[3 * t], y = colloc[3 * t + 1], z = colloc[3 * t + 2];
  const double nx = normals[3 * t], ny = normals[3 * t + 1], nz = normals[3 * t + 2];
  const double mp[3] = {mpx, mpy, mpz}, mn[3] = {mnx, mny, mnz};
  double a = fj_ring<KIND>(x, y, z, nx, ny, nz, zeta, n, i, j, 0, mp, mn, kp);
  if (mirror) a -= fj_ring<KIND>(x, y, z, nx, ny, nz, zeta, n, i, j, 1, mp, mn, kp);
  if (zeta_w && i == m - 1) {
    // add_wake_influence: every wake ring of column j lumped onto the TE ring
    double aw = 0.0;
    for (int iw = 0; iw < m_star; ++iw) {
      double w = fj_ring<KIND>(x, y, z, nx, ny, nz, zeta_w, n, iw, j, 0, mp, mn, kp);
      if (mirror) w -= fj_ring<KIND>(x, y, z, nx, ny, nz, zeta_w, n, iw, j, 1, mp, mn, kp);
      aw += w;
    }
    a += aw;
  }
  double* o = A + (size_t)t * lda + col0 + r;
  if (accumulate) *o += a; else *o = a;
}

extern "C" int fj_aic_assemble(const double* colloc, const double* normals, int n_tgt,
                               const double* zeta, int m, int n, const double* zeta_w,
                               int m_star, const double* mirror_point,
                               const double* mirror_normal, int kernel, double eps, double rcut,
                               double* A_out, int lda, int col0, int accumulate, void* stream) {
  if (n_tgt < 0 || m < 0 || n < 0 || m_star < 0 || col0 < 0) return FJ_ERR_ARG;
  if (kernel < FJ_KERNEL_EPSILON || kernel > FJ_KERNEL_PLAIN) return FJ_ERR_ARG;
  if (n_tgt == 0 || m * n == 0) return FJ_OK;
  if (!colloc || !normals || !zeta || !A_out || lda < col0 + m * n) return FJ_ERR_ARG;
  if (n_tgt > 65535) return FJ_ERR_ARG;
  const int mirror = (mirror_point && mirror_normal) ? 1 : 0;
  const double z3[3] = {0, 0, 0};
  const double* mp = mirror ? mirror_point : z3;
  const double* mn = mirror ? mirror_normal : z3;
  FjKernelParams kp;
  kp.eps = eps; kp.eps2 = eps * eps; kp.rcut = rcut;
  const double* zw = (zeta_w && m_star > 0) ? zeta_w : nullptr;
  const dim3 grid(fj_cdiv(m * n, 64), n_tgt);
  hipStream_t st = (hipStream_t)stream;
#define FJ_LAUNCH_AIC(K)                                                                   \
  hipLaunchKernelGGL(k_aic<K>, grid, dim3(64), 0, st, colloc, normals, n_tgt, zeta, m, n, zw, \
                     m_star, mirror, mp[0], mp[1], mp[2], mn[0], mn[1], mn[2], kp, A_out, lda, \
                     col0, accumulate)
  if (kernel == FJ_KERNEL_EPSILON) FJ_LAUNCH_AIC(FJ_KERNEL_EPSILON);
  else if (kernel == FJ_KERNEL_CUTOFF) FJ_LAUNCH_AIC(FJ_KERNEL_CUTOFF);
  else FJ_LAUNCH_AIC(FJ_KERNEL_PLAIN);
#undef FJ_LAUNCH_AIC
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}
