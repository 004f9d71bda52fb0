// K3: dense circulation solve, replacing jnp.linalg.solve (aero/uvlm.py:985).
//
// Blocked right-looking LU with partial pivoting on a row-major matrix.
//  * panel kernel (one 1024-thread workgroup): the (n-k) x jb panel is held in
//    registers, each thread owning up to R rows; per column: register argmax,
//    wave shuffles + LDS for the workgroup argmax, the pivot row broadcast
//    through LDS, row exchange between the two owner threads, scale and
//    rank-1 update in registers.  The composed row permutation is kept in
//    ipiv[n .. 2n) so the solves gather Pb in one parallel step.
//  * swap/trsm kernel, one thread per column outside the panel: the panel's
//    row exchanges (laswp) and, right of the panel, the unit-lower triangular
//    solve giving U12 (trsm) with L11 staged in LDS.
//  * trailing update A22 -= L21 U12 on the matrix cores
//    (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32), 64x64 tiles per
//    workgroup, K = jb staged in LDS.
// The triangular solves run in one workgroup: diagonal 64x64 blocks solved by
// one wave in LDS, the off-diagonal GEMV by all waves (one row per wave, lanes
// over the block columns, coalesced), x kept in LDS.
#include "fj_common.hpp"

#define FJ_LU_THREADS 1024
#define FJ_LU_WAVES (FJ_LU_THREADS / 64)

template <typename T>
__device__ __forceinline__ T fj_abs(T x) { return x < T(0) ? -x : x; }

// workgroup argmax of (value, index), smallest index among equal values
__device__ __forceinline__ void fj_argmax_pair(double& v, int& idx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off);
    const int oi = __shfl_xor(idx, off);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
}

template <typename T, int R, int JB>
__global__ __launch_bounds__(FJ_LU_THREADS) void k_lu_panel(T* __restrict__ A, int n, int lda,
                                                             int k, int32_t* __restrict__ ipiv,
                                                             int32_t* __restrict__ info) {
  __shared__ double s_val[FJ_LU_WAVES];
  __shared__ int s_idx[FJ_LU_WAVES];
  __shared__ T s_piv[2][JB];
  __shared__ T s_swp[2][JB];
  __shared__ int s_ipiv[JB];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int jb = min(JB, n - k);
  const int rows = n - k;

  // load the panel: thread tid owns local rows tid + r * 1024
  T a[R][JB];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int lr = tid + r * FJ_LU_THREADS;
#pragma unroll
    for (int c = 0; c < JB; ++c)
      a[r][c] = (lr < rows && c < jb) ? A[(size_t)(k + lr) * lda + k + c] : T(0);
  }

  // fully unrolled so that every register index below is a compile-time constant
#pragma unroll
  for (int jj = 0; jj < JB; ++jj) {
    if (jj >= jb) break;  // uniform: jb is the same for the whole workgroup
    const int buf = jj & 1;
    // argmax |a[lr][jj]| over local rows lr >= jj
    double best = -1.0;
    int bidx = 0x7fffffff;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int lr = tid + r * FJ_LU_THREADS;
      if (lr >= jj && lr < rows) {
        const double v = (double)fj_abs(a[r][jj]);
        if (v > best) { best = v; bidx = lr; }
      }
    }
    fj_argmax_pair(best, bidx);
    if (lane == 0) { s_val[wave] = best; s_idx[wave] = bidx; }
    __syncthreads();
    best = s_val[0]; bidx = s_idx[0];
#pragma unroll
    for (int w = 1; w < FJ_LU_WAVES; ++w) {
      const double v = s_val[w];
      const int i = s_idx[w];
      if (v > best || (v == best && i < bidx)) { best = v; bidx = i; }
    }
    const int p = bidx;  // local pivot row
    // owners publish the pivot row and the current row jj
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int lr = tid + r * FJ_LU_THREADS;
      if (lr == p) {
#pragma unroll
        for (int c = 0; c < JB; ++c) s_piv[buf][c] = a[r][c];
      }
      if (lr == jj) {
#pragma unroll
        for (int c = 0; c < JB; ++c) s_swp[buf][c] = a[r][c];
      }
    }
    if (tid == 0) {
      s_ipiv[jj] = k + p;
      if (best == 0.0 && *info == 0) *info = k + jj + 1;
    }
    __syncthreads();
    const T piv = s_piv[buf][jj];
    const T inv = (piv != T(0)) ? T(1) / piv : T(0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int lr = tid + r * FJ_LU_THREADS;
      if (lr == jj) {
#pragma unroll
        for (int c = 0; c < JB; ++c) a[r][c] = s_piv[buf][c];
      } else if (lr == p) {
#pragma unroll
        for (int c = 0; c < JB; ++c) a[r][c] = s_swp[buf][c];
      }
      if (lr > jj && lr < rows) {
        const T l = a[r][jj] * inv;
        a[r][jj] = l;
#pragma unroll
        for (int c = 0; c < JB; ++c)
          if (c > jj) a[r][c] -= l * s_piv[buf][c];
      }
    }
    // the next column writes the other LDS buffer; one barrier per column
  }
  __syncthreads();

  // write the panel back
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int lr = tid + r * FJ_LU_THREADS;
    if (lr < rows) {
#pragma unroll
      for (int c = 0; c < JB; ++c)
        if (c < jb) A[(size_t)(k + lr) * lda + k + c] = a[r][c];
    }
  }
  if (tid < jb) ipiv[k + tid] = s_ipiv[tid];
  // composed permutation ipiv[n + q]: only positions k..k+jb-1 and the pivot
  // rows change; each is traced back through this panel's exchanges (last
  // exchange first) to the position its content came from.
  {
    int32_t* perm = ipiv + n;
    int q = -1, src = -1, val = 0;
    if (tid < jb) q = k + tid;
    else if (tid < 2 * jb) q = s_ipiv[tid - jb];
    if (q >= 0) {
      src = q;
      for (int jj = jb - 1; jj >= 0; --jj) {
        const int a0 = k + jj, b0 = s_ipiv[jj];
        if (src == a0) src = b0;
        else if (src == b0) src = a0;
      }
      val = perm[src];
    }
    __syncthreads();
    if (q >= 0) perm[q] = val;
  }
}

// Columns outside the panel, one thread each: the panel's row exchanges
// (laswp), then U12 = L11^{-1} A12 for the columns right of the panel (trsm).
template <typename T, int JB>
__global__ __launch_bounds__(256) void k_lu_swap_trsm(T* __restrict__ A, int n, int lda, int k,
                                                      int jb, const int32_t* __restrict__ ipiv) {
  __shared__ T s_L11[JB][JB + 1];
  __shared__ T s_u[JB][256];
  __shared__ int s_ipiv[JB];
  const int tid = threadIdx.x;
  for (int e = tid; e < JB * JB; e += blockDim.x) {
    const int r = e / JB, q = e % JB;
    s_L11[r][q] = (r < jb && q < jb) ? A[(size_t)(k + r) * lda + k + q] : T(0);
  }
  if (tid < JB) s_ipiv[tid] = (tid < jb) ? ipiv[k + tid] : k + tid;
  __syncthreads();
  int c = blockIdx.x * blockDim.x + tid;
  if (c >= n - jb) return;
  if (c >= k) c += jb;  // skip the panel columns
  for (int jj = 0; jj < jb; ++jj) {
    const int p = s_ipiv[jj];
    if (p != k + jj) {
      const T t0 = A[(size_t)(k + jj) * lda + c];
      A[(size_t)(k + jj) * lda + c] = A[(size_t)p * lda + c];
      A[(size_t)p * lda + c] = t0;
    }
  }
  if (c >= k + jb) {
    // forward substitution with the column staged in this thread's LDS slot
    for (int r = 0; r < jb; ++r) s_u[r][tid] = A[(size_t)(k + r) * lda + c];
    for (int r = 1; r < jb; ++r) {
      T acc = s_u[r][tid];
      for (int q = 0; q < r; ++q) acc -= s_L11[r][q] * s_u[q][tid];
      s_u[r][tid] = acc;
      A[(size_t)(k + r) * lda + c] = acc;
    }
  }
}

__global__ void k_perm_init(int32_t* __restrict__ ipiv, int n, int32_t* __restrict__ info) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ipiv[n + i] = i;
  if (i == 0) *info = 0;
}

// ---- trailing update on MFMA ----------------------------------------------
// C (rows r0.., cols c0..) -= L (rows, K) * U (K, cols), K = kb <= 32.
template <typename T>
struct FjMfma;

// Operands of both 16x16x4 instructions: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16].  Output element v of lane l is column l % 16 and row
//   FP64: 4 v + l / 16   (four one-row groups)
//   FP32: 4 (l / 16) + v (one four-row group)
template <>
struct FjMfma<double> {
  typedef double __attribute__((ext_vector_type(4))) acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return 4 * v + lk; }
};
template <>
struct FjMfma<float> {
  typedef float __attribute__((ext_vector_type(4))) acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};

#define FJ_GT 64  // C tile edge
#define FJ_GK 32  // max K

template <typename T>
__global__ __launch_bounds__(256) void k_lu_update(T* __restrict__ A, int n, int lda, int k,
                                                   int kb) {
  __shared__ T sL[FJ_GT][FJ_GK + 1];
  __shared__ T sU[FJ_GK][FJ_GT + 1];
  const int base = k + kb;
  const int r0 = base + blockIdx.y * FJ_GT;
  const int c0 = base + blockIdx.x * FJ_GT;
  const int tid = threadIdx.x;
  for (int e = tid; e < FJ_GT * FJ_GK; e += 256) {
    const int i = e / FJ_GK, q = e % FJ_GK;
    sL[i][q] = (r0 + i < n && q < kb) ? A[(size_t)(r0 + i) * lda + k + q] : T(0);
  }
  for (int e = tid; e < FJ_GK * FJ_GT; e += 256) {
    const int q = e / FJ_GT, j = e % FJ_GT;
    sU[q][j] = (c0 + j < n && q < kb) ? A[(size_t)(k + q) * lda + c0 + j] : T(0);
  }
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  typedef typename FjMfma<T>::acc_t acc_t;
  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
  const int li = lane & 15, lk = lane >> 4;
  for (int q = 0; q < kb; q += 4) {
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const T av = sL[wr + a * 16 + li][q + lk];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const T bv = sU[q + lk][wc + b * 16 + li];
        acc[a][b] = FjMfma<T>::mma(av, bv, acc[a][b]);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = r0 + wr + a * 16 + FjMfma<T>::row(lk, v);
        const int j = c0 + wc + b * 16 + li;
        if (i < n && j < n) A[(size_t)i * lda + j] -= acc[a][b][v];
      }
}

template <typename T, int R, int JB>
static int fj_lu_factor_t(T* A, int n, int lda, int32_t* ipiv, int32_t* info, hipStream_t st) {
  hipLaunchKernelGGL(k_perm_init, dim3(fj_cdiv(n, 256)), dim3(256), 0, st, ipiv, n, info);
  FJ_CHECK_LAUNCH();
  for (int k = 0; k < n; k += JB) {
    const int jb = (n - k < JB) ? n - k : JB;
    hipLaunchKernelGGL((k_lu_panel<T, R, JB>), dim3(1), dim3(FJ_LU_THREADS), 0, st, A, n, lda, k,
                       ipiv, info);
    FJ_CHECK_LAUNCH();
    if (n - jb > 0) {
      hipLaunchKernelGGL((k_lu_swap_trsm<T, JB>), dim3(fj_cdiv(n - jb, 256)), dim3(256), 0, st, A, n,
                         lda, k, jb, ipiv);
      FJ_CHECK_LAUNCH();
    }
    const int rest = n - k - jb;
    if (rest > 0) {
      const dim3 g(fj_cdiv(rest, FJ_GT), fj_cdiv(rest, FJ_GT));
      hipLaunchKernelGGL(k_lu_update<T>, g, dim3(256), 0, st, A, n, lda, k, jb);
      FJ_CHECK_LAUNCH();
    }
  }
  return FJ_OK;
}

template <typename T>
static int fj_lu_dispatch(T* A, int n, int lda, int32_t* ipiv, int32_t* info, hipStream_t st) {
  if (n <= 0 || lda < n || !A || !ipiv || !info) return FJ_ERR_ARG;
  if (n <= 1024) return fj_lu_factor_t<T, 1, 32>(A, n, lda, ipiv, info, st);
  if (n <= 2048) return fj_lu_factor_t<T, 2, 16>(A, n, lda, ipiv, info, st);
  if (n <= 4096) return fj_lu_factor_t<T, 4, 16>(A, n, lda, ipiv, info, st);
  if (n <= 8192) return fj_lu_factor_t<T, 8, 4>(A, n, lda, ipiv, info, st);
  return FJ_ERR_ARG;
}

extern "C" size_t fj_lu_workspace_bytes(int n) {
  (void)n;
  return 256;
}

extern "C" int fj_lu_factor(double* A, int n, int lda, int32_t* ipiv, int32_t* info,
                            void* workspace, void* stream) {
  (void)workspace;
  return fj_lu_dispatch<double>(A, n, lda, ipiv, info, (hipStream_t)stream);
}

extern "C" int fj_lu_factor_f32(float* A, int n, int lda, int32_t* ipiv, int32_t* info,
                                void* workspace, void* stream) {
  (void)workspace;
  return fj_lu_dispatch<float>(A, n, lda, ipiv, info, (hipStream_t)stream);
}

// ---- triangular solves -------------------------------------------------------
#define FJ_TS_B 64

// x (LDS, n values) : solve with L (unit lower, trans=0) then U, or with U^T
// then L^T (trans=1).  Element accessors read the row-major factor.
template <typename T, typename X>
__device__ void fj_tri_forward(const T* __restrict__ LU, int n, int lda, X* x, bool unit,
                               bool transposed) {
  // forward substitution with a lower factor M: M = L (trans=0) or U^T (trans=1)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ X diag[FJ_TS_B][FJ_TS_B + 1];
  for (int b0 = 0; b0 < n; b0 += FJ_TS_B) {
    const int bs = min(FJ_TS_B, n - b0);
    // stage the diagonal block M[b0+i][b0+q]
    for (int e = tid; e < FJ_TS_B * FJ_TS_B; e += blockDim.x) {
      const int i = e / FJ_TS_B, q = e % FJ_TS_B;
      X v = 0;
      if (i < bs && q < bs)
        v = transposed ? (X)LU[(size_t)(b0 + q) * lda + b0 + i] : (X)LU[(size_t)(b0 + i) * lda + b0 + q];
      diag[i][q] = v;
    }
    __syncthreads();
    if (wave == 0) {
      X xi = (lane < bs) ? x[b0 + lane] : X(0);
      for (int r = 0; r < bs; ++r) {
        X xr = __shfl(xi, r);
        if (!unit) {
          xr = xr / diag[r][r];
          if (lane == r) xi = xr;
        }
        if (lane > r && lane < bs) xi -= diag[lane][r] * xr;
      }
      if (lane < bs) x[b0 + lane] = xi;
    }
    __syncthreads();
    // rows below the block: x[i] -= sum_q M[i][b0+q] x[b0+q]
    for (int i = b0 + bs + wave; i < n; i += blockDim.x / 64) {
      X s = 0;
      if (lane < bs) {
        const X mv = transposed ? (X)LU[(size_t)(b0 + lane) * lda + i] : (X)LU[(size_t)i * lda + b0 + lane];
        s = mv * x[b0 + lane];
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if (lane == 0) x[i] -= s;
    }
    __syncthreads();
  }
}

template <typename T, typename X>
__device__ void fj_tri_backward(const T* __restrict__ LU, int n, int lda, X* x, bool unit,
                                bool transposed) {
  // backward substitution with an upper factor M: M = U (trans=0) or L^T (trans=1)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ X diag[FJ_TS_B][FJ_TS_B + 1];
  const int nblk = (n + FJ_TS_B - 1) / FJ_TS_B;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int b0 = bi * FJ_TS_B;
    const int bs = min(FJ_TS_B, n - b0);
    for (int e = tid; e < FJ_TS_B * FJ_TS_B; e += blockDim.x) {
      const int i = e / FJ_TS_B, q = e % FJ_TS_B;
      X v = 0;
      if (i < bs && q < bs)
        v = transposed ? (X)LU[(size_t)(b0 + q) * lda + b0 + i] : (X)LU[(size_t)(b0 + i) * lda + b0 + q];
      diag[i][q] = v;
    }
    __syncthreads();
    if (wave == 0) {
      X xi = (lane < bs) ? x[b0 + lane] : X(0);
      for (int r = bs - 1; r >= 0; --r) {
        X xr = __shfl(xi, r);
        if (!unit) {
          xr = xr / diag[r][r];
          if (lane == r) xi = xr;
        }
        if (lane < r) xi -= diag[lane][r] * xr;
      }
      if (lane < bs) x[b0 + lane] = xi;
    }
    __syncthreads();
    for (int i = wave; i < b0; i += blockDim.x / 64) {
      X s = 0;
      if (lane < bs) {
        const X mv = transposed ? (X)LU[(size_t)(b0 + lane) * lda + i] : (X)LU[(size_t)i * lda + b0 + lane];
        s = mv * x[b0 + lane];
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if (lane == 0) x[i] -= s;
    }
    __syncthreads();
  }
}

#define FJ_SOLVE_MAXN 8192

// b (FP64, n) -> x (FP64, n) through the factor of type T; x may alias b.
template <typename T>
__global__ __launch_bounds__(FJ_LU_THREADS) void k_lu_solve(const T* __restrict__ LU, int n,
                                                             int lda,
                                                             const int32_t* __restrict__ ipiv,
                                                             const double* b, double* x,
                                                             int trans) {
  __shared__ double xs[FJ_SOLVE_MAXN];
  const int32_t* perm = ipiv + n;
  if (!trans) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) xs[i] = b[perm[i]];
    __syncthreads();
    fj_tri_forward<T, double>(LU, n, lda, xs, true, false);
    fj_tri_backward<T, double>(LU, n, lda, xs, false, false);
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = xs[i];
  } else {
    for (int i = threadIdx.x; i < n; i += blockDim.x) xs[i] = b[i];
    __syncthreads();
    fj_tri_forward<T, double>(LU, n, lda, xs, false, true);
    fj_tri_backward<T, double>(LU, n, lda, xs, true, true);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[perm[i]] = xs[i];
  }
}

extern "C" int fj_lu_solve(const double* LU, int n, int lda, const int32_t* ipiv, double* b,
                           int nrhs, int trans, void* stream) {
  if (n <= 0 || n > FJ_SOLVE_MAXN || lda < n || nrhs < 0 || !LU || !ipiv || !b) return FJ_ERR_ARG;
  for (int r = 0; r < nrhs; ++r) {
    hipLaunchKernelGGL(k_lu_solve<double>, dim3(1), dim3(FJ_LU_THREADS), 0, (hipStream_t)stream,
                       LU, n, lda, ipiv, b + (size_t)r * n, b + (size_t)r * n, trans);
    FJ_CHECK_LAUNCH();
  }
  return FJ_OK;
}

// ---- FP64 iterative refinement around an FP32 factor -----------------------
// r = b - A x (FP64); one wave per row, lanes over columns (coalesced rows).
// flags[0]: converged, flags[1]: sweeps done; amax[0]: |r|_inf scratch.
__global__ void k_residual(const double* __restrict__ A, int lda, int n,
                           const double* __restrict__ b, const double* __restrict__ x,
                           double* __restrict__ r, const int32_t* __restrict__ flags) {
  if (flags[0]) return;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (row >= n) return;
  const double* a = A + (size_t)row * lda;
  double s = 0.0;
  for (int c = lane; c < n; c += 64) s += a[c] * x[c];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) r[row] = b[row] - s;
}

// single workgroup: norms, convergence test, correction solve, update
template <typename T>
__global__ __launch_bounds__(FJ_LU_THREADS) void k_refine_step(
    const T* __restrict__ LU, int ldlu, const int32_t* __restrict__ ipiv, int n,
    const double* __restrict__ A_rownorm_max, const double* __restrict__ b,
    const double* __restrict__ r, double* __restrict__ x, double tol, int32_t* __restrict__ flags,
    int last) {
  __shared__ double xs[FJ_SOLVE_MAXN];
  __shared__ double red[3][FJ_LU_WAVES];
  if (flags[0]) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double rm = 0.0, xm = 0.0, bm = 0.0;
  for (int i = tid; i < n; i += blockDim.x) {
    rm = fmax(rm, fabs(r[i]));
    xm = fmax(xm, fabs(x[i]));
    bm = fmax(bm, fabs(b[i]));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    rm = fmax(rm, __shfl_xor(rm, off));
    xm = fmax(xm, __shfl_xor(xm, off));
    bm = fmax(bm, __shfl_xor(bm, off));
  }
  if (lane == 0) { red[0][wave] = rm; red[1][wave] = xm; red[2][wave] = bm; }
  __syncthreads();
  rm = red[0][0]; xm = red[1][0]; bm = red[2][0];
  for (int w = 1; w < FJ_LU_WAVES; ++w) {
    rm = fmax(rm, red[0][w]); xm = fmax(xm, red[1][w]); bm = fmax(bm, red[2][w]);
  }
  __syncthreads();
  if (rm <= tol * (A_rownorm_max[0] * xm + bm)) {
    if (tid == 0) flags[0] = 1;
    return;
  }
  if (last) {
    if (tid == 0) flags[0] = -1;
    return;
  }
  const int32_t* perm = ipiv + n;
  for (int i = tid; i < n; i += blockDim.x) xs[i] = r[perm[i]];
  __syncthreads();
  fj_tri_forward<T, double>(LU, n, ldlu, xs, true, false);
  fj_tri_backward<T, double>(LU, n, ldlu, xs, false, false);
  for (int i = tid; i < n; i += blockDim.x) x[i] += xs[i];
  if (tid == 0) flags[1] += 1;
}

// |A|_inf (max absolute row sum) for the stopping test
__global__ void k_rownorm(const double* __restrict__ A, int lda, int n, double* __restrict__ out) {
  __shared__ double red[FJ_LU_WAVES];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double best = 0.0;
  for (int row = wave; row < n; row += FJ_LU_WAVES) {
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s += fabs(A[(size_t)row * lda + c]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    best = fmax(best, s);
  }
  if (lane == 0) red[wave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < FJ_LU_WAVES; ++w) m = fmax(m, red[w]);
    out[0] = m;
  }
}

__global__ void k_refine_init(int32_t* flags) {
  flags[0] = 0;
  flags[1] = 0;
}

__global__ void k_refine_finish(const int32_t* flags, int32_t* iters) {
  iters[0] = (flags[0] == 1) ? flags[1] : -flags[1] - 1;
}

extern "C" size_t fj_refine_workspace_bytes(int n) {
  return ((size_t)(n + 8) * sizeof(double) + 64 + 255) / 256 * 256;
}

template <typename T>
static int fj_solve_refined_t(const double* A, int lda, const T* LU, int ldlu,
                              const int32_t* ipiv, int n, const double* b, double* x, int max_iter,
                              double tol, int32_t* iters, void* workspace, int x_is_guess,
                              hipStream_t st) {
  if (n <= 0 || n > FJ_SOLVE_MAXN || lda < n || ldlu < n || !A || !LU || !ipiv || !b || !x ||
      !iters || !workspace || max_iter < 0)
    return FJ_ERR_ARG;
  double* r = (double*)workspace;
  double* anorm = r + n;
  int32_t* flags = (int32_t*)(anorm + 4);
  hipLaunchKernelGGL(k_refine_init, dim3(1), dim3(1), 0, st, flags);
  hipLaunchKernelGGL(k_rownorm, dim3(1), dim3(FJ_LU_THREADS), 0, st, A, lda, n, anorm);
  // x0 = LU^{-1} b, unless the caller's x is the starting guess
  if (!x_is_guess)
    hipLaunchKernelGGL(k_lu_solve<T>, dim3(1), dim3(FJ_LU_THREADS), 0, st, LU, n, ldlu, ipiv, b, x,
                       0);
  FJ_CHECK_LAUNCH();
  for (int it = 0; it <= max_iter; ++it) {
    hipLaunchKernelGGL(k_residual, dim3(fj_cdiv(n, 4)), dim3(256), 0, st, A, lda, n, b, x, r,
                       flags);
    hipLaunchKernelGGL(k_refine_step<T>, dim3(1), dim3(FJ_LU_THREADS), 0, st, LU, ldlu, ipiv, n,
                       anorm, b, r, x, tol, flags, it == max_iter ? 1 : 0);
    FJ_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_refine_finish, dim3(1), dim3(1), 0, st, flags, iters);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

extern "C" int fj_solve_refined(const double* A, int lda, const float* LU32, int ldlu,
                                const int32_t* ipiv, int n, const double* b, double* x,
                                int max_iter, double tol, int32_t* iters, void* workspace,
                                void* stream) {
  return fj_solve_refined_t<float>(A, lda, LU32, ldlu, ipiv, n, b, x, max_iter, tol, iters,
                                   workspace, 0, (hipStream_t)stream);
}

extern "C" int fj_solve_refined_f64(const double* A, int lda, const double* LU, int ldlu,
                                    const int32_t* ipiv, int n, const double* b, double* x,
                                    int max_iter, double tol, int32_t* iters, void* workspace,
                                    void* stream) {
  return fj_solve_refined_t<double>(A, lda, LU, ldlu, ipiv, n, b, x, max_iter, tol, iters,
                                    workspace, 0, (hipStream_t)stream);
}
