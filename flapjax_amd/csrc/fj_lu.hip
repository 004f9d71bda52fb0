// K3: dense circulation solve, replacing jnp.linalg.solve (aero/uvlm.py:985).
//
// Blocked right-looking LU with partial pivoting on a row-major matrix.
//  * panel kernel (one 512-thread workgroup): the (n-k) x jb panel is held in
//    registers, each thread owning up to R rows; per column: register argmax,
//    a DPP wave maximum, each wave's candidate row and key published in LDS,
//    one barrier, the workgroup argmax reduced by every thread, row exchange
//    between the two owner threads, scale and rank-1 update in registers.  The composed row permutation is kept in
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
#include "fj_refine_dev.hpp"

#define FJ_LU_THREADS 1024
#define FJ_LU_WAVES (FJ_LU_THREADS / 64)

template <typename T>
__device__ __forceinline__ T fj_abs(T x) { return x < T(0) ? -x : x; }

// Pivot search key: |a| as the high bits (non-negative doubles order like
// their bit patterns; |a| clears the sign bit, so -0.0 keys like +0.0) with
// the low 13 mantissa bits replaced by 8191 - row, so one unsigned max picks
// the largest |a| and, among values equal to about 2^-39 relative, the
// smallest row.  Rows < 8192.
__device__ __forceinline__ unsigned long long fj_pivot_key(double a, int row) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(a) & 0x7FFFFFFFFFFFFFFFull;
  return (b & ~0x1FFFull) | (unsigned long long)(8191 - row);
}

// One step of a wave-wide unsigned maximum through the DPP crossbar: each
// lane takes the value of the lane CTRL names (a register move, no LDS round
// trip; the move folds into v_max_u32's DPP operand).  Lanes whose source is
// masked off or out of range read 0, the identity of the maximum.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned fj_dpp_max_u32(unsigned x) {
  const unsigned y = (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, 0xf, false);
  return x > y ? x : y;
}

// Maximum over the 64 lanes of a (fully active) wave: row maxima by quad
// permutations and row rotations, then the two row broadcasts carry them
// into lane 63, whose value every lane reads back.
__device__ __forceinline__ unsigned fj_wave_max_u32(unsigned x) {
  x = fj_dpp_max_u32<0xb1, 0xf>(x);   // quad_perm [1,0,3,2]
  x = fj_dpp_max_u32<0x4e, 0xf>(x);   // quad_perm [2,3,0,1]
  x = fj_dpp_max_u32<0x124, 0xf>(x);  // row_ror:4
  x = fj_dpp_max_u32<0x128, 0xf>(x);  // row_ror:8
  x = fj_dpp_max_u32<0x142, 0xa>(x);  // row_bcast:15 into rows 1, 3
  x = fj_dpp_max_u32<0x143, 0xc>(x);  // row_bcast:31 into rows 2, 3
  return (unsigned)__builtin_amdgcn_readlane((int)x, 63);
}

// 64-bit maximum in two 32-bit passes: the high words, then the low words of
// the lanes that hold the maximal high word.
__device__ __forceinline__ unsigned long long fj_wave_max_u64(unsigned long long key) {
  const unsigned hi = (unsigned)(key >> 32);
  const unsigned mh = fj_wave_max_u32(hi);
  const unsigned ml = fj_wave_max_u32(hi == mh ? (unsigned)key : 0u);
  return ((unsigned long long)mh << 32) | ml;
}

// NT threads, each owning the panel slots tid + r * NT (r < R), all JB
// columns of a slot's row in registers.  Rows never move between threads:
// each slot carries the logical row it currently holds (lg), a row exchange
// only swaps two logical indices, and the write-back stores each slot to its
// logical row.  Pivot keys use the logical row, so ties go to the smallest
// logical row as in LAPACK.  The column loop is unrolled at compile time
// (fj_panel_col<JJ> calls fj_panel_col<JJ + 1>), so every register index is
// a constant.  (Measured on the panel shapes below: a loop unrolled in groups
// of 2, 4 or 8 columns with the registers rotated between groups, which
// keeps the code inside the instruction cache, ran 3-20 % slower.)
//
// A column step is bound by the instructions every wave issues (~400 per
// column and wave, two waves per SIMD: ~3000 cycles whatever the panel
// height), so only the chain from one pivot to the next sits between the
// barriers:
//   keys of column jj -> DPP wave maximum -> LDS            | barrier A
//   (while waiting: the rank-1 update of column jj-1 on the columns
//    jj+1.., deferred from the previous step)
//   key reduction -> the global winner publishes its row   | barrier B
//   pivot, logical exchange, multipliers, and the update of column jj+1
//   only (the rest of column jj's update is deferred to the next step).
// Every element still receives its updates in column order, each as the
// same FMA, so the arithmetic is that of the plain rank-1 sequence.
template <typename T, int NT, int JB>
struct FjPanelLds {
  unsigned long long key[2][NT / 64];
  T prow[2][JB];
  int ipiv[JB];
  int keep[NT];
};

// a copy the optimiser cannot see through (an empty asm on the register)
template <typename T>
__device__ __forceinline__ T fj_opaque(T x) {
  asm("" : "+v"(x));
  return x;
}

// Workgroup barrier ordering LDS accesses only: global stores may stay in
// flight across it (__syncthreads() would wait for every outstanding global
// access of the wave first).
__device__ __forceinline__ void fj_lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T, int NT, int R, int JB, int JJ>
__device__ __forceinline__ void fj_panel_col(T (&a)[R][JB], int (&lg)[R], T (&l)[R],
                                             FjPanelLds<T, NT, JB>& sh, int tid, int rows, int k,
                                             int32_t* __restrict__ info) {
  if constexpr (JJ < JB) {
    constexpr int jj = JJ;
    if (jj >= rows) return;  // uniform: the panel is narrower than JB
    constexpr int NW = NT / 64;
    constexpr int buf = JJ & 1;
    const int lane = tid & 63, wave = tid >> 6;
    // this thread's best candidate among the logical rows jj .. (padding
    // slots hold zeros and larger logical rows than every real row, so they
    // never win)
    unsigned long long key = 0ull;
    int kr = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const unsigned long long kk = (lg[r] >= jj) ? fj_pivot_key((double)a[r][JJ], lg[r]) : 0ull;
      kr = kk > key ? r : kr;
      key = kk > key ? kk : key;
    }
    const unsigned long long wkey = fj_wave_max_u64(key);
    if (lane == 0) sh.key[buf][wave] = wkey;
    // the deferred rank-1 update of column jj-1 on columns jj+1 ..
    if constexpr (JJ > 0) {
#pragma unroll
      for (int c = JJ + 1; c < JB; ++c) {
        const T pc = sh.prow[buf ^ 1][c];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r][c] -= l[r] * pc;
      }
    }
    fj_lds_barrier();  // A
    unsigned long long kv[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) kv[w] = sh.key[buf][w];
#pragma unroll
    for (int s = 1; s < NW; s *= 2)
#pragma unroll
      for (int w = 0; w + s < NW; w += 2 * s) kv[w] = kv[w + s] > kv[w] ? kv[w + s] : kv[w];
    const unsigned long long best = kv[0];
    const int p = 8191 - (int)(best & 0x1FFFull);  // logical pivot row
    // keys are distinct (the row sits in the low bits): exactly one lane
    // holds the winning key, and it publishes its row's columns JJ ..  A
    // select chain picks the slot; it starts from an opaque copy, because a
    // select between two loads of `a` would be folded into a load from a
    // selected address, which keeps `a` in scratch memory.
    if (key == best) {
#pragma unroll
      for (int c = JJ; c < JB; ++c) {
        T v = fj_opaque(a[0][c]);
#pragma unroll
        for (int r = 1; r < R; ++r) v = (kr == r) ? a[r][c] : v;
        sh.prow[buf][c] = v;
      }
    }
    if (tid == 0) {
      sh.ipiv[jj] = k + p;
      if ((best >> 13) == 0ull && *info == 0) *info = k + jj + 1;  // zero column
    }
    fj_lds_barrier();  // B
    // exchange logical rows jj and p; multipliers of the rows below jj (the
    // other slots get l = 0, so the rank-1 updates need no per-slot
    // predicate: they leave them unchanged)
    const T piv = sh.prow[buf][JJ];
    const T inv = (piv != T(0)) ? T(1) / piv : T(0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      lg[r] = (lg[r] == jj) ? p : ((lg[r] == p) ? jj : lg[r]);
      const bool below = lg[r] > jj && lg[r] < rows;
      const T lv = below ? a[r][JJ] * inv : T(0);
      l[r] = lv;
      if (below) a[r][JJ] = lv;
    }
    if constexpr (JJ + 1 < JB) {
      const T pc = sh.prow[buf][JJ + 1];
#pragma unroll
      for (int r = 0; r < R; ++r) a[r][JJ + 1] -= l[r] * pc;
    }
    fj_panel_col<T, NT, R, JB, JJ + 1>(a, lg, l, sh, tid, rows, k, info);
  }
}

template <typename T, int NT, int R, int JB>
__global__ __launch_bounds__(NT) void k_lu_panel(T* __restrict__ A, int n, int lda, int k,
                                                 int32_t* __restrict__ ipiv,
                                                 int32_t* __restrict__ info) {
  static_assert(NT % 64 == 0 && JB <= NT && NT * R <= 8192 && JB % 2 == 0, "panel shape");
  // one row block of the panel staged through LDS on load and store
  // (padded against bank conflicts)
  __shared__ T s_stage[NT][JB + 1];
  __shared__ FjPanelLds<T, NT, JB> sh;
  const int tid = threadIdx.x;
  const int jb = min(JB, n - k);
  const int rows = n - k;

  // Load: block r (rows r*NT .. r*NT+NT-1) is read with consecutive lanes on
  // consecutive columns of a row (one cache line per 64 / JB rows instead of
  // one per lane) -- every block's loads issued at once, into the registers
  // that later hold the block -- then each block is transposed through the
  // LDS block.
  // (the write-back moves pairs of columns, JB being even, as one 16-byte
  // store for FP64 when the pair is aligned in memory)
  constexpr int JH = JB / 2;
  T a[R][JB];
  int lg[R];
  const bool al = (sizeof(T) == 8) && ((((size_t)lda | (size_t)k) & 1) == 0);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    lg[r] = tid + r * NT;
#pragma unroll
    for (int i = 0; i < JB; ++i) {
      const int e = tid + i * NT, lr = r * NT + e / JB, c = e % JB;
      a[r][i] = (lr < rows && c < jb) ? A[(size_t)(k + lr) * lda + k + c] : T(0);
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int i = 0; i < JB; ++i) {
      const int e = tid + i * NT;
      s_stage[e / JB][e % JB] = a[r][i];
    }
    fj_lds_barrier();
#pragma unroll
    for (int c = 0; c < JB; ++c) a[r][c] = s_stage[tid][c];
    fj_lds_barrier();
  }

  T l[R];
  fj_panel_col<T, NT, R, JB, 0>(a, lg, l, sh, tid, rows, k, info);
  fj_lds_barrier();

  // write-back: slots still holding their own row go through the LDS block
  // (coalesced); the at most 2 jb slots whose logical row changed store
  // their row directly (the exchange is a bijection, so no position is
  // written by both paths)
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const bool keep = lg[r] == tid + r * NT;
    if (keep) {
#pragma unroll
      for (int c = 0; c < JB; ++c) s_stage[tid][c] = a[r][c];
    }
    sh.keep[tid] = keep ? 1 : 0;
    fj_lds_barrier();
#pragma unroll
    for (int i = 0; i < JH; ++i) {
      const int e = tid + i * NT, q = e / JH, lr = r * NT + q, c = 2 * (e % JH);
      if (lr < rows && sh.keep[q]) {
        T* dst = A + (size_t)(k + lr) * lda + k + c;
        if (c + 1 < jb && al) {
          *reinterpret_cast<double2*>(dst) = double2{(double)s_stage[q][c], (double)s_stage[q][c + 1]};
        } else {
          if (c < jb) dst[0] = s_stage[q][c];
          if (c + 1 < jb) dst[1] = s_stage[q][c + 1];
        }
      }
    }
    fj_lds_barrier();
    if (!keep) {
#pragma unroll
      for (int c = 0; c < JB; ++c)
        if (c < jb) A[(size_t)(k + lg[r]) * lda + k + c] = a[r][c];
    }
  }
  if (tid < jb) ipiv[k + tid] = sh.ipiv[tid];
  // composed permutation ipiv[n + q]: only positions k..k+jb-1 and the pivot
  // rows change; each is traced back through this panel's exchanges (last
  // exchange first) to the position its content came from.
  {
    int32_t* perm = ipiv + n;
    int q = -1, src = -1, val = 0;
    if (tid < jb) q = k + tid;
    else if (tid < 2 * jb) q = sh.ipiv[tid - jb];
    if (q >= 0) {
      src = q;
      for (int jj = jb - 1; jj >= 0; --jj) {
        const int a0 = k + jj, b0 = sh.ipiv[jj];
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
// With a two-level blocking (below) the exchanges of a sub-panel reach every
// column while its U12 is only formed inside the current outer block
// (trsm_end); the outer block's own U12 is then formed once, without the
// exchanges (do_swap = 0).
//
// The exchanges are applied as one gather, not one after another: only rows
// k .. k+jb-1 and the pivot rows change, and the final content of each such
// row q is the original content of the row src(q) found by undoing the
// exchanges from the last to the first.  Each thread loads all its sources
// at once (independent loads in flight together, where a sequence of
// exchanges would wait out one memory round trip per exchange), solves the
// top jb rows in registers and stores.
#define FJ_SWAP_THREADS 128
template <typename T, int JB>
__global__ __launch_bounds__(FJ_SWAP_THREADS) void k_lu_swap_trsm(T* __restrict__ A, int n, int lda,
                                                                   int k, int jb,
                                                                   const int32_t* __restrict__ ipiv,
                                                                   int trsm_end, int do_swap) {
  __shared__ T s_L11[JB][JB];
  __shared__ int s_piv[JB], s_q[2 * JB], s_src[2 * JB];
  const int tid = threadIdx.x;
  for (int e = tid; e < JB * JB; e += FJ_SWAP_THREADS) {
    const int r = e / JB, q = e % JB;
    s_L11[r][q] = (r < jb && q < jb && q < r) ? A[(size_t)(k + r) * lda + k + q] : T(0);
  }
  if (tid < JB) s_piv[tid] = (tid < jb) ? ipiv[k + tid] : 0;
  __syncthreads();
  for (int t = tid; t < 2 * JB; t += FJ_SWAP_THREADS) {
    // target rows: k + t (t < jb), then the pivot rows; -1 marks "none"
    int q = -1, src = -1;
    if (t < jb) q = k + t;
    else if (do_swap && t >= JB && t - JB < jb) q = s_piv[t - JB];
    if (q >= 0) {
      src = q;
      if (do_swap)
        for (int jj = jb - 1; jj >= 0; --jj) {
          const int a0 = k + jj, b0 = s_piv[jj];
          if (src == a0) src = b0;
          else if (src == b0) src = a0;
        }
    }
    // a pivot row inside k .. k+jb-1 is written as a top row already
    if (t >= JB && q >= k && q < k + jb) q = -1;
    s_q[t] = q;
    s_src[t] = src;
  }
  __syncthreads();
  int c = blockIdx.x * FJ_SWAP_THREADS + tid;
  if (do_swap) {  // every column but the panel's
    if (c >= n - jb) return;
    if (c >= k) c += jb;
  } else {  // the trsm columns only
    c += k + jb;
    if (c >= trsm_end) return;
  }
  T v[2 * JB];
#pragma unroll
  for (int t = 0; t < 2 * JB; ++t) {
    const int src = s_src[t];
    v[t] = (s_q[t] >= 0) ? A[(size_t)src * lda + c] : T(0);
  }
  if (c >= k + jb && c < trsm_end) {
    // forward substitution with the unit lower L11 (rows past jb are zero)
#pragma unroll
    for (int r = 1; r < JB; ++r) {
      T acc = v[r];
#pragma unroll
      for (int q = 0; q < r; ++q) acc -= s_L11[r][q] * v[q];
      v[r] = acc;
    }
  }
#pragma unroll
  for (int t = 0; t < 2 * JB; ++t) {
    const int q = s_q[t];
    if (q >= 0) A[(size_t)q * lda + c] = v[t];
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

// Columns limited to [k + kb, c_end) (the outer block during a two-level
// factorisation), rows to [k + kb, n).
template <typename T>
__global__ __launch_bounds__(256) void k_lu_update(T* __restrict__ A, int n, int lda, int k,
                                                   int kb, int c_end) {
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
    sU[q][j] = (c0 + j < c_end && q < kb) ? A[(size_t)(k + q) * lda + c0 + j] : T(0);
  }
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  // the lane's 16 entries of C, loaded with the operand tiles so that their
  // memory latency overlaps the staging and the matrix-core loop (C lies
  // outside the L and U panels; the update is the same c - acc per entry)
  T cv[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = r0 + wr + a * 16 + FjMfma<T>::row(lk, v);
        const int j = c0 + wc + b * 16 + li;
        cv[a][b][v] = (i < n && j < c_end) ? A[(size_t)i * lda + j] : T(0);
      }
  __syncthreads();
  typedef typename FjMfma<T>::acc_t acc_t;
  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
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
        if (i < n && j < c_end) A[(size_t)i * lda + j] = cv[a][b][v] - acc[a][b][v];
      }
}

template <typename T>
__global__ __launch_bounds__(64) void k_diag_inverses(const T* __restrict__ LU, int n, int lda,
                                                      T* __restrict__ aux);
__host__ __device__ static inline int fj_nblk(int n);

template <typename T, int NT, int R, int JB>
static int fj_lu_factor_t(T* A, int n, int lda, int32_t* ipiv, int32_t* info, T* aux,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_perm_init, dim3(fj_cdiv(n, 256)), dim3(256), 0, st, ipiv, n, info);
  FJ_CHECK_LAUNCH();
  // Two-level blocking: register panels of JB columns (as wide as one
  // workgroup's registers allow at this n) inside outer blocks of FJ_GK
  // columns.  Sub-panel updates stay inside the outer block; the trailing
  // matrix gets one rank-FJ_GK matrix-core update per outer block instead of
  // FJ_GK / JB narrow ones, each of which streams the whole trailing matrix.
  constexpr int NB = FJ_GK;
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nb = (n - k0 < NB) ? n - k0 : NB;
    const int kend = k0 + nb;
    for (int k = k0; k < kend; k += JB) {
      const int jb = (n - k < JB) ? n - k : JB;
      hipLaunchKernelGGL((k_lu_panel<T, NT, R, JB>), dim3(1), dim3(NT), 0, st, A, n, lda, k, ipiv,
                         info);
      FJ_CHECK_LAUNCH();
      if (n - jb > 0) {
        hipLaunchKernelGGL((k_lu_swap_trsm<T, JB>), dim3(fj_cdiv(n - jb, FJ_SWAP_THREADS)),
                           dim3(FJ_SWAP_THREADS), 0, st, A,
                           n, lda, k, jb, ipiv, JB == NB ? n : kend, 1);
        FJ_CHECK_LAUNCH();
      }
      const int c_end = JB == NB ? n : kend;
      const int rows = n - k - jb, cols = c_end - k - jb;
      if (rows > 0 && cols > 0) {
        const dim3 g(fj_cdiv(cols, FJ_GT), fj_cdiv(rows, FJ_GT));
        hipLaunchKernelGGL(k_lu_update<T>, g, dim3(256), 0, st, A, n, lda, k, jb, c_end);
        FJ_CHECK_LAUNCH();
      }
    }
    if (JB < NB && kend < n) {
      // the outer block's U12 (its exchanges are already applied) and the
      // rank-nb trailing update
      hipLaunchKernelGGL((k_lu_swap_trsm<T, NB>), dim3(fj_cdiv(n - kend, FJ_SWAP_THREADS)),
                         dim3(FJ_SWAP_THREADS), 0, st, A, n, lda, k0, nb, ipiv, n, 0);
      FJ_CHECK_LAUNCH();
      const int rest = n - kend;
      const dim3 g(fj_cdiv(rest, FJ_GT), fj_cdiv(rest, FJ_GT));
      hipLaunchKernelGGL(k_lu_update<T>, g, dim3(256), 0, st, A, n, lda, k0, nb, n);
      FJ_CHECK_LAUNCH();
    }
  }
  hipLaunchKernelGGL(k_diag_inverses<T>, dim3(fj_nblk(n), 2), dim3(64), 0, st, A, n, lda, aux);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

template <typename T>
static int fj_lu_dispatch(T* A, int n, int lda, int32_t* ipiv, int32_t* info, T* aux,
                          hipStream_t st) {
  if (n <= 0 || n > 8192 || lda < n || !A || !ipiv || !info || !aux) return FJ_ERR_ARG;
  // panel shapes (threads, rows per thread, columns held in registers;
  // threads x rows >= n), the fastest measured at each size
  if (n <= 1024) return fj_lu_factor_t<T, 512, 2, 32>(A, n, lda, ipiv, info, aux, st);
  if (n <= 2048) return fj_lu_factor_t<T, 512, 4, 16>(A, n, lda, ipiv, info, aux, st);
  if (n <= 4096) return fj_lu_factor_t<T, 512, 8, 8>(A, n, lda, ipiv, info, aux, st);
  return fj_lu_factor_t<T, 1024, 8, 4>(A, n, lda, ipiv, info, aux, st);
}

extern "C" int fj_lu_factor(double* A, int n, int lda, int32_t* ipiv, int32_t* info, void* aux,
                            void* stream) {
  return fj_lu_dispatch<double>(A, n, lda, ipiv, info, (double*)aux, (hipStream_t)stream);
}

extern "C" int fj_lu_factor_f32(float* A, int n, int lda, int32_t* ipiv, int32_t* info, void* aux,
                                void* stream) {
  return fj_lu_dispatch<float>(A, n, lda, ipiv, info, (float*)aux, (hipStream_t)stream);
}

// ---- triangular solves -------------------------------------------------------
// The factor carries an auxiliary array: the inverses of the 64x64 diagonal
// blocks of L (unit lower) and U (upper), written by fj_lu_factor.  A solve is
// then a sequence of 64-row blocks, each a 64x64 mat-vec with the block
// inverse (one wave, lanes over rows) followed by the update of the remaining
// rows (one thread per row, a contiguous 64-element row segment each); x
// stays in LDS.  The transposed solve uses the transposed block inverses and
// reads the factor column-wise (coalesced across threads).
#define FJ_TS_B 64
#define FJ_SOLVE_MAXN 8192

__host__ __device__ static inline int fj_nblk(int n) { return (n + FJ_TS_B - 1) / FJ_TS_B; }

extern "C" size_t fj_lu_aux_bytes(int n) {
  return (size_t)2 * fj_nblk(n) * FJ_TS_B * FJ_TS_B * sizeof(double) + 256;
}

// aux layout (type T): Linv blocks then Uinv blocks, each (nblk, 64, 64) row-major
template <typename T>
__global__ __launch_bounds__(64) void k_diag_inverses(const T* __restrict__ LU, int n, int lda,
                                                      T* __restrict__ aux) {
  __shared__ double sb[FJ_TS_B][FJ_TS_B + 1];
  __shared__ double sy[FJ_TS_B][FJ_TS_B + 1];
  const int b = blockIdx.x, which = blockIdx.y;  // 0: L, 1: U
  const int nblk = fj_nblk(n);
  const int b0 = b * FJ_TS_B, bs = min(FJ_TS_B, n - b0);
  const int j = threadIdx.x;  // column of the inverse
  for (int r = 0; r < FJ_TS_B; ++r) sb[r][j] = (r < bs && j < bs) ? (double)LU[(size_t)(b0 + r) * lda + b0 + j] : 0.0;
  __syncthreads();
  if (which == 0) {
    // unit lower: y[r] = delta_rj - sum_{q<r} L[r][q] y[q]
    for (int r = 0; r < FJ_TS_B; ++r) {
      double acc = (r == j) ? 1.0 : 0.0;
      for (int q = 0; q < r; ++q) acc -= sb[r][q] * sy[q][j];
      sy[r][j] = (r < bs && j < bs) ? acc : 0.0;
    }
  } else {
    for (int r = FJ_TS_B - 1; r >= 0; --r) {
      double acc = (r == j) ? 1.0 : 0.0;
      for (int q = r + 1; q < FJ_TS_B; ++q) acc -= sb[r][q] * sy[q][j];
      sy[r][j] = (r < bs && j < bs) ? acc / sb[r][r] : 0.0;
    }
  }
  T* o = aux + ((size_t)which * nblk + b) * FJ_TS_B * FJ_TS_B;
  for (int r = 0; r < FJ_TS_B; ++r) o[r * FJ_TS_B + j] = (T)sy[r][j];
}

// forward sweep with a lower-triangular operator M (M = L, or U^T when trans)
template <typename T>
__device__ void fj_sweep_forward(const T* __restrict__ LU, const T* __restrict__ inv, int n,
                                 int lda, double* x, bool trans) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nblk = fj_nblk(n);
  for (int b = 0; b < nblk; ++b) {
    const int b0 = b * FJ_TS_B, bs = min(FJ_TS_B, n - b0);
    const T* Bi = inv + (size_t)b * FJ_TS_B * FJ_TS_B;
    if (wave == 0) {
      double acc = 0.0;
      if (lane < bs)
        for (int q = 0; q < bs; ++q)
          acc += (double)(trans ? Bi[q * FJ_TS_B + lane] : Bi[lane * FJ_TS_B + q]) * x[b0 + q];
      if (lane < bs) x[b0 + lane] = acc;  // all lanes read before any writes (one wave)
    }
    __syncthreads();
    for (int i = b0 + bs + tid; i < n; i += blockDim.x) {
      double acc = 0.0;
      if (!trans) {
        const T* row = LU + (size_t)i * lda + b0;
        for (int q = 0; q < bs; ++q) acc += (double)row[q] * x[b0 + q];
      } else {
        for (int q = 0; q < bs; ++q) acc += (double)LU[(size_t)(b0 + q) * lda + i] * x[b0 + q];
      }
      x[i] -= acc;
    }
    __syncthreads();
  }
}

// backward sweep with an upper-triangular operator M (M = U, or L^T when trans)
template <typename T>
__device__ void fj_sweep_backward(const T* __restrict__ LU, const T* __restrict__ inv, int n,
                                  int lda, double* x, bool trans) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nblk = fj_nblk(n);
  for (int b = nblk - 1; b >= 0; --b) {
    const int b0 = b * FJ_TS_B, bs = min(FJ_TS_B, n - b0);
    const T* Bi = inv + (size_t)b * FJ_TS_B * FJ_TS_B;
    if (wave == 0) {
      double acc = 0.0;
      if (lane < bs)
        for (int q = 0; q < bs; ++q)
          acc += (double)(trans ? Bi[q * FJ_TS_B + lane] : Bi[lane * FJ_TS_B + q]) * x[b0 + q];
      if (lane < bs) x[b0 + lane] = acc;
    }
    __syncthreads();
    for (int i = tid; i < b0; i += blockDim.x) {
      double acc = 0.0;
      if (!trans) {
        const T* row = LU + (size_t)i * lda + b0;
        for (int q = 0; q < bs; ++q) acc += (double)row[q] * x[b0 + q];
      } else {
        for (int q = 0; q < bs; ++q) acc += (double)LU[(size_t)(b0 + q) * lda + i] * x[b0 + q];
      }
      x[i] -= acc;
    }
    __syncthreads();
  }
}

// x <- solve(b) with the factor; x may alias b
template <typename T>
__device__ void fj_lu_solve_block(const T* __restrict__ LU, int n, int lda,
                                  const int32_t* __restrict__ ipiv, const T* __restrict__ aux,
                                  const double* b, double* x, int trans, double* xs) {
  const int nblk = fj_nblk(n);
  const T* Linv = aux;
  const T* Uinv = aux + (size_t)nblk * FJ_TS_B * FJ_TS_B;
  const int32_t* perm = ipiv + n;
  if (!trans) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) xs[i] = b[perm[i]];
    __syncthreads();
    fj_sweep_forward<T>(LU, Linv, n, lda, xs, false);
    fj_sweep_backward<T>(LU, Uinv, n, lda, xs, false);
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = xs[i];
  } else {
    for (int i = threadIdx.x; i < n; i += blockDim.x) xs[i] = b[i];
    __syncthreads();
    fj_sweep_forward<T>(LU, Uinv, n, lda, xs, true);
    fj_sweep_backward<T>(LU, Linv, n, lda, xs, true);
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[perm[i]] = xs[i];
  }
  __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(FJ_LU_THREADS) void k_lu_solve(const T* __restrict__ LU, int n,
                                                             int lda,
                                                             const int32_t* __restrict__ ipiv,
                                                             const T* __restrict__ aux,
                                                             const double* b, double* x,
                                                             int trans) {
  __shared__ double xs[FJ_SOLVE_MAXN];
  fj_lu_solve_block<T>(LU, n, lda, ipiv, aux, b, x, trans, xs);
}

extern "C" int fj_lu_solve(const double* LU, int n, int lda, const int32_t* ipiv, const void* aux,
                           double* b, int nrhs, int trans, void* stream) {
  if (n <= 0 || n > FJ_SOLVE_MAXN || lda < n || nrhs < 0 || !LU || !ipiv || !aux || !b)
    return FJ_ERR_ARG;
  for (int r = 0; r < nrhs; ++r) {
    hipLaunchKernelGGL(k_lu_solve<double>, dim3(1), dim3(FJ_LU_THREADS), 0, (hipStream_t)stream,
                       LU, n, lda, ipiv, (const double*)aux, b + (size_t)r * n, b + (size_t)r * n,
                       trans);
    FJ_CHECK_LAUNCH();
  }
  return FJ_OK;
}

// ---- FP64 iterative refinement ------------------------------------------------
// Mixed-precision refinement in the form of LAPACK dsgesv (Buttari, Dongarra
// et al., "Mixed precision iterative refinement techniques for the solution
// of dense linear systems", 2007): x0 = M b, then r = b - A x (FP64, the
// CURRENT A), x += M r, where M is an approximate inverse of A -- the LU
// factor of A (FP32 or FP64), or of a nearby matrix, or its explicit inverse.
// Stop test (dsgesv's backward-error bound): converged when
//   ||r||_inf <= tol ||A||_inf ||x||_inf       (tol default sqrt(n) eps64).
//
// Device state in the workspace: r (n doubles), then per sweep k = 0..max_iter
// a slot of three non-negative doubles kept as bit patterns for unsigned
// max-atomics {||A||_inf, ||r||_inf, ||x_k||_inf}, then flags {status
// (0 running, 1 converged), corrections made}.  Every launch reads the flags
// first, so once converged the remaining sweeps exit at entry.
__global__ void k_refine_init(fj_u64* __restrict__ slots, int nslots, int32_t* __restrict__ flags) {
  for (int i = threadIdx.x; i < 3 * nslots; i += blockDim.x) slots[i] = 0ull;
  if (threadIdx.x == 0) { flags[0] = 0; flags[1] = 0; }
}

// r = b - A x, one wave per row (lanes over the columns, coalesced);
// slot k gathers ||A||_inf and ||r||_inf.
__global__ __launch_bounds__(64 * FJ_RV_WAVES) void k_refine_residual(
    const double* __restrict__ A, int lda, int n, const double* __restrict__ b,
    const double* __restrict__ x, double* __restrict__ r, fj_u64* __restrict__ slot,
    const int32_t* __restrict__ flags) {
  __shared__ double sred[2][FJ_RV_WAVES];
  if (flags[0]) return;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * FJ_RV_WAVES + (threadIdx.x >> 6);
  double s = 0.0, a = 0.0;
  if (row < n) {
    const double* ar = A + (size_t)row * lda;
#pragma unroll 4
    for (int c = lane; c < n; c += 64) {
      const double v = ar[c];
      s += v * x[c];
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
    if (lane == 0) r[row] = rr;
  }
  fj_group_max(a, &slot[0], sred[0]);
  fj_group_max(fabs(rr), &slot[1], sred[1]);
}

// x = M v (first solve, v = b) or x += M v (correction, v = r) with the
// explicit inverse M (FP64 or FP32), one wave per row; ||x||_inf into xslot.
// A correction first applies the stop test of sweep k.
template <typename T, bool CORRECT>
__global__ __launch_bounds__(64 * FJ_RV_WAVES) void k_refine_apply_inv(
    const T* __restrict__ M, int ldm, int n, const double* __restrict__ v,
    double* __restrict__ x, fj_u64* __restrict__ slots, int k, double tol,
    int32_t* __restrict__ flags) {
  __shared__ double sred[FJ_RV_WAVES];
  if (flags[0]) return;
  if (CORRECT && fj_refine_converged(slots + 3 * k, tol)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) { flags[0] = 1; flags[1] = k; }
    return;
  }
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * FJ_RV_WAVES + (threadIdx.x >> 6);
  double s = 0.0;
  if (row < n) {
    const T* mr = M + (size_t)row * ldm;
#pragma unroll 4
    for (int c = lane; c < n; c += 64) s += (double)mr[c] * v[c];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  double xn = 0.0;
  if (row < n) {
    xn = CORRECT ? x[row] + s : s;
    if (lane == 0) x[row] = xn;
  }
  fj_group_max(fabs(xn), &slots[3 * (CORRECT ? k + 1 : 0) + 2], sred);
  if (CORRECT && blockIdx.x == 0 && threadIdx.x == 0) flags[1] = k + 1;
}

// x = LU^{-1} b (first solve) or x += LU^{-1} r (correction) in one
// workgroup with the triangular sweeps above; same slots and stop test.
template <typename T, bool CORRECT>
__global__ __launch_bounds__(FJ_LU_THREADS) void k_refine_apply_lu(
    const T* __restrict__ LU, int ldlu, const int32_t* __restrict__ ipiv,
    const T* __restrict__ aux, int n, const double* __restrict__ v, double* __restrict__ x,
    fj_u64* __restrict__ slots, int k, double tol, int32_t* __restrict__ flags) {
  __shared__ double xs[FJ_SOLVE_MAXN];
  __shared__ double sred[FJ_LU_WAVES];
  if (flags[0]) return;
  if (CORRECT && fj_refine_converged(slots + 3 * k, tol)) {
    if (threadIdx.x == 0) { flags[0] = 1; flags[1] = k; }
    return;
  }
  fj_lu_solve_block<T>(LU, n, ldlu, ipiv, aux, v, xs, 0, xs);
  double xm = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double xn = CORRECT ? x[i] + xs[i] : xs[i];
    x[i] = xn;
    xm = fmax(xm, fabs(xn));
  }
  fj_group_max(xm, &slots[3 * (CORRECT ? k + 1 : 0) + 2], sred);
  if (CORRECT && threadIdx.x == 0) flags[1] = k + 1;
}

// iters = corrections made, negated when the stop test never passed (the
// last slot is tested here, after the final residual)
__global__ void k_refine_finish(const fj_u64* __restrict__ slots, int max_iter, double tol,
                                const int32_t* __restrict__ flags, int32_t* __restrict__ iters) {
  if (flags[0] == 1) iters[0] = flags[1];
  else iters[0] = fj_refine_converged(slots + 3 * max_iter, tol) ? max_iter : -max_iter;
}

extern "C" size_t fj_refine_workspace_bytes(int n) {
  return ((size_t)(n + 4) * sizeof(double) + 3 * (FJ_REFINE_MAX_ITER + 1) * sizeof(fj_u64) + 64 +
          255) / 256 * 256;
}

// apply(first, k) launches x = M b (first) or the correction of sweep k
template <typename Apply>
static int fj_refine_run(const double* A, int lda, int n, const double* b, double* x,
                         int max_iter, double tol, int32_t* iters, void* workspace,
                         hipStream_t st, Apply apply) {
  const FjRefineWs w = fj_refine_ws(workspace, n);
  const int rgrid = fj_cdiv(n, FJ_RV_WAVES);
  hipLaunchKernelGGL(k_refine_init, dim3(1), dim3(256), 0, st, w.slots, max_iter + 1, w.flags);
  FJ_CHECK_LAUNCH();
  apply(true, 0, w);
  for (int k = 0; k < max_iter; ++k) {
    hipLaunchKernelGGL(k_refine_residual, dim3(rgrid), dim3(64 * FJ_RV_WAVES), 0, st, A, lda, n, b,
                       x, w.r, w.slots + 3 * k, w.flags);
    FJ_CHECK_LAUNCH();
    apply(false, k, w);
  }
  hipLaunchKernelGGL(k_refine_residual, dim3(rgrid), dim3(64 * FJ_RV_WAVES), 0, st, A, lda, n, b, x,
                     w.r, w.slots + 3 * max_iter, w.flags);
  hipLaunchKernelGGL(k_refine_finish, dim3(1), dim3(1), 0, st, w.slots, max_iter, tol, w.flags,
                     iters);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

template <typename T>
static int fj_solve_refined_t(const double* A, int lda, const T* LU, int ldlu,
                              const int32_t* ipiv, const T* aux, int n, const double* b,
                              double* x, int max_iter, double tol, int32_t* iters,
                              void* workspace, hipStream_t st) {
  if (n <= 0 || n > FJ_SOLVE_MAXN || lda < n || ldlu < n || !A || !LU || !ipiv || !aux || !b ||
      !x || !iters || !workspace || max_iter < 1 || max_iter > FJ_REFINE_MAX_ITER)
    return FJ_ERR_ARG;
  tol = fj_refine_tol(tol, n);
  return fj_refine_run(A, lda, n, b, x, max_iter, tol, iters, workspace, st,
                       [&](bool first, int k, const FjRefineWs& w) {
                         if (first)
                           hipLaunchKernelGGL((k_refine_apply_lu<T, false>), dim3(1),
                                              dim3(FJ_LU_THREADS), 0, st, LU, ldlu, ipiv, aux, n,
                                              b, x, w.slots, 0, tol, w.flags);
                         else
                           hipLaunchKernelGGL((k_refine_apply_lu<T, true>), dim3(1),
                                              dim3(FJ_LU_THREADS), 0, st, LU, ldlu, ipiv, aux, n,
                                              w.r, x, w.slots, k, tol, w.flags);
                       });
}

extern "C" int fj_solve_refined(const double* A, int lda, const float* LU32, int ldlu,
                                const int32_t* ipiv, const void* aux, int n, const double* b,
                                double* x, int max_iter, double tol, int32_t* iters,
                                void* workspace, void* stream) {
  return fj_solve_refined_t<float>(A, lda, LU32, ldlu, ipiv, (const float*)aux, n, b, x, max_iter,
                                   tol, iters, workspace, (hipStream_t)stream);
}

extern "C" int fj_solve_refined_f64(const double* A, int lda, const double* LU, int ldlu,
                                    const int32_t* ipiv, const void* aux, int n, const double* b,
                                    double* x, int max_iter, double tol, int32_t* iters,
                                    void* workspace, void* stream) {
  return fj_solve_refined_t<double>(A, lda, LU, ldlu, ipiv, (const double*)aux, n, b, x,
                                    max_iter, tol, iters, workspace, (hipStream_t)stream);
}

template <typename T>
static int fj_solve_refined_inv_t(const double* A, int lda, const T* M, int ldm, int n,
                                  const double* b, double* x, int max_iter, double tol,
                                  int32_t* iters, void* workspace, hipStream_t st) {
  if (n <= 0 || lda < n || ldm < n || !A || !M || !b || !x || !iters || !workspace ||
      max_iter < 1 || max_iter > FJ_REFINE_MAX_ITER)
    return FJ_ERR_ARG;
  tol = fj_refine_tol(tol, n);
  const int rgrid = fj_cdiv(n, FJ_RV_WAVES);
  return fj_refine_run(A, lda, n, b, x, max_iter, tol, iters, workspace, st,
                       [&](bool first, int k, const FjRefineWs& w) {
                         if (first)
                           hipLaunchKernelGGL((k_refine_apply_inv<T, false>), dim3(rgrid),
                                              dim3(64 * FJ_RV_WAVES), 0, st, M, ldm, n, b, x,
                                              w.slots, 0, tol, w.flags);
                         else
                           hipLaunchKernelGGL((k_refine_apply_inv<T, true>), dim3(rgrid),
                                              dim3(64 * FJ_RV_WAVES), 0, st, M, ldm, n, w.r, x,
                                              w.slots, k, tol, w.flags);
                       });
}

extern "C" int fj_solve_refined_inv(const double* A, int lda, const double* Ainv, int ldi, int n,
                                    const double* b, double* x, int max_iter, double tol,
                                    int32_t* iters, void* workspace, void* stream) {
  return fj_solve_refined_inv_t<double>(A, lda, Ainv, ldi, n, b, x, max_iter, tol, iters,
                                        workspace, (hipStream_t)stream);
}

extern "C" int fj_solve_refined_inv_f32(const double* A, int lda, const float* Ainv, int ldi,
                                        int n, const double* b, double* x, int max_iter,
                                        double tol, int32_t* iters, void* workspace,
                                        void* stream) {
  return fj_solve_refined_inv_t<float>(A, lda, Ainv, ldi, n, b, x, max_iter, tol, iters,
                                       workspace, (hipStream_t)stream);
}

// ---- explicit inverse from the factor ------------------------------------------
// A^{-1} = U^{-1} L^{-1} P, right-looking over 64-row block rows, in place in
// X (initialised to P: row i has its unit at column ipiv[n + i]):
//   forward,  k = 0 .. nb-1:  X_k <- Linv_kk X_k;  X_j -= L_jk X_k  (j > k)
//   backward, k = nb-1 .. 0:  X_k <- Uinv_kk X_k;  X_j -= U_jk X_k  (j < k)
// Linv_kk / Uinv_kk are the factor's diagonal-block inverses.  Each step is
// two launches: the 64 x n diagonal product (one workgroup per 64-column
// tile) and the update of every other block row in one launch (one
// workgroup per 64 x 64 tile), so each update is a wide matrix-core launch
// with K = 64 instead of a long K loop in a few workgroups.
__global__ void k_inv_perm_init_f64(double* __restrict__ X, int ldx, int n,
                                    const int32_t* __restrict__ ipiv) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * n) return;
  const int i = (int)(e / n), j = (int)(e % n);
  X[(size_t)i * ldx + j] = (ipiv[n + i] == j) ? 1.0 : 0.0;
}
__global__ void k_inv_perm_init_f32(float* __restrict__ X, int ldx, int n,
                                    const int32_t* __restrict__ ipiv) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * n) return;
  const int i = (int)(e / n), j = (int)(e % n);
  X[(size_t)i * ldx + j] = (ipiv[n + i] == j) ? 1.0f : 0.0f;
}

// C (64 x 64 tile in LDS-free registers) = Ablk (64 x K, row stride lda) *
// Bblk (K x 64, row stride ldb), K <= 64, staged through LDS in halves of 32;
// rows/columns beyond (ra, cb) and K beyond kk read as zero.
template <typename T>
__device__ __forceinline__ void fj_tile_mma(const T* __restrict__ Ab, int lda, int ra,
                                            const T* __restrict__ Bb, int ldb, int cb, int kk,
                                            T (*sA)[FJ_GK + 1], T (*sB)[FJ_GT + 1],
                                            typename FjMfma<T>::acc_t (&acc)[2][2]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  for (int q0 = 0; q0 < kk; q0 += FJ_GK) {
    const int kq = min(FJ_GK, kk - q0);
    for (int e = tid; e < FJ_GT * FJ_GK; e += 256) {
      const int i = e / FJ_GK, q = e % FJ_GK;
      sA[i][q] = (i < ra && q < kq) ? Ab[(size_t)i * lda + q0 + q] : T(0);
    }
    for (int e = tid; e < FJ_GK * FJ_GT; e += 256) {
      const int q = e / FJ_GT, j = e % FJ_GT;
      sB[q][j] = (q < kq && j < cb) ? Bb[(size_t)(q0 + q) * ldb + j] : T(0);
    }
    __syncthreads();
    for (int q = 0; q < kq; q += 4) {
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const T av = sA[wr + a * 16 + li][q + lk];
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[a][c] = FjMfma<T>::mma(av, sB[q + lk][wc + c * 16 + li], acc[a][c]);
      }
    }
    __syncthreads();
  }
}

// X_k (block row kb, this workgroup's 64-column tile) <- D X_k
template <typename T>
__global__ __launch_bounds__(256) void k_inv_diag(T* __restrict__ X, int ldx, int n,
                                                  const T* __restrict__ D, int kb) {
  __shared__ T sA[FJ_GT][FJ_GK + 1];
  __shared__ T sB[FJ_GK][FJ_GT + 1];
  typedef typename FjMfma<T>::acc_t acc_t;
  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = acc_t{0, 0, 0, 0};
  const int r0 = kb * FJ_TS_B, rs = min(FJ_TS_B, n - r0);
  const int c0 = blockIdx.x * FJ_GT, cb = min(FJ_GT, n - c0);
  fj_tile_mma<T>(D, FJ_TS_B, rs, X + (size_t)r0 * ldx + c0, ldx, cb, rs, sA, sB, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = wr + a * 16 + FjMfma<T>::row(lk, v), j = wc + c * 16 + li;
        if (i < rs && j < cb) X[(size_t)(r0 + i) * ldx + c0 + j] = acc[a][c][v];
      }
}

// X_j -= F_{j,kb} X_kb for the block rows j after kb (lower) or before it
// (upper); grid (column tiles, block rows updated)
template <typename T>
__global__ __launch_bounds__(256) void k_inv_update(const T* __restrict__ LU, int lda,
                                                    T* __restrict__ X, int ldx, int n, int kb,
                                                    int lower) {
  __shared__ T sA[FJ_GT][FJ_GK + 1];
  __shared__ T sB[FJ_GK][FJ_GT + 1];
  typedef typename FjMfma<T>::acc_t acc_t;
  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = acc_t{0, 0, 0, 0};
  const int jblk = lower ? kb + 1 + (int)blockIdx.y : (int)blockIdx.y;
  const int r0 = jblk * FJ_TS_B, rs = min(FJ_TS_B, n - r0);
  const int k0 = kb * FJ_TS_B, ks = min(FJ_TS_B, n - k0);
  const int c0 = blockIdx.x * FJ_GT, cb = min(FJ_GT, n - c0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  // the lane's entries of X_j, loaded before the products (X_j is another
  // block row than the operand X_kb; the update is the same x - acc)
  T xv[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = wr + a * 16 + FjMfma<T>::row(lk, v), j = wc + c * 16 + li;
        xv[a][c][v] = (i < rs && j < cb) ? X[(size_t)(r0 + i) * ldx + c0 + j] : T(0);
      }
  fj_tile_mma<T>(LU + (size_t)r0 * lda + k0, lda, rs, X + (size_t)k0 * ldx + c0, ldx, cb, ks, sA,
                 sB, acc);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = wr + a * 16 + FjMfma<T>::row(lk, v), j = wc + c * 16 + li;
        if (i < rs && j < cb) X[(size_t)(r0 + i) * ldx + c0 + j] = xv[a][c][v] - acc[a][c][v];
      }
}

template <typename T>
static void fj_inv_perm_init(T* X, int ldx, int n, const int32_t* ipiv, hipStream_t st);
template <>
void fj_inv_perm_init<double>(double* X, int ldx, int n, const int32_t* ipiv, hipStream_t st) {
  hipLaunchKernelGGL(k_inv_perm_init_f64, dim3((unsigned)(((size_t)n * n + 255) / 256)), dim3(256),
                     0, st, X, ldx, n, ipiv);
}
template <>
void fj_inv_perm_init<float>(float* X, int ldx, int n, const int32_t* ipiv, hipStream_t st) {
  hipLaunchKernelGGL(k_inv_perm_init_f32, dim3((unsigned)(((size_t)n * n + 255) / 256)), dim3(256),
                     0, st, X, ldx, n, ipiv);
}

template <typename T>
static int fj_lu_inverse_t(const T* LU, int n, int lda, const int32_t* ipiv, const T* aux, T* X,
                           int ldx, hipStream_t st) {
  if (n <= 0 || n > FJ_SOLVE_MAXN || lda < n || ldx < n || !LU || !ipiv || !aux || !X)
    return FJ_ERR_ARG;
  const int nblk = fj_nblk(n);
  const int tiles = fj_cdiv(n, FJ_GT);
  const T* Linv = aux;
  const T* Uinv = aux + (size_t)nblk * FJ_TS_B * FJ_TS_B;
  fj_inv_perm_init<T>(X, ldx, n, ipiv, st);
  FJ_CHECK_LAUNCH();
  for (int kb = 0; kb < nblk; ++kb) {
    hipLaunchKernelGGL(k_inv_diag<T>, dim3(tiles), dim3(256), 0, st, X, ldx, n,
                       Linv + (size_t)kb * FJ_TS_B * FJ_TS_B, kb);
    FJ_CHECK_LAUNCH();
    if (nblk - kb - 1 > 0) {
      hipLaunchKernelGGL(k_inv_update<T>, dim3(tiles, nblk - kb - 1), dim3(256), 0, st, LU, lda, X,
                         ldx, n, kb, 1);
      FJ_CHECK_LAUNCH();
    }
  }
  for (int kb = nblk - 1; kb >= 0; --kb) {
    hipLaunchKernelGGL(k_inv_diag<T>, dim3(tiles), dim3(256), 0, st, X, ldx, n,
                       Uinv + (size_t)kb * FJ_TS_B * FJ_TS_B, kb);
    FJ_CHECK_LAUNCH();
    if (kb > 0) {
      hipLaunchKernelGGL(k_inv_update<T>, dim3(tiles, kb), dim3(256), 0, st, LU, lda, X, ldx, n, kb,
                         0);
      FJ_CHECK_LAUNCH();
    }
  }
  return FJ_OK;
}

extern "C" int fj_lu_inverse(const double* LU, int n, int lda, const int32_t* ipiv,
                             const void* aux, double* Ainv, int ldi, void* stream) {
  return fj_lu_inverse_t<double>(LU, n, lda, ipiv, (const double*)aux, Ainv, ldi,
                                 (hipStream_t)stream);
}

extern "C" int fj_lu_inverse_f32(const float* LU, int n, int lda, const int32_t* ipiv,
                                 const void* aux, float* Ainv, int ldi, void* stream) {
  return fj_lu_inverse_t<float>(LU, n, lda, ipiv, (const float*)aux, Ainv, ldi,
                                (hipStream_t)stream);
}

// ---- Banded solve (the beam's Newton systems) ---------------------------------
// A x = b for a matrix whose nonzeros lie within kl sub- and ku super-diagonals
// (a chain of two-node beam elements couples each node's 6 DOFs with its
// neighbours' only: kl = ku = 11), read from a dense row-major array.  The
// arithmetic is LAPACK's banded LU with partial pivoting (dgbtf2: first
// maximal |a| of the column, scaling by the reciprocal pivot, rank-1 update of
// the band, fill confined to kl extra super-diagonals) followed by dgbtrs's
// substitutions (row exchanges and L applied to b column by column, then the
// column-oriented upper-triangular sweep of dtbsv).  Partial pivoting on a
// band picks the same rows as on the dense matrix, because the entries of a
// column outside the band are zero.
// One wave does the whole solve with the band in LDS (row i holds columns
// i - kl .. i + kl + ku), so each of the n sequential pivot steps is a few
// LDS round trips instead of a dense panel launch.
#define FJ_BAND_LDS_DOUBLES 19968  // 156 KiB of the 160 KiB LDS

template <int KLMAX>
__global__ __launch_bounds__(64) void k_band_solve(const double* __restrict__ A, int n, int lda,
                                                   int kl, int ku, const double* __restrict__ b,
                                                   double* __restrict__ x,
                                                   int32_t* __restrict__ info,
                                                   const int32_t* __restrict__ guard) {
  if (guard && guard[1] != 0) return;
  __shared__ double s[FJ_BAND_LDS_DOUBLES];
  const int lane = threadIdx.x;
  const int W = 2 * kl + ku + 1;
  double* xs = s + (size_t)n * W;
  // band -> LDS in batches of 32 independent loads per lane (one wave has
  // to keep many loads in flight to hide the memory latency)
  const int total = n * W;
  for (int base = 0; base < total; base += 64 * 32) {
    double v[32];
#pragma unroll
    for (int u = 0; u < 32; ++u) {
      const int q = base + 64 * u + lane;
      const int i = q / W, t = q - i * W;
      const int c = i - kl + t;
      v[u] = (q < total && c >= 0 && c < n && t <= kl + ku) ? A[(size_t)i * lda + c] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 32; ++u) {
      const int q = base + 64 * u + lane;
      if (q < total) s[q] = v[u];
    }
  }
  for (int i = lane; i < n; i += 64) xs[i] = b[i];
  int first_zero = 0;
  int ju = 0;  // last column reached by fill so far (dgbtf2's JU)
  __syncthreads();
  // Each column step is written without data-dependent branches (clamped
  // addresses and selects): a taken branch costs a single wave more than
  // the arithmetic of a whole step.
  for (int k = 0; k < n; ++k) {
    const int km = min(kl, n - 1 - k);
    // column k, rows k .. k + km, read by every lane (broadcast reads; rows
    // past k + km read row k + km again and are ignored)
    double a[KLMAX + 1];
#pragma unroll
    for (int i = 0; i <= KLMAX; ++i) {
      const int ii = min(i, km);
      a[i] = s[(size_t)(k + ii) * W + (kl - ii)];
    }
    // pivot: the first largest |a| (idamax)
    double amax = fabs(a[0]);
    int ip = 0;
#pragma unroll
    for (int i = 1; i <= KLMAX; ++i) {
      const double t = fabs(a[i]);
      const bool take = i <= km && t > amax;
      amax = take ? t : amax;
      ip = take ? i : ip;
    }
    if (amax != 0.0) {
      const int p = k + ip;
      ju = max(ju, min(p + ku, n - 1));
      const int nj = ju - k;
      // rows k and p exchanged: multipliers l_i = a'_i / a'_0 (dgbtf2 scales
      // by the reciprocal pivot); l_i for i > km repeats l_km
      double l[KLMAX + 1];
      double piv = a[0];
#pragma unroll
      for (int i = 1; i <= KLMAX; ++i) piv = i == ip ? a[i] : piv;
      const double rp = 1.0 / piv;
      double lkm = 0.0, lme = 0.0;
#pragma unroll
      for (int i = 1; i <= KLMAX; ++i) {
        l[i] = (i == ip ? a[0] : a[i]) * rp;
        lkm = i == km ? l[i] : lkm;
        lme = i == lane ? l[i] : lme;
      }
#pragma unroll
      for (int i = 1; i <= KLMAX; ++i) l[i] = i <= km ? l[i] : lkm;
      // each lane owns columns k + j: exchange, then the rank-1 update of
      // rows k + 1 .. k + km (all reads first, then the writes; rows past
      // k + km rewrite row k + km with the same value)
      if (km > 0) {
        for (int j = 1 + lane; j <= nj; j += 64) {
          double* uk = s + (size_t)k * W + (j + kl);
          double* up = s + (size_t)p * W + (j - ip + kl);
          const double u = *up;
          const double t = *uk;
          *uk = u;
          *up = t;
          double r[KLMAX + 1];
#pragma unroll
          for (int i = 1; i <= KLMAX; ++i) {
            const int ii = min(i, km);
            r[i] = s[(size_t)(k + ii) * W + (j - ii + kl)];
          }
#pragma unroll
          for (int i = 1; i <= KLMAX; ++i) {
            const int ii = min(i, km);
            s[(size_t)(k + ii) * W + (j - ii + kl)] = r[i] - l[i] * u;
          }
        }
      }  // km == 0: no rows below, so p == k and there is nothing to update
      // column k: pivot and multipliers; b: exchange and update
      const double bk = xs[p];
      const bool mine = lane >= 1 && lane <= km;
      const int rl = mine ? k + lane : k;
      const double bi = (mine && lane == ip) ? xs[k] : xs[rl];
      if (lane == 0) {
        s[(size_t)k * W + kl] = piv;
        xs[k] = bk;
      }
      if (mine) {
        s[(size_t)rl * W + (kl - lane)] = lme;
        xs[rl] = bi - lme * bk;
      }
    } else if (first_zero == 0) {
      first_zero = k + 1;
    }
    __syncthreads();
  }
  // U x = y, column by column (dtbsv, upper, no transpose)
  for (int j = n - 1; j >= 0; --j) {
    const double xj = xs[j];
    double v = xj;
    if (xj != 0.0) v = xj / s[(size_t)j * W + kl];
    __syncthreads();
    if (xj != 0.0) {
      const int kk = min(kl + ku, j);
      for (int i = 1 + lane; i <= kk; i += 64) xs[j - i] -= v * s[(size_t)(j - i) * W + (i + kl)];
      if (lane == 0) xs[j] = v;
    }
    __syncthreads();
  }
  for (int i = lane; i < n; i += 64) x[i] = xs[i];
  if (lane == 0 && info) *info = first_zero;
}

// First maximal |a| of a column (idamax) by a tree: the left operand always
// holds the lower rows and the right one wins only when strictly larger.
template <int KL>
struct FjBandPick {
  // first maximal |a[lo .. lo + cnt)|: signed value and offset
  template <int LO, int CNT>
  static __device__ __forceinline__ void run(const double (&a)[KL + 1], double& v, int& ix) {
    if constexpr (CNT == 1) {
      v = a[LO];
      ix = LO;
    } else {
      constexpr int H = (CNT + 1) / 2;
      double v1, v2;
      int i1, i2;
      run<LO, H>(a, v1, i1);
      run<LO + H, CNT - H>(a, v2, i2);
      const bool t = fabs(v2) > fabs(v1);
      v = t ? v2 : v1;
      ix = t ? i2 : i1;
    }
  }
};

__device__ __forceinline__ double fj_readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Ordering of LDS accesses between the lanes of one wave: a wave's LDS
// instructions execute in issue order, so a later read sees an earlier write
// of any lane of the same wave once the compiler keeps the order (a
// wavefront-scope fence; no wait and no barrier instruction is needed).
__device__ __forceinline__ void fj_wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// The same solve with the bandwidths fixed at compile time (kl = ku = 11:
// any chain of two-node beam elements), spread over three waves.  KL zero
// rows below the matrix let every column step read and update rows
// k .. k + KL unconditionally (a zero row is never chosen as a pivot and
// never reaches the solution).
//
// Elimination: 16-lane groups, group g < KL owning columns k + 1 + g and
// k + 12 + g of the step's rank-1 update (lane i: row k + i; lane 0 the
// exchanged row k), group KL the right-hand side; one wave per SIMD, since a
// wave sharing its SIMD reached each barrier 100-200 cycles late.  A column
// step is one barrier: each lane reads its entry of column k, and a 16-lane
// DPP maximum of the high words of |a| below the diagonal tells whether the
// diagonal is idamax's choice (every entry below it smaller in magnitude);
// then, with the reciprocal of the diagonal issued meanwhile, the rank-1
// update is one FMA per lane and column and one store.  A tie or a larger
// high word (NaN included, and a zero diagonal) takes the exact search (the
// tree above over the whole column) and the general exchange.  A group lies
// inside one wave and within a step only the group of a column reads or
// writes that column, so the reads of the pivot row before its rows are
// rewritten need no barrier.
//
// Substitution: one wave with x in registers (lane l, register q: row
// 64 q + l); each row is two lane reads, a division and one FMA per lane,
// with U's entries read one pointer step a row.  The arithmetic is dgbtf2's
// and dgbtrs's to the bit: the same pivot rows, the same scaling by the
// reciprocal pivot, the same FMA per entry in the same order, dtbsv's
// division per row.
#define FJ_BAND_FIXED_GROUPS 12
#define FJ_BAND_FIXED_THREADS (16 * FJ_BAND_FIXED_GROUPS)

template <int KL, int KU>
__global__ __launch_bounds__(FJ_BAND_FIXED_THREADS) void k_band_solve_fixed(const double* __restrict__ A, int n,
                                                                      int lda, const double* __restrict__ b,
                                                                      double* __restrict__ x,
                                                                      int32_t* __restrict__ info,
                                                                      const int32_t* __restrict__ guard) {
  if (guard && guard[1] != 0) return;
  constexpr int W = 2 * KL + KU + 1;
  constexpr int D = W - 1;
  constexpr int JM = KL + KU;
  static_assert(KL == KU && KL + 1 == FJ_BAND_FIXED_GROUPS && KL < 16, "band shape");
  constexpr int QMAX = (FJ_BAND_LDS_DOUBLES / (W + 2) + 63) / 64;  // x registers per lane
  __shared__ double s[FJ_BAND_LDS_DOUBLES];
  // a lane's slot for the stores it does not own in a column step, so that
  // every lane stores unconditionally (no divergent branch around a store)
  __shared__ double sink[FJ_BAND_FIXED_THREADS];
  const int tid = threadIdx.x;
  const int rows = n + KL;
  double* xs = s + (size_t)rows * W;  // b, then y = L^-1 P b
  double* dg = xs + rows;             // U's diagonal
  const int total = rows * W;
  for (int base = 0; base < total; base += FJ_BAND_FIXED_THREADS * 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int q = base + FJ_BAND_FIXED_THREADS * u + tid;
      const int i = q / W, t = q - i * W;
      const int c = i - KL + t;
      v[u] = (q < total && i < n && c >= 0 && c < n && t <= KL + KU) ? A[(size_t)i * lda + c] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int q = base + FJ_BAND_FIXED_THREADS * u + tid;
      if (q < total) s[q] = v[u];
    }
  }
  for (int i = tid; i < rows; i += FJ_BAND_FIXED_THREADS) xs[i] = i < n ? b[i] : 0.0;
  const int g = tid >> 4, i = tid & 15;
  const bool act = i <= KL;        // lanes 0 .. KL of a group: rows k .. k + KL
  const int ic = act ? i : 0;      // idle lanes read row k and write nothing
  const bool updg = g < KL, bgrp = g == KL;
  const int j1 = g + 1, j2 = g + 1 + KL;  // update groups: columns k + j1 and k + j2
  const int c1 = updg ? j1 : 1, c2 = updg ? j2 : 1;
  // Each lane reads r and t and, without an exchange, stores r - l t over
  // r: update lanes (row k + i, column k + j) and (row k, column k + j) for
  // their two columns, b lanes y[k + i] and y[k]; the pointers advance by
  // rs doubles a column.
  double* const rb1 = updg ? s + KL + ic * D + c1 : (bgrp ? xs + ic : s + KL);
  double* const tb1 = updg ? s + KL + c1 : (bgrp ? xs : s + KL);
  double* const rb2 = updg ? s + KL + ic * D + c2 : s + KL;
  double* const tb2 = updg ? s + KL + c2 : s + KL;
  const int rs = bgrp ? 1 : W;
  const bool st1 = (updg && act && i != 0) || (bgrp && act), st2 = updg && act && i != 0;
  const bool b0 = bgrp && i == 0;  // stores U's diagonal instead
  const int jl1 = updg ? j1 : 0, jl2 = updg ? j2 : 0;
  int first_zero = 0;
  int ju = 0;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    double* colk = s + (size_t)k * W + KL;  // (row k, column k)
    double* const pr1 = rb1 + (size_t)k * rs;
    double* const pr2 = rb2 + (size_t)k * rs;
    const double a0 = colk[0];              // (row k, column k)
    const double al = colk[ic * D];         // (row k + i, column k)
    const double r1 = *pr1, t1 = tb1[(size_t)k * rs];
    const double r2 = *pr2, t2 = tb2[(size_t)k * rs];
    // The diagonal is idamax's choice when every entry below it is smaller
    // in magnitude, which holds when its high word (sign cleared) is
    // smaller: lane i of every group holds entry i, so a 16-lane DPP
    // maximum of those words (entries 1 .. KL; the rest count 0) decides
    // it, and the same in every group.  A tie or a larger word (NaN
    // included, and a zero diagonal) takes the exact search over the
    // column.  The reciprocal of the diagonal is issued meanwhile.
    const unsigned hw = (unsigned)((unsigned long long)__double_as_longlong(al) >> 32) & 0x7FFFFFFFu;
    unsigned m = (i >= 1 && i <= KL) ? hw : 0u;
    m = fj_dpp_max_u32<0x111, 0xf>(m);  // row_shr:1 .. 8: lane 15 holds the row maximum
    m = fj_dpp_max_u32<0x112, 0xf>(m);
    m = fj_dpp_max_u32<0x114, 0xf>(m);
    m = fj_dpp_max_u32<0x118, 0xf>(m);
    const unsigned below = __builtin_amdgcn_readlane(m, 15);
    const unsigned h0 =
        __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)__double_as_longlong(a0) >> 32) & 0x7FFFFFFFu);
    double rp = 1.0 / a0;
    if (__builtin_expect(below < h0, 1)) {
      // no exchange: row k is the pivot row (t)
      ju = max(ju, min(k + KU, n - 1));
      const int nj = ju - k;
      const double li = al * rp;
      const double v1 = r1 - li * t1, v2 = r2 - li * t2;
      double* const d1 = (st1 && jl1 <= nj) ? (b0 ? dg + k : pr1) : sink + tid;
      double* const d2 = (st2 && jl2 <= nj) ? pr2 : sink + tid;
      *d1 = b0 ? a0 : v1;
      *d2 = v2;
    } else {
      double a[KL + 1];
#pragma unroll
      for (int q = 0; q <= KL; ++q) a[q] = colk[q * D];
      double piv;
      int ipv;
      FjBandPick<KL>::template run<0, KL + 1>(a, piv, ipv);
      const int ip = __builtin_amdgcn_readfirstlane(ipv);
      if (piv != 0.0) {
        if (ip != 0) rp = 1.0 / piv;
        const int p = k + ip;
        ju = max(ju, min(p + KU, n - 1));
        const int nj = ju - k;
        const double li = ((i == ip) ? a0 : al) * rp;
        if (updg && act) {
          // rows k and p exchanged in the lane's columns, then the update
          const double u1 = colk[ip * D + j1], u2 = colk[ip * D + j2];  // (row p, column k + j)
          fj_wave_lds_order();
          if (i == 0) {
            if (ip != 0) {
              if (j1 <= nj) colk[j1] = u1;
              if (j2 <= nj) colk[j2] = u2;
            }
          } else {
            if (j1 <= nj) colk[i * D + j1] = ((i == ip) ? t1 : r1) - li * u1;
            if (j2 <= nj) colk[i * D + j2] = ((i == ip) ? t2 : r2) - li * u2;
          }
        } else if (bgrp && act) {
          const double bk = xs[p];
          fj_wave_lds_order();
          if (i == 0) {
            xs[k] = bk;
            dg[k] = piv;
          } else {
            xs[k + i] = ((i == ip) ? t1 : r1) - li * bk;
          }
        }
      } else {
        if (first_zero == 0) first_zero = k + 1;
        if (b0) dg[k] = 0.0;
      }
    }
    __syncthreads();
  }
  if (tid >= 64) return;
  // U x = y (dtbsv, upper, no transpose), rows from the last
  const int lane = tid;
  double y[QMAX];
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    const int row = 64 * q + lane;
    y[q] = row < n ? xs[row] : 0.0;
  }
#pragma unroll
  for (int q = QMAX - 1; q >= 0; --q) {
    if (64 * q >= n) continue;
    const int rq = 64 * q + lane;  // this lane's row in register q (and rq - 64 in q - 1)
    const double dq = dg[rq < n ? rq : 0];  // the block's diagonal, read by lane
    // U's entries (row rq, column jj) and (row rq - 64, column jj) sit at
    // rq D + KL + jj and rq D + KL + jj - 64 D: a pointer moving one double
    // a row (every address stays inside the LDS array; an entry outside the
    // band is read and not used)
    const double* p0 = s + (size_t)rq * D + KL;
    const double* p1 = s + (size_t)(q > 0 ? rq - 64 : 0) * D + KL;
    for (int jj = min(n - 1, 64 * q + 63); jj >= 64 * q; --jj) {
      const int d0 = jj - rq;  // column offset of jj from row rq
      const int d1 = d0 + 64;  // ... and from row rq - 64
      const bool u0 = d0 >= 1 && d0 <= JM, u1 = q > 0 && d1 >= 1 && d1 <= JM;
      const double a0 = p0[jj];
      const double a1 = p1[jj];
      const double dj = fj_readlane_f64(dq, jj - 64 * q);
      const double yj = fj_readlane_f64(y[q], jj - 64 * q);
      // dtbsv leaves a zero y_j and its column alone; the step is written
      // without a branch (selects) so that the reads above are issued
      // before the division
      const bool nz = yj != 0.0;
      const double xj = yj / dj;
      y[q] = (d0 == 0) ? (nz ? xj : yj) : ((nz && u0) ? y[q] - xj * a0 : y[q]);
      if (q > 0) y[q - 1] = (nz && u1) ? y[q - 1] - xj * a1 : y[q - 1];
    }
  }
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    const int row = 64 * q + lane;
    if (row < n) x[row] = y[q];
  }
  if (lane == 0 && info) *info = first_zero;
}

extern "C" size_t fj_band_solve_capacity(void) { return FJ_BAND_LDS_DOUBLES; }

// The band solve launch; a non-NULL guard (a Newton status word) skips it
// once the loop it belongs to has finished.
int fj_band_launch(const double* A, int n, int lda, int kl, int ku, const double* b, double* x, int32_t* info,
                   const int32_t* guard, hipStream_t st) {
  if (n <= 0 || lda < n || kl < 0 || ku < 0 || kl > 63 || !A || !b || !x) return FJ_ERR_ARG;
  if ((size_t)n * (size_t)(2 * kl + ku + 2) > (size_t)FJ_BAND_LDS_DOUBLES) return FJ_ERR_ARG;
  if (kl == 11 && ku == 11 && (size_t)(n + 11) * 36 <= (size_t)FJ_BAND_LDS_DOUBLES)
    hipLaunchKernelGGL((k_band_solve_fixed<11, 11>), dim3(1), dim3(FJ_BAND_FIXED_THREADS), 0, st, A, n, lda, b, x, info,
                       guard);
  else if (kl <= 15)
    hipLaunchKernelGGL(k_band_solve<15>, dim3(1), dim3(64), 0, st, A, n, lda, kl, ku, b, x, info, guard);
  else
    hipLaunchKernelGGL(k_band_solve<63>, dim3(1), dim3(64), 0, st, A, n, lda, kl, ku, b, x, info, guard);
  FJ_CHECK_LAUNCH();
  return FJ_OK;
}

extern "C" int fj_band_solve(const double* A, int n, int lda, int kl, int ku, const double* b,
                             double* x, int32_t* info, void* stream) {
  return fj_band_launch(A, n, lda, kl, ku, b, x, info, nullptr, (hipStream_t)stream);
}
