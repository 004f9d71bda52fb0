/* flapjax_amd: C-ABI of the MI355X-native UVLM aerodynamic hot path.
 *
 * Every entry point takes caller-owned DEVICE pointers (FP64 unless stated),
 * plain sizes, and an explicit HIP stream (hipStream_t passed as void*; NULL
 * is the default stream).  Work is enqueued asynchronously on that stream.
 * The return value is a status code: 0 on success, a negative FJ_ERR_* value
 * otherwise (nothing is enqueued when an argument check fails).
 *
 * Layouts (the reference's, aero/uvlm.py:623-640 and aero/aic.py):
 *   vertex grid      (m+1, n+1, 3) row-major, xyz innermost
 *   ring strengths   (m, n) row-major
 *   points / vectors (T, 3)
 *   AIC              (rows, ld) row-major, rows = targets, columns = rings
 *   segment table    FJ_SEG_STRIDE doubles per filament (see fj_segments_from_grid)
 *
 * Reference paths below are relative to src/flapjax/ of ben-l-p/flapjax v1.2.1.
 */
#ifndef FLAPJAX_AMD_H
#define FLAPJAX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FJ_OK 0
#define FJ_ERR_ARG (-1)      /* bad size / null pointer / unknown enum */
#define FJ_ERR_LAUNCH (-2)   /* a kernel launch failed */
#define FJ_ERR_ALLOC (-3)    /* workspace allocation failed */
#define FJ_ERR_SINGULAR (-4) /* zero pivot met in the LU factorisation */
#define FJ_ERR_NOCONV (-5)   /* iterative refinement did not converge */
#define FJ_ERR_SCRATCH (-6)  /* the call needs more than the library's fixed scratch memory */

/* Size of the fixed device scratch array of the allocation-free entry point
 * fj_pair_sum_lattices_noalloc (one copy per GPU, held until the process ends). */
#define FJ_NOALLOC_SCRATCH_BYTES (32u << 20)

/* Biot-Savart kernel kinds: the three KernelFunction implementations the
 * reference ships (aero/utils.py:19 type, :580-648 implementations). */
#define FJ_KERNEL_EPSILON 0 /* biot_savart_epsilon, aero/utils.py:606-623 (default, aero/uvlm.py:280-285) */
#define FJ_KERNEL_CUTOFF 1  /* biot_savart_cutoff,  aero/utils.py:626-648 */
#define FJ_KERNEL_PLAIN 2   /* biot_savart,         aero/utils.py:580-592 */

/* Arithmetic forms of the pair sums */
#define FJ_PREC_FP64 0  /* plain FP64 drop-in */
#define FJ_PREC_MIXED 1 /* FP32 far field, FP64 near field, FP64 accumulation */

/* Epilogues of a pair sum (what is done with v at target t) */
#define FJ_OUT_STORE 0   /* out[t] = v */
#define FJ_OUT_ADD 1     /* out[t] += v */
#define FJ_OUT_CONVECT 2 /* out[t] = x[t] + dt * (U(x[t], time) + v): wake convection,
                            aero/utils.py:445-450 with the free-wake closure aero/uvlm.py:917-929 */

/* Flow-field kinds, aero/flowfields.py */
#define FJ_FLOW_CONSTANT 0          /* Constant, aero/flowfields.py:88-97 */
#define FJ_FLOW_ONE_MINUS_COSINE 1  /* OneMinusCosine, aero/flowfields.py:101-200 */

#define FJ_SEG_STRIDE_DOUBLES 16

typedef struct fj_flowfield {
  int kind;
  int relative_motion;
  double u_inf[3];
  double gust_length;
  double gust_amplitude;
  double gust_travel_dir[3];    /* unit */
  double gust_amplitude_dir[3]; /* unit */
  double gust_x0[3];
} fj_flowfield;

/* ---- library ----------------------------------------------------------- */

/* Version string of the library build. */
const char* fj_version(void);

/* Number of bytes of device workspace fj_pair_sum needs for n_tgt targets and
 * n_seg segments (caller allocates; 256-byte alignment). */
size_t fj_pair_sum_workspace_bytes(int n_tgt, int n_seg, int precision);

/* ---- K1: de-duplicated filaments and pair sums --------------------------- */

/* Unique filaments of one vertex grid with their effective circulations,
 * replacing the segment construction of compute_v_ind (aero/aic.py:277-287):
 * chordwise filaments (m, n+1) first, then spanwise (m+1, n), each row-major.
 * gamma may be NULL (all circulations zero; geometry only).
 * If mirror_point/mirror_normal (HOST pointers, 3 doubles each, normal unit)
 * are non-NULL the grid is first reflected (aero/utils.py:651-663) and the
 * circulations negated, which reproduces the subtraction at aero/aic.py:293-308.
 * seg_out receives (m*(n+1) + (m+1)*n) records of FJ_SEG_STRIDE_DOUBLES. */
int fj_segments_from_grid(const double* zeta, const double* gamma, int m, int n,
                          const double* mirror_point, const double* mirror_normal,
                          int kernel, double eps, double* seg_out, void* stream);

/* One source grid of fj_segments_from_grids: device vertex grid (m+1, n+1, 3),
 * ring strengths (m, n) or NULL, and whether to take its mirror image. */
typedef struct {
  const double* zeta;
  const double* gamma;
  int m;
  int n;
  int mirror;
} fj_grid_ref;

/* The tables of several grids (host array `grids`) concatenated in list order
 * into seg_out, in as few launches as possible (one per 16 grids): the table
 * of all source surfaces that compute_v_ind sums over (aero/aic.py:276-308,
 * bound and wake surfaces, each followed by its mirror image when mirrored).
 * mirror_point/mirror_normal (HOST, 3 doubles) are required if any grid has
 * mirror set. */
int fj_segments_from_grids(int n_grids, const fj_grid_ref* grids, const double* mirror_point,
                           const double* mirror_normal, int kernel, double eps, double* seg_out,
                           void* stream);

/* Filament table from end points ends (n_seg, 2, 3) and circulations gamma
 * (n_seg; NULL = unit), the filament-grid input of v_ind_vmap
 * (aero/aic.py:208-245). */
int fj_segments_from_endpoints(const double* ends, const double* gamma, int n_seg, int kernel,
                               double eps, double* seg_out, void* stream);

/* v(x_t) = sum_s K(x_t; seg_s) for n_tgt targets against n_seg filaments:
 * the contraction done by v_ind_vmap (aero/aic.py:208-245) and summed by
 * compute_v_ind (aero/aic.py:248-310).  The epilogue `mode` decides what is
 * written to out (n_tgt, 3); for FJ_OUT_CONVECT, `flow`, `time` and `dt`
 * give U(x, time) and the step; otherwise they are ignored (flow may be NULL).
 * workspace: fj_pair_sum_workspace_bytes(n_tgt, n_seg, precision) bytes. */
int fj_pair_sum(const double* seg, int n_seg, const double* targets, int n_tgt, double* out,
                int kernel, double eps, double rcut, int precision, int mode,
                const fj_flowfield* flow, double time, double dt, void* workspace,
                void* stream);

/* fj_pair_sum over two target ranges in one pair-kernel launch: targets
 * [0, n_tgt_a) get epilogue mode_a into out_a, targets [n_tgt_a, n_tgt)
 * mode_b into out_b; flow, time and dt serve whichever range uses
 * FJ_OUT_CONVECT.  The filament-table counterpart of
 * fj_pair_sum_lattices_split (the UVLM step's forces at step n and rollup
 * at step n + 1 in the mixed form: aero/uvlm.py:1018-1050 and 912-941).
 * Workspace as fj_pair_sum for n_tgt. */
int fj_pair_sum_split(const double* seg, int n_seg, const double* targets, int n_tgt, int n_tgt_a,
                      double* out_a, int mode_a, double* out_b, int mode_b, int kernel, double eps,
                      double rcut, int precision, const fj_flowfield* flow, double time, double dt,
                      void* workspace, void* stream);

/* One vortex lattice inside a filament table: the S(m, n) = m(n+1) + (m+1)n
 * records written for an (m+1, n+1) vertex grid by fj_segments_from_grid(s)
 * (chordwise filaments first, then spanwise), starting at row seg_off. */
typedef struct {
  int seg_off;
  int m;
  int n;
} fj_lattice_ref;

/* fj_pair_sum (FP64) over the filaments of whole lattices -- the surfaces
 * that compute_v_ind sums over (aero/aic.py:248-310) -- with the per-vertex
 * part of biot_savart_epsilon / _cutoff / biot_savart (aero/utils.py:580-648:
 * r = x - y, 1/sqrt(|r|^2 + eps^2)) evaluated once per vertex and shared by
 * the filaments that meet there.  Same epilogue modes as fj_pair_sum.  At
 * most 16 lattices per call; lattices with m = 0 or n = 0 (all filament
 * strengths zero) contribute nothing.
 * workspace: fj_pair_sum_lattices_workspace_bytes(n_tgt, n_lat, lats) bytes. */
size_t fj_pair_sum_lattices_workspace_bytes(int n_tgt, int n_lat, const fj_lattice_ref* lats);
int fj_pair_sum_lattices(const double* seg, int n_lat, const fj_lattice_ref* lats,
                         const double* targets, int n_tgt, double* out, int kernel, double eps,
                         double rcut, int mode, const fj_flowfield* flow, double time, double dt,
                         void* workspace, void* stream);

/* fj_pair_sum_lattices over two target ranges that share one filament table:
 * targets [0, n_tgt_a) get epilogue mode_a into out_a ((n_tgt_a, 3)),
 * targets [n_tgt_a, n_tgt) mode_b into out_b ((n_tgt - n_tgt_a, 3)); flow,
 * time and dt serve whichever range uses FJ_OUT_CONVECT.  One launch of the
 * pair kernel for both.  No single counterpart in the reference: it fuses
 * the forces' induced velocity at the bound filament midpoints of step n
 * (aero/uvlm.py:1018-1050, compute_v_ind over bound and wake) with the
 * free-wake velocity of step n + 1 at the wake vertices
 * (aero/uvlm.py:912-941, propagate_wake), which sum over the same lattices
 * with the same circulations.  Workspace as fj_pair_sum_lattices for n_tgt. */
int fj_pair_sum_lattices_split(const double* seg, int n_lat, const fj_lattice_ref* lats,
                               const double* targets, int n_tgt, int n_tgt_a, double* out_a, int mode_a,
                               double* out_b, int mode_b, int kernel, double eps, double rcut,
                               const fj_flowfield* flow, double time, double dt, void* workspace,
                               void* stream);

/* fj_pair_sum_lattices without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  It
 * runs the same kernels and plan on a fixed file-scope device array of
 * FJ_NOALLOC_SCRATCH_BYTES and calls no allocating routine; a call whose
 * fj_pair_sum_lattices_workspace_bytes exceeds that size returns
 * FJ_ERR_SCRATCH before any device work.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (two streams, two host threads)
 * race.  Results equal fj_pair_sum_lattices' bit for bit. */
int fj_pair_sum_lattices_noalloc(const double* seg, int n_lat, const fj_lattice_ref* lats,
                                 const double* targets, int n_tgt, double* out, int kernel,
                                 double eps, double rcut, int mode, const fj_flowfield* flow,
                                 double time, double dt, void* stream);

/* Reverse mode of fj_pair_sum (FJ_OUT_STORE) for hand-written adjoints:
 * with upstream weights grad_v (n_tgt, 3),
 *   grad_targets (n_tgt, 3)   = dL/dx          (written, or added if accumulate)
 *   grad_seg     (n_seg, 7)   = [dL/dy0 (3), dL/dy1 (3), dL/dg]  (written)
 * Either output may be NULL.  No counterpart in the reference (which gets
 * this from jax.vjp of compute_v_ind, aero/aic.py:248-310). */
int fj_pair_sum_backward(const double* seg, int n_seg, const double* targets, int n_tgt,
                         const double* grad_v, int kernel, double eps, double rcut,
                         double* grad_targets, int accumulate, double* grad_seg, void* workspace,
                         void* stream);
/* workspace of fj_pair_sum_backward (may be NULL: dL/dx then takes one pass
 * of a thread per target, which leaves most of the chip idle for a few
 * thousand targets). */
size_t fj_pair_sum_backward_workspace_bytes(int n_tgt, int n_seg);

/* Fused pair sums of many independent cases in one launch (serving loop).
 * Per task i (HOST arrays of n_tasks): targets rows [tgt_off, tgt_off+n_tgt)
 * of `targets`, filaments rows [seg_off, seg_off+n_seg) of `seg`, results to
 * rows [out_off, out_off+n_tgt) of `out` with epilogue mode[i] (dt, time,
 * flows used by FJ_OUT_CONVECT; may be NULL otherwise).  Tasks must be sorted
 * by tgt_off with disjoint target ranges.  task_table: device buffer of
 * fj_batched_task_bytes(n_tasks); host_staging: host buffer of the same size
 * (free to reuse once the call returns); workspace:
 * fj_batched_workspace_bytes(...).  No counterpart in the reference (which
 * vmaps/pmaps whole cases, tests/coupled/dynamic/test_parallel_cantilever.py). */
size_t fj_batched_task_bytes(int n_tasks);
size_t fj_batched_workspace_bytes(int n_tasks, const int* n_tgt, const int* n_seg);
int fj_pair_sum_batched(int n_tasks, const int* tgt_off, const int* n_tgt, const int* seg_off,
                        const int* n_seg, const int* out_off, const int* mode, const double* dt,
                        const double* time, const fj_flowfield* flows, const double* seg,
                        const double* targets, double* out, int kernel, double eps, double rcut,
                        void* task_table, void* host_staging, void* workspace, void* stream);

/* ---- K2: AIC assembly ----------------------------------------------------- */

/* Normal-projected ring influence of the grid zeta (m+1, n+1, 3) at n_tgt
 * collocation points with (non-unit) normals: compute_aic_grid
 * (aero/aic.py:11-63), row t, column i*n + j:
 *   A = -(M[i,j+1] - M[i,j]) + (N[i+1,j] - N[i,j]),  M, N = K . n.
 * Writes A_out[t*lda + col0 + i*n + j] (accumulate=0) or adds to it
 * (accumulate=1).  sign = +1 or -1 multiplies the block (mirror images are
 * subtracted, aero/aic.py:101-110).
 * If zeta_w is non-NULL (m_star >= 1 wake rows, (m_star+1, n+1, 3)), the
 * summed influence of each wake column is added onto the trailing-edge row of
 * the bound block (add_wake_influence, aero/aic.py:190-205; static solves). */
int fj_aic_assemble(const double* colloc, const double* normals, int n_tgt, const double* zeta,
                    int m, int n, const double* zeta_w, int m_star,
                    const double* mirror_point, const double* mirror_normal, int kernel,
                    double eps, double rcut, double* A_out, int lda, int col0, int accumulate,
                    void* stream);

/* ---- K3: circulation solve -------------------------------------------------
 * Replaces jnp.linalg.solve(aic, -v_bc_n) at aero/uvlm.py:985 (and :1634). */

/* Bytes of the auxiliary array that travels with a factor of order n: the
 * inverses of its 64x64 diagonal blocks (L unit-lower and U), used by every
 * solve with that factor. */
size_t fj_lu_aux_bytes(int n);

/* In-place LU with partial pivoting of a row-major n x n FP64 matrix
 * (P A = L U, unit-lower L), n <= 8192.  ipiv (device, 2n int32) receives the
 * 0-based pivot rows in LAPACK order, then the composed row permutation
 * (row i of P A is row ipiv[n + i] of A).  info (device, 1 int32) receives 0
 * or 1 + the first column without a non-zero pivot.  aux (device,
 * fj_lu_aux_bytes(n)) receives the diagonal-block inverses. */
int fj_lu_factor(double* A, int n, int lda, int32_t* ipiv, int32_t* info, void* aux,
                 void* stream);

/* Solve (trans = 0) A x = b or (trans = 1) A^T x = b in place in b (n, nrhs
 * column-major: right-hand side r at b + r*n) from fj_lu_factor's output. */
int fj_lu_solve(const double* LU, int n, int lda, const int32_t* ipiv, const void* aux,
                double* b, int nrhs, int trans, void* stream);

/* Banded solve of A x = b (the beam's Newton systems, which
 * jnp.linalg.solve solves at structure/beam.py:2273): A is read from a dense
 * row-major n x n array whose nonzeros lie within kl sub- and ku
 * super-diagonals (kl <= 63); LAPACK's banded LU with partial pivoting
 * (dgbtf2) and substitutions (dgbtrs), which pick the same pivot rows as the
 * dense LU.  A and b are not modified; x may alias b.  info (device, 1 int32,
 * may be NULL) receives 0 or 1 + the first column without a nonzero pivot.
 * Needs n (2 kl + ku + 2) <= fj_band_solve_capacity() doubles; otherwise
 * FJ_ERR_ARG (use fj_lu_factor). */
size_t fj_band_solve_capacity(void);
int fj_band_solve(const double* A, int n, int lda, int kl, int ku, const double* b, double* x,
                  int32_t* info, void* stream);

/* FP32 LU factor (row-major, in place; aux holds float block inverses). */
int fj_lu_factor_f32(float* A, int n, int lda, int32_t* ipiv, int32_t* info, void* aux,
                     void* stream);

/* Mixed-precision iterative refinement (LAPACK dsgesv's scheme): x0 = M b,
 * then r = b - A x in FP64 with the CURRENT FP64 A and x += M r, where M is an
 * approximation of A^{-1}: an LU factor of A or of a nearby matrix (e.g. the
 * AIC of an earlier step of a rigid-body motion, which changes by rounding
 * only), FP32 or FP64, or its explicit inverse.  Stop test (dsgesv's):
 * ||r||_inf <= tol ||A||_inf ||x||_inf, tol <= 0 meaning sqrt(n) * 2^-53.
 * At most max_iter (1..64) corrections; iters (device int32) receives the
 * corrections made, negated if the stop test never passed (also for NaN).
 * workspace: fj_refine_workspace_bytes(n).  All work stays on the stream; no
 * host synchronisation. */
size_t fj_refine_workspace_bytes(int n);

/* ... with the FP32 factor LU32 (ipiv, aux from fj_lu_factor_f32). */
int fj_solve_refined(const double* A, int lda, const float* LU32, int ldlu, const int32_t* ipiv,
                     const void* aux, int n, const double* b, double* x, int max_iter,
                     double tol, int32_t* iters, void* workspace, void* stream);

/* ... with an FP64 factor (fj_lu_factor). */
int fj_solve_refined_f64(const double* A, int lda, const double* LU, int ldlu,
                         const int32_t* ipiv, const void* aux, int n, const double* b, double* x,
                         int max_iter, double tol, int32_t* iters, void* workspace, void* stream);

/* Explicit inverse (row-major, leading dimension ldi) of the matrix whose
 * factor fj_lu_factor / fj_lu_factor_f32 produced: A^{-1} = U^{-1} L^{-1} P by
 * blocked triangular sweeps on the matrix cores.  Used as M above when one
 * factor serves many solves: each solve is then a few mat-vecs over all CUs. */
int fj_lu_inverse(const double* LU, int n, int lda, const int32_t* ipiv, const void* aux,
                  double* Ainv, int ldi, void* stream);
int fj_lu_inverse_f32(const float* LU, int n, int lda, const int32_t* ipiv, const void* aux,
                      float* Ainv, int ldi, void* stream);

/* Refinement (as above) with M = an explicit inverse, FP64 or FP32. */
int fj_solve_refined_inv(const double* A, int lda, const double* Ainv, int ldi, int n,
                         const double* b, double* x, int max_iter, double tol, int32_t* iters,
                         void* workspace, void* stream);
int fj_solve_refined_inv_f32(const double* A, int lda, const float* Ainv, int ldi, int n,
                             const double* b, double* x, int max_iter, double tol,
                             int32_t* iters, void* workspace, void* stream);

/* ---- K4-K8: stencils, convection, forces ------------------------------- */

/* zeta = R_j zeta0_ij + x_j and, if hg_dot/zeta_dot are non-NULL,
 * zeta_dot = W_j zeta0_ij + v_j (hg_to_zeta_b / hg_dot_to_zeta_b_dot,
 * aero/uvlm.py:537-621, se3.vect_product algebra/se3.py:209-216).  hg and
 * hg_dot are (n_nodes, 4, 4) row-major; dof_map (device, n+1 int32) selects
 * the beam node of each spanwise station. */
int fj_grid_from_hg(const double* hg, const double* hg_dot, const int32_t* dof_map,
                    const double* zeta0, int m, int n, double* zeta, double* zeta_dot,
                    void* stream);

/* Collocation points (4-vertex average, aero/utils.py:87-93), normals
 * (diag1 x diag2, aero/utils.py:96-105) and, if zeta_dot/c_dot non-NULL,
 * collocation velocities (aero/uvlm.py:888-893).  Any output may be NULL. */
int fj_colloc_normals(const double* zeta, const double* zeta_dot, int m, int n, double* c,
                      double* nc, double* c_dot, void* stream);

/* Filament midpoints of a vertex grid: chordwise (m, n+1) then spanwise
 * (m+1, n), the targets of calculate_steady_forcing (aero/utils.py:162-164). */
int fj_filament_midpoints(const double* zeta, int m, int n, double* mid, void* stream);

/* Everything a time step derives from the beam poses, in one launch (the
 * case_solve preamble, aero/uvlm.py:1162-1185, with the stencils of
 * base_solve_from_grid, :888-893, and the force midpoints, utils.py:162-164):
 * zeta_n from hg_n, zeta_nm1 from hg_nm1, zeta_dot_n from hg_n/hg_dot_n as in
 * fj_grid_from_hg; c, nc, c_dot as in fj_colloc_normals; mid as in
 * fj_filament_midpoints.  hg_nm1 / hg_dot_n and every output may be NULL.
 * Results are identical to those separate entry points. */
int fj_step_geometry(const double* hg_n, const double* hg_nm1, const double* hg_dot_n,
                     const int32_t* dof_map, const double* zeta0, int m, int n, double* zeta_n,
                     double* zeta_nm1, double* zeta_dot_n, double* c, double* nc, double* c_dot,
                     double* mid, void* stream);

/* rhs[t] = -(U(c_t, time) - c_dot_t + v_w_t) . n_t  (aero/uvlm.py:961-985);
 * c_dot and v_w may be NULL (static solve). */
int fj_rhs(const double* c, const double* nc, const double* c_dot, const double* v_w, int n_tgt,
           const fj_flowfield* flow, double time, double* rhs, void* stream);

/* Uniform-discretisation wake shift (aero/utils.py:427-442): gamma_w_n =
 * [gamma_b_nm1[-1]; gamma_w_nm1[:-1]] and zeta_w_n[0] = zeta_b_n[-1] (TE
 * row).  The convected rows 1..m_star of zeta_w_n come from fj_pair_sum with
 * FJ_OUT_CONVECT (or fj_wake_convect for a prescribed wake). */
int fj_wake_shift(const double* gamma_b_nm1, const double* gamma_w_nm1, const double* zeta_b_n,
                  int m, int n, int m_star, double* gamma_w_n, double* zeta_w_n, void* stream);

/* Convection of rows 0..m_star-1 of zeta_w_nm1 into rows 1..m_star of
 * zeta_w_n: x + dt * vel (vel (m_star*(n+1), 3) given, e.g. from a caller's
 * v_func) or, if vel is NULL, x + dt * U(x, time) (prescribed wake),
 * aero/utils.py:445-450. */
int fj_wake_convect(const double* zeta_w_nm1, const double* vel, int m_star, int n,
                    const fj_flowfield* flow, double time, double dt, double* zeta_w_n,
                    void* stream);

/* Variable wake discretisation (aero/utils.py:452-491): zeta_pre
 * (m_star+2, n+1, 3) is [TE row; all m_star+1 convected rows], gamma_pre
 * (m_star+1, n) is [TE circulation; all m_star wake rows].  Each column is
 * re-sampled by arc length onto s_disc (device, m_star+1) = [0,
 * cumsum(delta_w)] (vertices) and onto its midpoints (circulations).
 * workspace: fj_wake_rediscretise_workspace_bytes(m_star, n). */
size_t fj_wake_rediscretise_workspace_bytes(int m_star, int n);
int fj_wake_rediscretise(const double* zeta_pre, const double* gamma_pre, int m_star, int n,
                         const double* s_disc, double* zeta_w, double* gamma_w, void* workspace,
                         void* stream);

/* Initial flat wake TE + s_i * dir (initialise_wake, aero/uvlm.py:662-689):
 * s (device, m_star+1) = [0, cumsum(delta_w)], dir (HOST, 3) = unit U_inf. */
int fj_wake_initialise(const double* zeta_b, int m, int n, const double* s, int m_star,
                       const double* dir, double* zeta_w, void* stream);

/* Static wake circulation: every one of `rows` wake rows = TE row of gamma_b
 * (aero/uvlm.py:995-1012). */
int fj_static_wake_gamma(const double* gamma_b, int m, int n, int rows, double* gamma_w,
                         void* stream);

/* Background flow at n_pts points (FlowField.vmap_call, aero/flowfields.py:49-60). */
int fj_flow_eval(const fj_flowfield* flow, const double* x, int n_pts, double time, double* out,
                 void* stream);

/* Steady vertex forces (calculate_steady_forcing, aero/utils.py:126-223)
 * from the velocity v_mid at the filament midpoints (layout of
 * fj_filament_midpoints): f = rho * Gamma_fil * (v_rel x r), split to the
 * filament ends, v_rel = U(mid, time) + v_mid - zeta_dot(mid).  With flow NULL,
 * v_mid is taken as the total velocity (a caller's v_func).  gamma_w0 (n) is
 * the first wake row or NULL.  zeta_dot may be NULL (static). */
int fj_steady_forces(const double* zeta, const double* zeta_dot, const double* gamma_b,
                     const double* gamma_w0, const double* v_mid, int m, int n,
                     const fj_flowfield* flow, double time, double rho, double* f_out,
                     void* stream);

/* Filtered circulation rate and unsteady vertex force (aero/uvlm.py:691-715,
 * :1086-1103): gdot = g*(G_n - G_nm1)/dt + (1-g)*gdot_nm1,
 * f = split_to_vertex(rho * gdot * nc). */
int fj_gamma_dot_unsteady(const double* gamma_n, const double* gamma_nm1,
                          const double* gamma_dot_nm1, const double* nc, int m, int n, double dt,
                          double relax, double rho, double* gamma_dot_n, double* f_unsteady,
                          void* stream);

/* The force stage of a dynamic step in one launch (aero/uvlm.py:1023-1103):
 * steady vertex forces as fj_steady_forces and, if gamma_b_nm1 is non-NULL,
 * the circulation rate and unsteady force as fj_gamma_dot_unsteady
 * (f_unsteady may be NULL).  Results are identical to those entry points. */
int fj_step_forces(const double* zeta, const double* zeta_dot, const double* gamma_b,
                   const double* gamma_b_nm1, const double* gamma_dot_nm1, const double* gamma_w0,
                   const double* v_mid, const double* nc, int m, int n, const fj_flowfield* flow,
                   double time, double rho, double dt, double relax, double* f_steady,
                   double* gamma_dot_n, double* f_unsteady, void* stream);

/* Per-strip effective angle of attack (strip_alpha, aero/utils.py:226-273). */
int fj_strip_alpha(const double* zeta, const double* f_steady, int m, int n,
                   const fj_flowfield* flow, double time, double rho, double* alpha,
                   void* stream);

/* ---- Serving: one time step of many independent cases ----------------- */

/* One active case of a batched serving tick.  All pointers are device
 * pointers to this case's own buffers; the grid / wake / circulation pairs
 * are (previous step, this step) and must not alias. */
typedef struct {
  int m, n, m_star;          /* bound panels (m chordwise, n spanwise), wake rows */
  double dt, time;           /* step size, time of this step */
  const double* hg;          /* (n+1, 4, 4) beam-node poses of this step */
  const double* hg_dot;      /* (n+1, 4, 4) pose rates */
  const int32_t* dof;        /* (n+1) beam node of each spanwise station */
  const double* zeta0;       /* (m+1, n+1, 3) reference grid */
  const double* zeta_b_nm1;  /* (m+1, n+1, 3) bound grid of the previous step */
  double* zeta_b;            /* out: bound grid of this step */
  double* zeta_b_dot;        /* out: its velocity */
  const double* gamma_b_nm1; /* (m, n) */
  const double* gamma_w_nm1; /* (m_star, n) */
  const double* zeta_w_nm1;  /* (m_star+1, n+1, 3) */
  double* gamma_w;           /* out (m_star, n) */
  double* zeta_w;            /* out (m_star+1, n+1, 3) */
  double* aic;               /* (m n, m n): assembled here (bound-only dynamic AIC) */
  const double* aic_inv;     /* (m n, m n): kept inverse of that AIC (fj_lu_inverse) */
  double* gamma_b;           /* out (m, n) */
  double* gamma_b_dot;       /* (m, n): previous rate in, this step's rate out */
  double* f_steady;          /* out (m+1, n+1, 3) steady vertex forces */
  double* lift;              /* out: 1 double, sum of the z force over the vertices */
  double* refine_ws;         /* fj_refine_workspace_bytes(m n) bytes */
  int32_t* iters;            /* out: refinement corrections, negative if the stop test failed */
} fj_step_case;

/* Sizes for fj_step_batched: device workspace, device table and host
 * staging (which must stay valid until the tick's work has run). */
size_t fj_step_batched_workspace_bytes(int n_cases, const fj_step_case* cases);
size_t fj_step_batched_table_bytes(int n_cases);

/* One dynamic UVLM step (the per-step work of base_solve_from_grid,
 * aero/uvlm.py:824-1117, uniform wake, no mirror) for n_cases independent
 * cases in one set of launches: geometry and wake shift, the three filament
 * tables, the rollup / upwash / force pair sums as fused multi-case launches,
 * the AIC, right-hand side and kept-inverse refinement of every case, and
 * the steady forces, circulation rate and lift.  free_wake: 1 free-wake
 * rollup, 0 prescribed convection.  precision: FJ_PREC_MIXED evaluates the
 * rollup and force pair sums in the mixed form (the upwash, AIC and solve
 * stay FP64), as fj_pair_sum does.  No host synchronisation; the stop-test
 * outcome of each case is in its `iters`. */
int fj_step_batched(int n_cases, const fj_step_case* cases, const fj_flowfield* flow,
                    double rho, double relax, int kernel, double eps, double rcut, int precision,
                    int free_wake, int max_iter, double tol, void* device_table,
                    void* host_staging, void* workspace, void* stream);

/* The host half of fj_step_batched with the same arguments: writes the case
 * records and pair-sum task tables of this step into host_staging exactly as
 * fj_step_batched does, and makes no device call.  For replaying a captured
 * fj_step_batched (its copies read host_staging, its kernels read the
 * device table) with the next step's records while the set of cases and
 * every buffer stay the same.  No counterpart in the reference (its step is
 * a jitted function); it serves a serving loop that replays one captured
 * graph per tick. */
int fj_step_batched_restage(int n_cases, const fj_step_case* cases, const fj_flowfield* flow,
                            double rho, double relax, int kernel, double eps, double rcut, int precision,
                            int free_wake, int max_iter, double tol, void* device_table,
                            void* host_staging, void* workspace, void* stream);

/* ---- Structure side of the coupling: the SE(3) two-node beam -------------
 * The Newton iterations of BaseBeamStructure.static_solve and
 * base_dynamic_solve (structure/beam.py:1878-2092, 2129-2861) with the
 * element formulas of structure/utils.py:_k_t_entry, _integrate_m_l and
 * _integrate_c_t (the reference's default OptionalJacobians: no geometric,
 * dead-load or gravity stiffness, no q_dot term in C_T), and the coupling
 * seam of coupled/coupled.py.  Node frames are row-major 4x4 homogeneous
 * transforms (16 doubles), nodal vectors 6 doubles (u, omega). */

/* Element tables of a beam (device pointers), set up once per design. */
typedef struct fj_beam_model {
  int n_nodes, n_elem;
  const int32_t* conn;      /* (n_elem, 2) node indices A, B */
  const double* o0;         /* (n_elem, 3, 3) reference element triads, row-major, columns x y z */
  const double* l0;         /* (n_elem,) reference lengths */
  const double* d0;         /* (n_elem, 6) reference configurations (l0, 0, 0, 0, 0, 0) */
  const double* k_cs;       /* (n_elem, 6, 6) cross-section stiffness of each element */
  const double* m_cs;       /* (n_elem, 6, 6) cross-section mass of each element */
  const int32_t* node_off;  /* (n_nodes + 1) CSR offsets of the elements at each node */
  const int32_t* node_elem; /* (2 n_elem) entries 2 e + end, end 0 for node A and 1 for node B */
} fj_beam_model;

/* Device workspace of fj_beam_system for n_elem elements. */
size_t fj_beam_workspace_bytes(int n_elem);

/* One Newton system: the reduced matrix A (n_solve x n_solve, row-major) and
 * residual r with the absolute-force sums f_abs (n_solve) over the solve DOFs
 * solve_dofs (ascending, 6 node + component).  Replaces the assembly in
 * structure/beam.py:1934-1995 (dynamic = 0: A = K_t, r = f_int + follower +
 * R^T dead) and :2170-2290 (dynamic = 1: A = beta' M + gamma' C_T + K_t T(phi)
 * with t_phi (n_nodes, 6, 6) = T(phi_alpha); r adds the inertial and
 * gyroscopic forces of velocities v and accelerations v_dot, all local).
 * hg: the configuration the iteration linearises about; f_follower (local)
 * and f_dead (global, dead and aerodynamic loads) may be NULL. */
int fj_beam_system(const fj_beam_model* model, int dynamic, const double* hg, const double* v,
                   const double* v_dot, const double* t_phi, const double* f_follower,
                   const double* f_dead, double beta_p, double gamma_p, const int32_t* solve_dofs,
                   int n_solve, double* A, double* r, double* f_abs, void* workspace, void* stream);

/* Exact derivatives of the element forces of the beam residual, for the
 * implicit adjoint of the Newton solves: the reference obtains them with
 * jax.vjp / jax.jacrev of the same residual (coupled/gradients/coupled.py:
 * 195-353 static, :540-629 timestep residual; the residual is
 * structure/beam.py:1749-1843).  Forward-mode dual numbers on the device,
 * nested for the gyroscopic term's Q_dot (algebra/se3.py:q_dot).  Per element
 * e, with d = log(H_A^-1 H_B) at the configuration hg (n_nodes, 4, 4) and the
 * element velocities v and accelerations v_dot (n_nodes, 6, local; dynamic
 * only): fe (n_elem, 12) the element force -P^T K eps [- M_t v_dot - C_L v],
 * jd (n_elem, 12, 6) d fe / d d, jv (n_elem, 12, 12) d fe / d v_e, mt
 * (n_elem, 12, 12) M_t (d fe / d v_dot_e = -M_t), p (n_elem, 6, 12) P(d) =
 * d d / d(right increments of the two node frames) and eps (n_elem, 6) the
 * strain.  k_elem (n_elem, 6, 6): the stiffness each element uses (the
 * scaled table entry).  jv, mt, v, v_dot: dynamic only (may be NULL when
 * dynamic = 0).  Returns FJ_ERR_ARG on a missing pointer. */
int fj_beam_element_jacobians(const fj_beam_model* model, int dynamic, const double* hg, const double* v,
                              const double* v_dot, const double* k_elem, double* fe, double* jd, double* jv,
                              double* mt, double* p, double* eps, void* stream);

/* Newton update with dx = relax x (structure/beam.py:1974-1990, 2274-2296):
 * dynamic = 1 adds dx, gamma' dx and beta' dx to phi, v and v_dot at the solve
 * DOFs; dynamic = 0 writes the full increment (zero at prescribed DOFs) to
 * phi.  norms (5 doubles) = |dx|, |phi| (dynamic; 0 when static), |r|,
 * max |f_abs|, 1 if dx holds a NaN: the inputs of ConvergenceStatus.update
 * (utils/data_structures.py:81-151). */
int fj_beam_update(int dynamic, int n_nodes, const int32_t* solve_dofs, int n_solve, const double* x,
                   double relax, double gamma_p, double beta_p, double* phi, double* v, double* v_dot,
                   const double* r, const double* f_abs, double* norms, void* stream);

/* Convergence settings of a Newton loop (ConvergenceSettings,
 * utils/data_structures.py:16-30): a NaN tolerance is unset (it never
 * passes); max_n_iter <= 0 sets no limit. */
typedef struct fj_convergence {
  int max_n_iter;
  double rel_disp_tol, abs_disp_tol, rel_force_tol, abs_force_tol;
} fj_convergence;

/* Up to n_iter iterations of the dynamic Newton loop at the alpha point
 * (structure/beam.py:2170-2316, 2727-2792) driven by a device status word,
 * so the host waits once per batch instead of once per iteration.  Each
 * iteration: hg_u = hg_prev exp(phi_a) and t_phi = T(phi_a); the Newton
 * system (fj_beam_system, dynamic); x = A^-1 r by the banded solve
 * (bandwidths kl, ku, as fj_band_solve); the update of phi_a, v_a, v_dot_a
 * and the norms (fj_beam_update); then ConvergenceStatus.update on the
 * device.  status (4 int32): [0] iterations run, [1] done (converged, NaN or
 * the iteration limit), [2] converged, [3] NaN met; every launch of an
 * iteration returns at once after done, so iterations may be enqueued past
 * convergence.  reset != 0 zeroes status first (a new loop).  No counterpart
 * entry in the reference (its loop is a Python while around jnp.linalg.solve). */
int fj_beam_newton_dynamic(const fj_beam_model* model, int n_iter, int reset, const double* hg_prev,
                           double* phi_a, double* v_a, double* v_dot_a, const double* f_follower,
                           const double* f_dead, double beta_p, double gamma_p, double relax,
                           const int32_t* solve_dofs, int n_solve, int kl, int ku,
                           const fj_convergence* conv, double* hg_u, double* t_phi, double* A, double* r,
                           double* f_abs, double* x, double* norms, int32_t* status, int32_t* band_info,
                           void* workspace, void* stream);

/* out = |x[idx]| (the static solve's total-displacement norm). */
int fj_beam_subnorm(const double* x, const int32_t* idx, int n, double* out, void* stream);

#define FJ_GA_PREDICT 0 /* p0 v, p1 v_dot, p2 a (step n-1) -> o0 phi_a, o1 v_a, o2 v_dot_a:
                           predict_q + calculate_q_alpha, structure/time_integration.py:86-150 */
#define FJ_GA_ALPHA 1   /* p0 phi_a, p1 v_a, p2 v_prev -> o0 phi_n, o1 v_n, o2 copy of phi_a (or NULL):
                           the step-n state of an FSI iteration, structure/beam.py:2612-2725 */
#define FJ_GA_FINISH 2  /* p0 phi_a, p1 v_a, p2 v_dot_a, p3 v_prev, p4 v_dot_prev, p5 a_prev ->
                           o0 phi, o1 v, o2 v_dot, o3 a: calculate_q_n_from_q_alpha + calculate_a_n,
                           structure/time_integration.py:53-84 */
#define FJ_GA_FORCE 3   /* p0 f_nm1, p1 f_n -> o0 (1 - af) f_n + af f_nm1 (the alpha-point force) */

/* Generalised-alpha bookkeeping of a time step over n values, one launch per
 * op (am, af, gamma, beta, dt: the integrator's constants; unused ones are
 * ignored).  Each product and sum is rounded on its own, in the order of the
 * reference's array expressions (structure/time_integration.py:53-150). */
int fj_beam_ga(int op, int n, double am, double af, double g, double b, double dt, const double* p0,
               const double* p1, const double* p2, const double* p3, const double* p4, const double* p5,
               double* o0, double* o1, double* o2, double* o3, void* stream);

/* fj_beam_ga's FJ_GA_ALPHA or FJ_GA_FINISH over the 6 n_nodes DOFs fused with
 * the node update that follows it (structure/beam.py:2612-2725, 2794-2834):
 * hg = hg_prev exp(phi) per node (FJ_BEAM_NODE_UPDATE) and, for FJ_GA_ALPHA
 * with hg_dot not NULL, hg_dot = hg tilde(v) (FJ_BEAM_NODE_RATE).  ALPHA:
 * phi_a, v_a, v_prev -> phi (phi_n), v (v_n), phi_copy (copy of phi_a, or
 * NULL); FINISH: phi_a, v_a, v_dot_a, v_prev, v_dot_prev, a_prev -> phi, v,
 * v_dot, a.  The same values as the unfused launches. */
int fj_beam_ga_nodes(int op, int n_nodes, double am, double af, const double* phi_a, const double* v_a,
                     const double* v_dot_a, const double* v_prev, const double* v_dot_prev, const double* a_prev,
                     const double* hg_prev, double* phi, double* v, double* v_dot, double* a, double* phi_copy,
                     double* hg, double* hg_dot, void* stream);

/* The FSI convergence norms of a coupled iteration (structure/beam.py:
 * 2700-2720) over the solve DOFs, one workgroup in a fixed order:
 * out = {|phi_old - phi_a|, |phi_a|, |f_al - f_prev|, max |f_al|, 1 if
 * phi_old - phi_a holds a NaN else 0}. */
int fj_beam_fsi_norms(int n_solve, const int32_t* solve_dofs, const double* phi_old, const double* phi_a,
                      const double* f_al, const double* f_prev, double* out, void* stream);

#define FJ_BEAM_NODE_UPDATE 0 /* hg_out = hg exp(x) (structure/beam.py:1845-1856); t_out = T(x)
                                 and h_out = log(hg_ref^-1 hg_out) when not NULL */
#define FJ_BEAM_NODE_REL 1    /* h_out = log(hg^-1 hg_ref) (algebra/se3.py:hg_to_d) */
#define FJ_BEAM_NODE_RATE 2   /* hg_out = hg tilde(x) (structure/beam.py:1521-1531) */
#define FJ_BEAM_NODE_LOCAL 3  /* h_out = (R^T x[0:3], R^T x[3:6]), R = hg[0:3, 0:3]
                                 (structure/beam.py:1387-1395, make_f_dead_ext) */
#define FJ_BEAM_NODE_GLOBAL 4 /* h_out = (R x[0:3], R x[3:6]) */

/* Per-node SE(3) maps over n_nodes frames. */
int fj_beam_nodes(int op, int n_nodes, const double* hg, const double* x, const double* hg_ref,
                  double* hg_out, double* t_out, double* h_out, void* stream);

/* Aerodynamic vertex forces of one surface onto the beam nodes
 * (aero/utils.py:666-697): node dof_mapping[j] receives the sum over the
 * m1 chordwise vertices of column j of f = f_a (+ f_b) and of (R x0) x f,
 * R = rmat[node] (n_nodes, 3, 3); global frame.  clear zeroes all n_nodes
 * rows first (the first surface); later surfaces overwrite shared nodes. */
int fj_beam_project_forces(int n_nodes, const double* rmat, int m1, int n1, const double* f_a,
                           const double* f_b, const double* x0, const int32_t* dof_mapping, int clear,
                           double* out, void* stream);

/* The end of a coupled time step in one launch: fj_beam_ga_nodes
 * (FJ_GA_FINISH: phi, v, v_dot, a of the step from the alpha point, and the
 * node frames hg = hg_prev exp(phi)), fj_beam_project_forces of one surface
 * onto those frames (f_a + f_b, (m1, n1, 3); reference grid x0; node of
 * each of the n1 stations in dof; f_global (n_nodes, 6) zeroed first), the
 * local-frame copy R^T f (f_local, as fj_beam_nodes FJ_BEAM_NODE_LOCAL) and,
 * when phi_a_next is given, the next step's predictor (as fj_beam_ga
 * FJ_GA_PREDICT from v, v_dot, a).  The same arithmetic as those entry
 * points, in one workgroup; n_nodes <= 512.  Restates structure/beam.py:
 * 2727-2792 (step completion), aero/utils.py:666-697 (projection),
 * structure/beam.py:1387-1395 (local frame) and
 * structure/time_integration.py:86-150 (predictor). */
int fj_beam_step_end(int n_nodes, double alpha_m, double alpha_f, double gamma, double beta, double dt,
                     const double* phi_a, const double* v_a, const double* v_dot_a, const double* v_prev,
                     const double* v_dot_prev, const double* a_prev, const double* hg_prev, int m1, int n1,
                     const double* f_a, const double* f_b, const double* x0, const int32_t* dof, double* phi,
                     double* v, double* v_dot, double* a, double* hg, double* f_global, double* f_local,
                     double* phi_a_next, double* v_a_next, double* v_dot_a_next, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FLAPJAX_AMD_H */
