// Shared device-side definitions for the flapjax_amd UVLM hot path (gfx950).
//
// Data layout in HBM (all FP64 unless noted):
//   vertex grids   (m+1, n+1, 3) row-major, xyz innermost (chordwise index first)
//   ring strengths (m, n) row-major
//   segment table  FJ_SEG_STRIDE doubles per de-duplicated filament (see below)
//   targets        (T, 3) AoS
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/flapjax_amd.h"

// One vortex filament, 16 doubles = 128 bytes, read by whole waves through the
// scalar unit (every lane of a wave evaluates the same filament at once):
//   [0..2] y0   start vertex          [3..5] y1   end vertex
//   [6..8] r0 = y1 - y0               [9]   aux  (epsilon kernel: eps*|r0|^4,
//                                                 cutoff kernel: 1/|r0|^2)
//   [10]  g    filament circulation   [11]  |r0|^2
//   [12..15] FP32 copies for the mixed form: r0x, r0y, r0z, aux, g, 0, 0, 0
#define FJ_SEG_STRIDE 16

#define FJ_INV_4PI 0.079577471545947667884441881686257181017229822870228

struct FjFlow {
  int kind;             // FJ_FLOW_CONSTANT or FJ_FLOW_ONE_MINUS_COSINE
  int relative_motion;  // 1: the air moves (u_inf added), 0: the body moves
  double u_inf[3];
  double gust_length, gust_amplitude;
  double travel_dir[3], amp_dir[3], x0[3];
};

static inline FjFlow fj_flow_from_desc(const fj_flowfield* f) {
  FjFlow o;
  o.kind = f ? f->kind : FJ_FLOW_CONSTANT;
  o.relative_motion = f ? f->relative_motion : 1;
  for (int k = 0; k < 3; ++k) {
    o.u_inf[k] = f ? f->u_inf[k] : 0.0;
    o.travel_dir[k] = f ? f->gust_travel_dir[k] : 0.0;
    o.amp_dir[k] = f ? f->gust_amplitude_dir[k] : 0.0;
    o.x0[k] = f ? f->gust_x0[k] : 0.0;
  }
  o.gust_length = f ? f->gust_length : 1.0;
  o.gust_amplitude = f ? f->gust_amplitude : 0.0;
  return o;
}

// Background flow at (x, t): aero/flowfields.py:88-97 (Constant) and
// aero/flowfields.py:158-181 (OneMinusCosine, 0.5*A*(1-cos(pi*s/L)) on 0<s<2L).
__device__ __forceinline__ void fj_flow_eval(const FjFlow& f, double x, double y, double z,
                                             double t, double* u) {
  u[0] = 0.0; u[1] = 0.0; u[2] = 0.0;
  if (f.kind == FJ_FLOW_ONE_MINUS_COSINE) {
    double rx = x - f.x0[0], ry = y - f.x0[1], rz = z - f.x0[2];
    if (f.relative_motion) {
      rx -= f.u_inf[0] * t; ry -= f.u_inf[1] * t; rz -= f.u_inf[2] * t;
    }
    const double s = rx * f.travel_dir[0] + ry * f.travel_dir[1] + rz * f.travel_dir[2];
    if (s > 0.0 && s < 2.0 * f.gust_length) {
      const double a = f.gust_amplitude * 0.5 * (1.0 - cos(M_PI * s / f.gust_length));
      u[0] = f.amp_dir[0] * a; u[1] = f.amp_dir[1] * a; u[2] = f.amp_dir[2] * a;
    }
  }
  if (f.relative_motion) {
    u[0] += f.u_inf[0]; u[1] += f.u_inf[1]; u[2] += f.u_inf[2];
  }
}

struct FjKernelParams {
  double eps;   // regularisation (utils/constants.py:16, EPSILON = 1e-7)
  double eps2;  // eps * eps
  double rcut;  // cutoff radius (utils/constants.py:19, R_CUTOFF = 1e-4)
};

// FP64 1/sqrt(x) for positive normal x: hardware estimate (about 2^-23
// relative) refined by one third-order step y(1 + e/2 + 3e^2/8), e = 1 - x y^2,
// which leaves an error of a few ulp.  No special-value fix-up: callers only
// pass arguments >= eps^2 > 0 (or +0 for the un-regularised kernel, where the
// reference itself yields inf/nan).
__device__ __forceinline__ double fj_rsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Influence of one filament scaled by its circulation g, WITHOUT the 1/(4 pi)
// factor (callers apply it once per target):
//   FJ_KERNEL_EPSILON: aero/utils.py:606-623
//     K = c (r0 . d) / (|c|^2 + eps |r0|^4),  c = r1 x r2,
//     d = r1 / sqrt(|r1|^2 + eps^2) - r2 / sqrt(|r2|^2 + eps^2)
//   FJ_KERNEL_CUTOFF : aero/utils.py:626-648 (0 within rcut of the line,
//                      else c (r0 . d) / |c|^2 with the same regularised d)
//   FJ_KERNEL_PLAIN  : aero/utils.py:580-592 (un-regularised d, / |c|^2)
// The division by D = |c|^2 (+ eps |r0|^4) is folded into the two reciprocal
// square roots: (r0 . d) / D = r0 . (r1 / sqrt(s1 D^2) - r2 / sqrt(s2 D^2)),
// which is the same quantity up to rounding and saves the FP64 divide.
// r1 = x - y0 and r2 = x - y1 are formed as the reference forms them, and
// c = r1 x r2 as well: at a target ON an end point (every wake vertex in the
// rollup is one) one of r1, r2 is exactly zero, so c is exactly zero -- the
// identity c = r0 x r1 would instead leave the rounding error of an FMA-
// contracted cross product of a vector with itself.  r0 . d uses
// r0 . r2 = r0 . r1 - |r0|^2 (one subtraction instead of a dot product).
template <int KIND>
__device__ __forceinline__ void fj_bs_eval(double x, double y, double z, const double* s,
                                           const FjKernelParams& kp, double& vx, double& vy,
                                           double& vz) {
  const double r0x = s[6], r0y = s[7], r0z = s[8];
  const double r1x = x - s[0], r1y = y - s[1], r1z = z - s[2];
  const double r2x = x - s[3], r2y = y - s[4], r2z = z - s[5];
  const double cx = r1y * r2z - r1z * r2y;
  const double cy = r1z * r2x - r1x * r2z;
  const double cz = r1x * r2y - r1y * r2x;
  const double e2 = (KIND == FJ_KERNEL_PLAIN) ? 0.0 : kp.eps2;
  // |r|^2 + eps^2 as FMA chains seeded with eps^2
  const double s1 = fma(r1z, r1z, fma(r1y, r1y, fma(r1x, r1x, e2)));
  const double s2 = fma(r2z, r2z, fma(r2y, r2y, fma(r2x, r2x, e2)));
  const double a1 = r0x * r1x + r0y * r1y + r0z * r1z;
  const double a2 = a1 - s[11];
  // |c|^2 then + eps |r0|^4, the reference's order (and no register copy of
  // the scalar operand)
  double D = fma(cz, cz, fma(cy, cy, cx * cx));
  if (KIND == FJ_KERNEL_EPSILON) D += s[9];
  const double D2 = D * D;
  const double w1 = fj_rsqrt(s1 * D2);
  const double w2 = fj_rsqrt(s2 * D2);
  double q = s[10] * (a1 * w1 - a2 * w2);
  if (KIND == FJ_KERNEL_CUTOFF) {
    // squared distance from x to the filament line: |r1|^2 - (r0.r1)^2 / |r0|^2
    const double rad2 = (s1 - e2) - a1 * (a1 * s[9]);
    q = (rad2 > kp.rcut * kp.rcut) ? q : 0.0;
  }
  vx = cx * q; vy = cy * q; vz = cz * q;
}

// Mixed form (FP32 far field, FP64 near field): a pair is "near" when the
// target lies within 5 filament lengths of the start point (so at least 4 of
// the end point); FP32 contributions are summed FJ_MIXED_FLUSH filaments at a
// time before they are folded into the FP64 accumulators.
#define FJ_NEAR_FACTOR2 25.0f  // (5 filament lengths)^2 from the start point
#define FJ_MIXED_FLUSH 16

// FP32 far-field evaluation from r1 (rounded from the FP64 difference) and
// the filament's FP32 r0 / aux / g: r2 = r1 - r0 (relative error ~eps32 since
// far-field pairs have |r2| >= 4 |r0|), c = r0 x r1, r0.d = a1 w1 - a2 w2.
template <int KIND>
__device__ __forceinline__ void fj_bs_eval_f32(float r1x, float r1y, float r1z, float s1,
                                               const float* sf, float r0sq_f, float eps2f,
                                               float rcut2f, float& vx, float& vy, float& vz) {
  const float r0x = sf[0], r0y = sf[1], r0z = sf[2];
  const float r2x = r1x - r0x, r2y = r1y - r0y, r2z = r1z - r0z;
  const float cx = r0y * r1z - r0z * r1y;
  const float cy = r0z * r1x - r0x * r1z;
  const float cz = r0x * r1y - r0y * r1x;
  const float e2 = (KIND == FJ_KERNEL_PLAIN) ? 0.0f : eps2f;
  const float s2 = r2x * r2x + r2y * r2y + r2z * r2z + e2;
  const float a1 = r0x * r1x + r0y * r1y + r0z * r1z;
  const float a2 = a1 - r0sq_f;
  float D = cx * cx + cy * cy + cz * cz;
  if (KIND == FJ_KERNEL_EPSILON) D += sf[3];
  const float w1 = __builtin_amdgcn_rsqf(s1 + e2);
  const float w2 = __builtin_amdgcn_rsqf(s2);
  float q = sf[4] * (a1 * w1 - a2 * w2) * __builtin_amdgcn_rcpf(D);
  if (KIND == FJ_KERNEL_CUTOFF) {
    const float rad2 = s1 - a1 * (a1 * sf[3]);
    q = (rad2 > rcut2f) ? q : 0.0f;
  }
  vx = cx * q; vy = cy * q; vz = cz * q;
}

// Launch shape of the pair-sum kernels (FP64 and mixed share the plan):
// FJ_BS_WAVES waves per workgroup take interleaved filaments of a chunk;
// grid.x = target tiles of 64 * tpl, grid.y = filament chunks sized so that
// about FJ_BS_TARGET_BLOCKS workgroups run, each wave with at least
// FJ_BS_MIN_SEG_PER_WAVE filaments; two targets per lane from
// FJ_BS_TPL2_MIN_TARGETS targets on.
#ifndef FJ_BS_WAVES
#define FJ_BS_WAVES 8
#endif
#define FJ_BS_THREADS (64 * FJ_BS_WAVES)
#ifndef FJ_BS_TARGET_BLOCKS
#define FJ_BS_TARGET_BLOCKS 8192
#endif
#ifndef FJ_BS_MIN_SEG_PER_WAVE
#define FJ_BS_MIN_SEG_PER_WAVE 8
#endif
#ifndef FJ_BS_TPL2_MIN_TARGETS
#define FJ_BS_TPL2_MIN_TARGETS 16384
#endif

// (tile, chunk) of this workgroup in a (tiles, chunks) pair-sum grid, mapped
// so that each XCD works on a contiguous band of chunks: workgroups are
// handed to the 8 XCDs round-robin by linear id, so XCD k runs ids
// k, k + 8, ...; giving those ids consecutive chunk-major slots means each
// XCD's L2 holds only its own chunks' filament records.
#define FJ_NUM_XCD 8
__device__ __forceinline__ void fj_tile_chunk(int& tile, int& chunk) {
  const int tiles = gridDim.x;
  const int total = tiles * gridDim.y;
  const int lin = blockIdx.x + blockIdx.y * tiles;
  const int k = lin % FJ_NUM_XCD, idx = lin / FJ_NUM_XCD;
  int base = 0;  // workgroups on XCDs 0 .. k-1
  for (int j = 0; j < k; ++j) base += (total - j + FJ_NUM_XCD - 1) / FJ_NUM_XCD;
  const int slot = base + idx;
  chunk = slot / tiles;
  tile = slot - chunk * tiles;
}

// Launch plan of a pair sum (fj_biot_savart.hip)
struct FjPairPlan {
  int tpl;        // targets per lane
  int tiles;      // target tiles (grid.x)
  int chunks;     // filament chunks (grid.y)
  int chunk_len;  // filaments per chunk
  int ld;         // padded target count of the partial buffer
};

// one task of a fused multi-case pair sum (fj_batched.hip): targets and
// outputs by pointer, filaments by row range of a shared table
// n_lat > 0: the task's filaments are whole lattices (rows lat_off of the
// table, m x n rings each), which lets the FP64 form run the lattice kernel
#define FJ_TASK_MAX_LAT 2
struct FjTaskDesc {
  const double* tgt;
  double* out;
  int n_tgt, seg_off, n_seg, mode;
  double dt, time;
  FjFlow flow;
  int n_lat;
  int lat_off[FJ_TASK_MAX_LAT], lat_m[FJ_TASK_MAX_LAT], lat_n[FJ_TASK_MAX_LAT];
};
size_t fj_batched_ws_bytes(const FjTaskDesc* d, int n, int precision);
size_t fj_batched_desc_bytes(int n_tasks);
// launch = false: fill host_staging only (no device call), exactly as a
// launching call would, for replaying a captured launch with new tasks
int fj_batched_launch(const FjTaskDesc* d, int n_tasks, const double* seg, int kernel,
                      int precision, const FjKernelParams& kp, void* task_table,
                      void* host_staging, void* workspace, hipStream_t st, bool launch = true);

// pair-sum epilogue (fj_biot_savart.hip): chunk partials -> out by mode
int fj_pair_epilogue(const double* part, int chunks, int ld, int n_tgt, const double* targets,
                     double* out, int mode, const FjFlow& fl, double time, double dt,
                     hipStream_t st);

// mixed-precision partial sums (fj_biot_savart_mixed.hip)
int fj_pair_partial_mixed(const FjPairPlan* p, const double* seg, int n_seg, const double* tgt,
                          int n_tgt, double* part, int kernel, const FjKernelParams* kp,
                          hipStream_t st);

// Launch-shape helpers
__host__ __device__ static inline int fj_cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

#define FJ_CHECK_LAUNCH()                         \
  do {                                            \
    hipError_t e_ = hipGetLastError();            \
    if (e_ != hipSuccess) return FJ_ERR_LAUNCH;   \
  } while (0)
