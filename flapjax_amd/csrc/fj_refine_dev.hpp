// Device helpers of the FP64 iterative refinement (fj_lu.hip), shared with
// the batched serving tick (fj_serving.hip).
#pragma once
#include "fj_common.hpp"

#define FJ_REFINE_MAX_ITER 64
#define FJ_RV_WAVES 4

typedef unsigned long long fj_u64;

__device__ __forceinline__ fj_u64 fj_bits(double v) { return (fj_u64)__double_as_longlong(v); }
__device__ __forceinline__ double fj_from_bits(fj_u64 b) { return __longlong_as_double((long long)b); }

// max of a non-negative double over the workgroup, then one atomic per group
__device__ __forceinline__ void fj_group_max(double v, fj_u64* dst, double* sred) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  if (lane == 0) sred[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sred[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, sred[w]);
    atomicMax(dst, fj_bits(m));
  }
}

__device__ __forceinline__ bool fj_refine_converged(const fj_u64* slot, double tol) {
  const double an = fj_from_bits(slot[0]), rn = fj_from_bits(slot[1]), xn = fj_from_bits(slot[2]);
  return rn <= tol * an * xn;
}


// workspace of one refinement: r (n doubles), per-sweep norm slots, flags
struct FjRefineWs {
  double* r;
  fj_u64* slots;
  int32_t* flags;
};

__host__ __device__ inline FjRefineWs fj_refine_ws(void* workspace, int n) {
  FjRefineWs w;
  w.r = (double*)workspace;
  w.slots = (fj_u64*)(w.r + n + 4);
  w.flags = (int32_t*)(w.slots + 3 * (FJ_REFINE_MAX_ITER + 1));
  return w;
}

__host__ __device__ inline double fj_refine_tol(double tol, int n) {
  return tol > 0.0 ? tol : sqrt((double)n) * 1.1102230246251565e-16;
}
