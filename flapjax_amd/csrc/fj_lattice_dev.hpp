// Device side of the lattice pair sums (fj_lattice.hip, fj_batched.hip):
// the strip-column record layout, the per-vertex terms and the filament
// evaluation sharing them.  The method is described in fj_lattice.hip.
#pragma once
#include "fj_common.hpp"

#ifndef FJ_LAT_ROWS
#define FJ_LAT_ROWS 2
#endif

// strip-column record: vertices (FJ_LAT_ROWS + 1) x 3, chordwise filaments
// FJ_LAT_ROWS x 5, spanwise filaments (to the previous column)
// (FJ_LAT_ROWS + 1) x 5; a filament's 5 are g r0 (3), aux, g (r0 for the
// cutoff kernel)
#define FJ_LAT_COFF (3 * (FJ_LAT_ROWS + 1))
#define FJ_LAT_SOFF (FJ_LAT_COFF + 5 * FJ_LAT_ROWS)
#define FJ_LAT_RAW (FJ_LAT_SOFF + 5 * (FJ_LAT_ROWS + 1))
#define FJ_LAT_REC ((FJ_LAT_RAW + 3) / 4 * 4)

struct FjVtx {
  double rx, ry, rz, s, inv;
};

__device__ __forceinline__ FjVtx fj_lat_vertex(const double* __restrict__ v, double x, double y,
                                               double z, double e2) {
  FjVtx o;
  o.rx = x - v[0];
  o.ry = y - v[1];
  o.rz = z - v[2];
  o.s = fma(o.rz, o.rz, fma(o.ry, o.ry, fma(o.rx, o.rx, e2)));
  o.inv = fj_rsqrt(o.s);
  return o;
}

// filament A -> B with data f: (g r0, eps |r0|^4, -) for the epsilon
// kernel, (g r0, -, -) for the plain kernel, (r0, 1/|r0|^2, g) for the
// cutoff kernel (r0 = B - A)
template <int KIND>
__device__ __forceinline__ void fj_lat_eval(const FjVtx& A, const FjVtx& B,
                                            const double* __restrict__ f, double e2,
                                            double rcut2, double& ax, double& ay, double& az) {
  const double cx = A.ry * B.rz - A.rz * B.ry;
  const double cy = A.rz * B.rx - A.rx * B.rz;
  const double cz = A.rx * B.ry - A.ry * B.rx;
  const double dx = fma(A.rx, A.inv, -(B.rx * B.inv));
  const double dy = fma(A.ry, A.inv, -(B.ry * B.inv));
  const double dz = fma(A.rz, A.inv, -(B.rz * B.inv));
  const double a = f[0] * dx + f[1] * dy + f[2] * dz;
  // |c|^2 + eps |r0|^4 as an FMA chain seeded with eps |r0|^4
  const double D = fma(cz, cz, fma(cy, cy, fma(cx, cx, KIND == FJ_KERNEL_EPSILON ? f[3] : 0.0)));
  // 1/D: hardware estimate, then y (1 + e + e^2) with e = 1 - D y
  double w = __builtin_amdgcn_rcp(D);
  const double e = fma(-D, w, 1.0);
  w = fma(w, fma(e, e, e), w);
  double q = (KIND == FJ_KERNEL_CUTOFF) ? (f[4] * a) * w : a * w;
  if (KIND == FJ_KERNEL_CUTOFF) {
    // squared distance from x to the filament line: |r1|^2 - (r0.r1)^2 / |r0|^2
    const double a1 = f[0] * A.rx + f[1] * A.ry + f[2] * A.rz;
    const double rad2 = (A.s - e2) - a1 * (a1 * f[3]);
    q = (rad2 > rcut2) ? q : 0.0;
  }
  ax = fma(cx, q, ax);
  ay = fma(cy, q, ay);
  az = fma(cz, q, az);
}

// one column j of a strip: the vertices of column j into C, the chordwise
// filaments between them, the spanwise filaments from P (column j - 1)
template <int KIND, int R, int SR>
__device__ __forceinline__ void fj_lat_column(const double* __restrict__ rp, bool span,
                                              const FjVtx (&P)[R + 1], FjVtx (&C)[R + 1],
                                              double x, double y, double z, double e2,
                                              double rcut2, double& ax, double& ay, double& az) {
  double r[FJ_LAT_RAW];
#pragma unroll
  for (int k = 0; k < FJ_LAT_RAW; ++k) r[k] = rp[k];
#pragma unroll
  for (int k = 0; k <= R; ++k) C[k] = fj_lat_vertex(r + 3 * k, x, y, z, e2);
#pragma unroll
  for (int k = 0; k < R; ++k)
    fj_lat_eval<KIND>(C[k], C[k + 1], r + FJ_LAT_COFF + 5 * k, e2, rcut2, ax, ay, az);
  if (span) {
#pragma unroll
    for (int k = 0; k < SR; ++k)
      fj_lat_eval<KIND>(P[k], C[k], r + FJ_LAT_SOFF + 5 * k, e2, rcut2, ax, ay, az);
  }
}

// one task: R chordwise filament rows and SR spanwise rows (R, plus the
// closing row on a lattice's last strip) of the strip whose column records
// start at `col`, chordwise columns jb .. je-1
template <int KIND, int R, int SR>
__device__ __forceinline__ void fj_lat_task(const double* __restrict__ col, int jb, int je,
                                            double x, double y, double z, double e2, double rcut2,
                                            double& ax, double& ay, double& az) {
  FjVtx P[R + 1], C[R + 1];
  if (jb > 0) {
    const double* r = col + (size_t)(jb - 1) * FJ_LAT_REC;
#pragma unroll
    for (int k = 0; k <= R; ++k) P[k] = fj_lat_vertex(r + 3 * k, x, y, z, e2);
  }
  // two columns per trip with the roles of P and C swapped, so the vertex
  // window is never copied
  int j = jb;
  for (; j + 1 < je; j += 2) {
    fj_lat_column<KIND, R, SR>(col + (size_t)j * FJ_LAT_REC, j > 0, P, C, x, y, z, e2, rcut2, ax,
                               ay, az);
    fj_lat_column<KIND, R, SR>(col + (size_t)(j + 1) * FJ_LAT_REC, true, C, P, x, y, z, e2, rcut2,
                               ax, ay, az);
  }
  if (j < je)
    fj_lat_column<KIND, R, SR>(col + (size_t)j * FJ_LAT_REC, j > 0, P, C, x, y, z, e2, rcut2, ax,
                               ay, az);
}

// word f of the strip-column record (strip s, column j) of the lattice whose
// filaments start at row `off` of the filament table (fj_segments_from_grid
// order: chordwise (m, n+1), then spanwise (m+1, n); r0 at [6..8], aux at
// [9], g at [10] of a filament record)
__device__ __forceinline__ double fj_lat_record_word(const double* __restrict__ seg, int off,
                                                     int m, int n, int s, int j, int f,
                                                     int kernel) {
  const int i0 = s * FJ_LAT_ROWS;
  const int rows = min(FJ_LAT_ROWS, m - i0);
  const int srows = rows + ((i0 + rows == m) ? 1 : 0);
  const double* chord = seg + (size_t)off * FJ_SEG_STRIDE;
  const double* span = seg + (size_t)(off + m * (n + 1)) * FJ_SEG_STRIDE;
  if (f < FJ_LAT_COFF) {
    // vertex (i, j): start of chordwise filament (i, j), or the end of
    // chordwise filament (m-1, j) on the last vertex row
    const int k = f / 3, c = f - 3 * k, i = i0 + k;
    if (k > rows) return 0.0;
    return (i < m) ? chord[(size_t)(i * (n + 1) + j) * FJ_SEG_STRIDE + c]
                   : chord[(size_t)((m - 1) * (n + 1) + j) * FJ_SEG_STRIDE + 3 + c];
  }
  if (f >= FJ_LAT_RAW) return 0.0;
  const bool ch = f < FJ_LAT_SOFF;
  const int base = ch ? FJ_LAT_COFF : FJ_LAT_SOFF;
  const int k = (f - base) / 5, c = f - base - 5 * k;
  const double* rec = nullptr;
  if (ch && k < rows) rec = chord + (size_t)((i0 + k) * (n + 1) + j) * FJ_SEG_STRIDE;
  if (!ch && j > 0 && k < srows) rec = span + (size_t)((i0 + k) * n + j - 1) * FJ_SEG_STRIDE;
  if (!rec) return 0.0;
  return (kernel == FJ_KERNEL_CUTOFF || c >= 3) ? rec[6 + c] : rec[10] * rec[6 + c];
}

// one strip task (R chordwise rows, SR spanwise rows) dispatched on the
// strip's actual row count; every strip but a lattice's last has
// FJ_LAT_ROWS rows
template <int KIND>
__device__ __forceinline__ void fj_lat_strip(const double* __restrict__ col, int rows, int last,
                                             int jb, int je, double x, double y, double z,
                                             double e2, double rcut2, double& ax, double& ay,
                                             double& az) {
#define FJ_LAT_ARGS col, jb, je, x, y, z, e2, rcut2, ax, ay, az
  if (rows == FJ_LAT_ROWS && !last) fj_lat_task<KIND, FJ_LAT_ROWS, FJ_LAT_ROWS>(FJ_LAT_ARGS);
  else if (rows == FJ_LAT_ROWS) fj_lat_task<KIND, FJ_LAT_ROWS, FJ_LAT_ROWS + 1>(FJ_LAT_ARGS);
#if FJ_LAT_ROWS > 3
  else if (rows == 3) fj_lat_task<KIND, 3, 4>(FJ_LAT_ARGS);
#endif
#if FJ_LAT_ROWS > 2
  else if (rows == 2) fj_lat_task<KIND, 2, 3>(FJ_LAT_ARGS);
#endif
#if FJ_LAT_ROWS > 1
  else if (rows == 1) fj_lat_task<KIND, 1, 2>(FJ_LAT_ARGS);
#endif
#undef FJ_LAT_ARGS
}
