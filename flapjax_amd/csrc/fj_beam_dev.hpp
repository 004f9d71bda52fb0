// SE(3) algebra and two-node beam element quantities on the device
// (algebra/so3.py, algebra/se3.py, structure/utils.py of the reference).
//
// Every function is templated on the scalar: double, or FjDual (a value and
// one directional derivative) where the reference takes a Jacobian-vector
// product (q_dot = JVP of Q(s, d) along d_dot, algebra/se3.py:q_dot).  The
// small-angle switches follow the reference exactly, including which of them
// test |w| and which |w|^2; a dual number takes the branch of its value, as
// the reference's JVP differentiates the branch its predicate selects.
#pragma once
#include <hip/hip_runtime.h>

#define FJ_SMALL_ANG 1e-10  // utils/constants.py:13

// A value and one directional derivative.  T = double for first
// derivatives; T = FjDualT<double> nests them (the derivative of a JVP, as
// the exact Jacobian of the gyroscopic forces needs).
template <typename T>
struct FjDualT {
  T v, d;
  __device__ FjDualT() : v(0.0), d(0.0) {}
  __device__ FjDualT(double x) : v(x), d(0.0) {}
  __device__ FjDualT(T x, T dx) : v(x), d(dx) {}
};
typedef FjDualT<double> FjDual;

template <typename T>
__device__ __forceinline__ FjDualT<T> operator+(FjDualT<T> a, FjDualT<T> b) { return FjDualT<T>(a.v + b.v, a.d + b.d); }
template <typename T>
__device__ __forceinline__ FjDualT<T> operator-(FjDualT<T> a, FjDualT<T> b) { return FjDualT<T>(a.v - b.v, a.d - b.d); }
template <typename T>
__device__ __forceinline__ FjDualT<T> operator-(FjDualT<T> a) { return FjDualT<T>(-a.v, -a.d); }
template <typename T>
__device__ __forceinline__ FjDualT<T> operator*(FjDualT<T> a, FjDualT<T> b) {
  return FjDualT<T>(a.v * b.v, a.d * b.v + a.v * b.d);
}
template <typename T>
__device__ __forceinline__ FjDualT<T> operator/(FjDualT<T> a, FjDualT<T> b) {
  const T q = a.v / b.v;
  return FjDualT<T>(q, (a.d - q * b.d) / b.v);
}
template <typename T>
__device__ __forceinline__ FjDualT<T> operator+(FjDualT<T> a, double b) { return FjDualT<T>(a.v + b, a.d); }
template <typename T>
__device__ __forceinline__ FjDualT<T> operator+(double a, FjDualT<T> b) { return FjDualT<T>(a + b.v, b.d); }
template <typename T>
__device__ __forceinline__ FjDualT<T> operator-(FjDualT<T> a, double b) { return FjDualT<T>(a.v - b, a.d); }
template <typename T>
__device__ __forceinline__ FjDualT<T> operator-(double a, FjDualT<T> b) { return FjDualT<T>(a - b.v, -b.d); }
template <typename T>
__device__ __forceinline__ FjDualT<T> operator*(FjDualT<T> a, double b) { return FjDualT<T>(a.v * b, a.d * b); }
template <typename T>
__device__ __forceinline__ FjDualT<T> operator*(double a, FjDualT<T> b) { return FjDualT<T>(a * b.v, a * b.d); }
template <typename T>
__device__ __forceinline__ FjDualT<T> operator/(FjDualT<T> a, double b) { return FjDualT<T>(a.v / b, a.d / b); }
template <typename T>
__device__ __forceinline__ FjDualT<T> operator/(double a, FjDualT<T> b) {
  const T q = a / b.v;
  return FjDualT<T>(q, -q * b.d / b.v);
}
__device__ __forceinline__ double fj_sqrt(double a) { return sqrt(a); }
__device__ __forceinline__ double fj_sin(double a) { return sin(a); }
__device__ __forceinline__ double fj_cos(double a) { return cos(a); }
__device__ __forceinline__ double fj_val(double a) { return a; }
template <typename T>
__device__ __forceinline__ FjDualT<T> fj_sqrt(FjDualT<T> a) {
  const T s = fj_sqrt(a.v);
  return FjDualT<T>(s, 0.5 * a.d / s);
}
template <typename T>
__device__ __forceinline__ FjDualT<T> fj_sin(FjDualT<T> a) { return FjDualT<T>(fj_sin(a.v), fj_cos(a.v) * a.d); }
template <typename T>
__device__ __forceinline__ FjDualT<T> fj_cos(FjDualT<T> a) { return FjDualT<T>(fj_cos(a.v), -fj_sin(a.v) * a.d); }
template <typename T>
__device__ __forceinline__ double fj_val(FjDualT<T> a) { return fj_val(a.v); }

// ---- 3x3 / 6x6 helpers (row-major) -----------------------------------------
template <typename S>
__device__ __forceinline__ S fj_dot3(const S* a, const S* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

template <typename S>
__device__ __forceinline__ void fj_skew(const S* v, S m[3][3]) {
  m[0][0] = S(0.0); m[0][1] = -v[2]; m[0][2] = v[1];
  m[1][0] = v[2]; m[1][1] = S(0.0); m[1][2] = -v[0];
  m[2][0] = -v[1]; m[2][1] = v[0]; m[2][2] = S(0.0);
}

template <typename S, int N>
__device__ __forceinline__ void fj_matmul(const S a[N][N], const S b[N][N], S c[N][N]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      S acc = a[i][0] * b[0][j];
#pragma unroll
      for (int k = 1; k < N; ++k) acc = acc + a[i][k] * b[k][j];
      c[i][j] = acc;
    }
}

// ---- SO(3) (algebra/so3.py) ---------------------------------------------------
// alpha = sin|b| / |b|, beta = 2 (1 - cos|b|) / |b|^2 (Sonneville 2013 eq. A.4),
// both switching on |b|.
template <typename S>
__device__ __forceinline__ S fj_alpha(const S* b) {
  const S nb = fj_sqrt(fj_dot3(b, b));
  if (fj_val(nb) > FJ_SMALL_ANG) return fj_sin(nb) / nb;
  return 1.0 - nb * nb / 6.0;
}

template <typename S>
__device__ __forceinline__ S fj_beta(const S* b) {
  const S nb2 = fj_dot3(b, b);
  const S nb = fj_sqrt(nb2);
  if (fj_val(nb) > FJ_SMALL_ANG) return 2.0 * (1.0 - fj_cos(nb)) / nb2;
  return 0.5 - nb2 / 24.0;
}

// T(w) (eq. A.6), switch on |w|^2; small branch I - S/2 + S^2/6
template <typename S>
__device__ void fj_t_so3(const S* w, S t[3][3]) {
  S s[3][3], s2[3][3];
  fj_skew(w, s);
  fj_matmul<S, 3>(s, s, s2);
  const S n2 = fj_dot3(w, w);
  S c1, c2;
  if (fj_val(n2) > FJ_SMALL_ANG) {
    c1 = -0.5 * fj_beta(w);
    c2 = (1.0 - fj_alpha(w)) / n2;
  } else {
    c1 = S(-0.5);
    c2 = S(1.0 / 6.0);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) t[i][j] = (i == j ? S(1.0) : S(0.0)) + c1 * s[i][j] + c2 * s2[i][j];
}

// T^-1(w) (eq. A.7), switch on |w|; small branch I + S/2 + S^2/12
template <typename S>
__device__ void fj_t_inv_so3(const S* w, S t[3][3]) {
  S s[3][3], s2[3][3];
  fj_skew(w, s);
  fj_matmul<S, 3>(s, s, s2);
  const S n2 = fj_dot3(w, w);
  S c2;
  if (fj_val(fj_sqrt(n2)) > FJ_SMALL_ANG) {
    c2 = (1.0 - fj_alpha(w) / fj_beta(w)) / n2;
  } else {
    c2 = S(1.0 / 12.0);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) t[i][j] = (i == j ? S(1.0) : S(0.0)) + 0.5 * s[i][j] + c2 * s2[i][j];
}

// ---- SE(3) (algebra/se3.py), h = (u, w) ------------------------------------------
template <typename S>
__device__ __forceinline__ void fj_hat6(const S* h, S a[6][6]) {
  S sw[3][3], su[3][3];
  fj_skew(h + 3, sw);
  fj_skew(h, su);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      a[i][j] = sw[i][j];
      a[i][j + 3] = su[i][j];
      a[i + 3][j] = S(0.0);
      a[i + 3][j + 3] = sw[i][j];
    }
}

// T_U-omega+ (eq. A.12), switch on |w|^2
template <typename S>
__device__ void fj_t_u_omega_plus(const S* h, S out[3][3]) {
  const S* a = h;
  const S* b = h + 3;
  S sa[3][3], sb[3][3];
  fj_skew(a, sa);
  fj_skew(b, sb);
  const S n2 = fj_dot3(b, b);
  if (fj_val(n2) > FJ_SMALL_ANG) {
    const S al = fj_alpha(b), be = fj_beta(b);
    const S ba = fj_dot3(b, a);
    const S k1 = -0.5 * be;
    const S k2 = (1.0 - al) / n2;
    const S k3 = ba / n2 * (be - al);
    const S k4 = ba / n2 * (0.5 * be - 3.0 * (1.0 - al) / n2);
    S ab[3][3], bab[3][3], bb[3][3];
    fj_matmul<S, 3>(sa, sb, ab);
    fj_matmul<S, 3>(sb, sa, bab);
    fj_matmul<S, 3>(sb, sb, bb);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        out[i][j] = k1 * sa[i][j] + k2 * (ab[i][j] + bab[i][j]) + k3 * sb[i][j] + k4 * bb[i][j];
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) out[i][j] = -0.5 * sa[i][j];
  }
}

// series branch of the SE(3) tangents: I + c1 A + c2 A^2 with A = hat(h)
template <typename S>
__device__ void fj_hat_series(const S* h, double c1, double c2, S t[6][6]) {
  S a[6][6], a2[6][6];
  fj_hat6(h, a);
  fj_matmul<S, 6>(a, a, a2);
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) t[i][j] = (i == j ? S(1.0) : S(0.0)) + c1 * a[i][j] + c2 * a2[i][j];
}

// T(h) (eq. A.11); small branch I - A/2 + A^2/6
template <typename S>
__device__ void fj_t_se3(const S* h, S t[6][6]) {
  if (fj_val(fj_dot3(h + 3, h + 3)) > FJ_SMALL_ANG) {
    S r[3][3], x[3][3];
    fj_t_so3(h + 3, r);
    fj_t_u_omega_plus(h, x);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        t[i][j] = r[i][j];
        t[i][j + 3] = x[i][j];
        t[i + 3][j] = S(0.0);
        t[i + 3][j + 3] = r[i][j];
      }
  } else {
    fj_hat_series(h, -0.5, 1.0 / 6.0, t);
  }
}

// T^-1(h) (eq. A.13/A.14); small branch I + A/2 + A^2/12
template <typename S>
__device__ void fj_t_inv_se3(const S* h, S t[6][6]) {
  if (fj_val(fj_dot3(h + 3, h + 3)) > FJ_SMALL_ANG) {
    S ti[3][3], x[3][3], y[3][3], z[3][3];
    fj_t_inv_so3(h + 3, ti);
    fj_t_u_omega_plus(h, x);
    fj_matmul<S, 3>(ti, x, y);
    fj_matmul<S, 3>(y, ti, z);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        t[i][j] = ti[i][j];
        t[i][j + 3] = -z[i][j];
        t[i + 3][j] = S(0.0);
        t[i + 3][j + 3] = ti[i][j];
      }
  } else {
    fj_hat_series(h, 0.5, 1.0 / 12.0, t);
  }
}

// ---- group maps (double only) ----------------------------------------------------
// 4x4 homogeneous transforms stored row-major, 16 doubles.
__device__ __forceinline__ void fj_exp_se3(const double* h, double g[16]) {
  const double* w = h + 3;
  const double n2 = fj_dot3(w, w);
  if (n2 > FJ_SMALL_ANG) {
    const double a = sqrt(n2);
    double s[3][3], s2[3][3], t[3][3];
    fj_skew(w, s);
    fj_matmul<double, 3>(s, s, s2);
    const double c1 = sin(a) / a, c2 = (1.0 - cos(a)) / (a * a);
    fj_t_so3(w, t);
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) g[4 * i + j] = (i == j ? 1.0 : 0.0) + c1 * s[i][j] + c2 * s2[i][j];
      g[4 * i + 3] = t[0][i] * h[0] + t[1][i] * h[1] + t[2][i] * h[2];  // T(w)^T u
    }
  } else {
    // I + X + X^2/2 with X = [[S(w), u], [0, 0]]
    double s[3][3], s2[3][3];
    fj_skew(w, s);
    fj_matmul<double, 3>(s, s, s2);
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) g[4 * i + j] = (i == j ? 1.0 : 0.0) + s[i][j] + 0.5 * s2[i][j];
      g[4 * i + 3] = h[i] + 0.5 * (s[i][0] * h[0] + s[i][1] * h[1] + s[i][2] * h[2]);
    }
  }
  g[12] = 0.0; g[13] = 0.0; g[14] = 0.0; g[15] = 1.0;
}

__device__ __forceinline__ void fj_unskew(const double m[3][3], double* v) {
  v[0] = 0.5 * (m[2][1] - m[1][2]);
  v[1] = 0.5 * (m[0][2] - m[2][0]);
  v[2] = 0.5 * (m[1][0] - m[0][1]);
}

// log_so3 (switch on cos(theta) < cos(1e-10), i.e. cos(theta) < 1)
__device__ __forceinline__ void fj_log_so3(const double* g, double* w) {
  const double ct = 0.5 * (g[0] + g[5] + g[10] - 1.0);
  double m[3][3];
  if (ct < cos(FJ_SMALL_ANG)) {
    const double th = acos(ct);
    const double f = th / (2.0 * sin(th));
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) m[i][j] = f * (g[4 * i + j] - g[4 * j + i]);
  } else {
    // (R - I) - (R - I)^2 / 2
    double e[3][3], e2[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) e[i][j] = g[4 * i + j] - (i == j ? 1.0 : 0.0);
    fj_matmul<double, 3>(e, e, e2);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) m[i][j] = e[i][j] - 0.5 * e2[i][j];
  }
  fj_unskew(m, w);
}

// log_se3 (eq. A.15): h = (T^-1(w)^T y, w); small branch from (G - I) - (G - I)^2/2
__device__ __forceinline__ void fj_log_se3(const double* g, double* h) {
  double w[3];
  fj_log_so3(g, w);
  if (fj_dot3(w, w) > FJ_SMALL_ANG) {
    double ti[3][3];
    fj_t_inv_so3(w, ti);
    for (int i = 0; i < 3; ++i) h[i] = ti[0][i] * g[3] + ti[1][i] * g[7] + ti[2][i] * g[11];
    h[3] = w[0]; h[4] = w[1]; h[5] = w[2];
  } else {
    double e[4][4], e2[4][4], l[3][3];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) e[i][j] = g[4 * i + j] - (i == j ? 1.0 : 0.0);
    fj_matmul<double, 4>(e, e, e2);
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) l[i][j] = e[i][j] - 0.5 * e2[i][j];
      h[i] = e[i][3] - 0.5 * e2[i][3];
    }
    fj_unskew(l, h + 3);
  }
}

// inv(a) @ b for homogeneous transforms
__device__ __forceinline__ void fj_hg_inv_mul(const double* a, const double* b, double* c) {
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 4; ++j) {
      double acc = a[i] * b[j] + a[4 + i] * b[4 + j] + a[8 + i] * b[8 + j];  // R^T b
      if (j == 3) acc -= a[i] * a[3] + a[4 + i] * a[7] + a[8 + i] * a[11];   // - R^T y
      c[4 * i + j] = acc;
    }
  }
  c[12] = 0.0; c[13] = 0.0; c[14] = 0.0; c[15] = 1.0;
}

__device__ __forceinline__ void fj_hg_mul(const double* a, const double* b, double* c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) {
      double acc = a[4 * i] * b[j] + a[4 * i + 1] * b[4 + j] + a[4 * i + 2] * b[8 + j];
      if (j == 3) acc += a[4 * i + 3];
      c[4 * i + j] = acc;
    }
  c[12] = 0.0; c[13] = 0.0; c[14] = 0.0; c[15] = 1.0;
}

// ---- element operators (algebra/se3.py:p, t_star, q) ---------------------------------
// Right-multiplication by Ad = chi(O^T) = blockdiag(O^T, O^T), O the element's
// reference triad (columns): (X Ad)[i][3B + j] = sum_k X[i][3B + k] O[j][k].
template <typename S>
__device__ __forceinline__ S fj_times_ad(const S x[6][6], const double* o, int i, int c) {
  const int b = (c / 3) * 3, j = c % 3;
  return x[i][b] * o[3 * j] + x[i][b + 1] * o[3 * j + 1] + x[i][b + 2] * o[3 * j + 2];
}

// P(d) = [-T^-1(-d) Ad, T^-1(d) Ad], T^-1(-d) = T^-1(d) - hat(d)
template <typename S>
__device__ void fj_p_mat(const S* d, const double* o, S p[6][12]) {
  S tp[6][6], tn[6][6], a[6][6];
  fj_t_inv_se3(d, tp);
  fj_hat6(d, a);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) tn[i][j] = tp[i][j] - a[i][j];
  for (int i = 0; i < 6; ++i)
    for (int c = 0; c < 6; ++c) {
      p[i][c] = -fj_times_ad(tn, o, i, c);
      p[i][c + 6] = fj_times_ad(tp, o, i, c);
    }
}

// Q(s, d) = [(I - T*) Ad, T* Ad], T* = s T(s d) T^-1(d)
template <typename S>
__device__ void fj_q_mat(double s, const S* d, const double* o, S q[6][12]) {
  S sd[6], ts[6][6], ti[6][6], st[6][6];
  for (int k = 0; k < 6; ++k) sd[k] = s * d[k];
  fj_t_se3(sd, ts);
  fj_t_inv_se3(d, ti);
  fj_matmul<S, 6>(ts, ti, st);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      st[i][j] = s * st[i][j];
      ts[i][j] = (i == j ? S(1.0) : S(0.0)) - st[i][j];  // reuse ts for I - T*
    }
  for (int i = 0; i < 6; ++i)
    for (int c = 0; c < 6; ++c) {
      q[i][c] = fj_times_ad(ts, o, i, c);
      q[i][c + 6] = fj_times_ad(st, o, i, c);
    }
}
