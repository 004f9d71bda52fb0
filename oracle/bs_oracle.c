/* FP64 C restatement of the regularised Biot-Savart pair sum (test/baseline
 * infrastructure only: the CPU baseline timed beside the GPU, and a faster
 * checker for large oracle runs).  Same formula as oracle/uvlm_oracle.py
 * bs_epsilon, which follows aero/utils.py:606-623 of the reference:
 *   K = c (r0 . d) / (4 pi (|c|^2 + eps |r0|^4)), c = r1 x r2,
 *   d = r1 / sqrt(|r1|^2 + eps^2) - r2 / sqrt(|r2|^2 + eps^2).
 * out[t] = sum_s g[s] K(x_t; y0[s], y1[s]); OpenMP over targets. */
#include <math.h>
#include <omp.h>
#include <stddef.h>

/* OpenMP team size of the following calls (CPU-baseline runs at 1 core and
 * at all host cores in one process). */
void fj_oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

void fj_oracle_pair_sum_eps(const double* x, long nt, const double* y0, const double* y1,
                            const double* g, long ns, double eps, double* out) {
  const double inv4pi = 1.0 / (4.0 * M_PI);
#pragma omp parallel for schedule(static)
  for (long t = 0; t < nt; ++t) {
    const double px = x[3 * t], py = x[3 * t + 1], pz = x[3 * t + 2];
    double vx = 0.0, vy = 0.0, vz = 0.0;
    for (long s = 0; s < ns; ++s) {
      const double* a = y0 + 3 * s;
      const double* b = y1 + 3 * s;
      const double r0x = b[0] - a[0], r0y = b[1] - a[1], r0z = b[2] - a[2];
      const double r1x = px - a[0], r1y = py - a[1], r1z = pz - a[2];
      const double r2x = px - b[0], r2y = py - b[1], r2z = pz - b[2];
      const double i1 = 1.0 / sqrt(r1x * r1x + r1y * r1y + r1z * r1z + eps * eps);
      const double i2 = 1.0 / sqrt(r2x * r2x + r2y * r2y + r2z * r2z + eps * eps);
      const double dx = r1x * i1 - r2x * i2, dy = r1y * i1 - r2y * i2, dz = r1z * i1 - r2z * i2;
      const double cx = r1y * r2z - r1z * r2y, cy = r1z * r2x - r1x * r2z, cz = r1x * r2y - r1y * r2x;
      const double r0sq = r0x * r0x + r0y * r0y + r0z * r0z;
      const double den = cx * cx + cy * cy + cz * cz + eps * r0sq * r0sq;
      const double q = g[s] * (r0x * dx + r0y * dy + r0z * dz) * inv4pi / den;
      vx += cx * q; vy += cy * q; vz += cz * q;
    }
    out[3 * t] = vx; out[3 * t + 1] = vy; out[3 * t + 2] = vz;
  }
}
