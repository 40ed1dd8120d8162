/*
 * FP32 C/OpenMP restatement of MesoLF's CPU phase-space RL path.
 *
 * TEST INFRASTRUCTURE ONLY: the timed CPU baseline of bench.py and a checker
 * in tests/, never the product (the product is the HIP library under
 * mesolf_amd/).  Pinned against the FP64 numpy oracle in
 * tests/test_cpu_baseline.py.
 *
 * One view step (recon:107-137) as the reference CPU branch computes it:
 * full-resolution conv2(X_z, H_ps(:,:,u,v,z), 'same') per plane, summed and
 * sampled at the view's lenslet centres (forwardProjectACC_phasespace_theory,
 * recon:191-198), the ratio (recon:116-117), conv2(big, Ht(:,:,u,v,z), 'same')
 * of the ratio and of ones placed at the view's pixels
 * (backwardProject_phasespace_theory, recon:201-208), and the update
 * (recon:130-137).  225x redundant: every full-resolution output is computed
 * and 1/225 of them is used.  Images and volumes are column-major
 * [I1, I2(, Z)].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NP 225

int mlf_cpu_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* conv2(x, h, 'same') for odd K1, K2 (MATLAB): out[p] = sum_q x[q] h[p - q + m],
 * m = (K - 1) / 2, zero outside x.  Column-major x [I1, I2], h [K1, K2]. */
void mlf_cpu_conv2_same(const float* x, int I1, int I2, const float* h, int K1, int K2, float* out) {
  const int m1 = (K1 - 1) / 2, m2 = (K2 - 1) / 2;
#pragma omp parallel for schedule(static)
  for (int p2 = 0; p2 < I2; ++p2) {
    float* o = out + (size_t)p2 * I1;
    for (int p1 = 0; p1 < I1; ++p1) o[p1] = 0.f;
    for (int k2 = 0; k2 < K2; ++k2) {
      const int q2 = p2 - k2 + m2;
      if (q2 < 0 || q2 >= I2) continue;
      const float* xc = x + (size_t)q2 * I1;
      const float* hc = h + (size_t)k2 * K1;
      for (int k1 = 0; k1 < K1; ++k1) {
        const float hv = hc[k1];
        /* out[p1] += x[p1 - k1 + m1] * hv for p1 - k1 + m1 in [0, I1); zero taps
         * are not skipped: conv2's cost does not depend on the values */
        const int lo = k1 - m1 > 0 ? k1 - m1 : 0;
        const int hi = I1 + k1 - m1 < I1 ? I1 + k1 - m1 : I1;
        const float* xs = xc - k1 + m1;
#pragma omp simd
        for (int p1 = lo; p1 < hi; ++p1) o[p1] += xs[p1] * hv;
      }
    }
  }
}

static float ratio_of(float lf, double y) {
  const double r = (double)lf / y;
  return isfinite(r) ? (float)r : 0.f;
}

static float update_of(float x, double b, double b1, double w) {
  double c = (double)x * b / b1;
  if (!isfinite(c) || c < 0) c = 0; /* NaN, +-Inf and negatives -> 0 (recon:131-134) */
  double v = w * c + (1.0 - w) * x;
  return v < 0 ? 0.f : (float)v;
}

/* One literal view step.  X [I1, I2, Z] updated in place; LF [I1, I2];
 * Hv, Htv [IMG, IMG, Z]: the view's slices H_ps(:,:,u,v,:) and Ht(:,:,u,v,:);
 * (u, v) 0-based; work: 3 * I1 * I2 floats. */
void mlf_cpu_view_step_literal(float* X, const float* LF, int I1, int I2, int Z, const float* Hv,
                               const float* Htv, int IMG, int u, int v, double w, float* work) {
  const int N = 15, c0 = 7;
  const size_t npx = (size_t)I1 * I2;
  const int n1 = I1 / N, n2 = I2 / N;
  float* full = work;
  float* big = work + npx;
  float* B = work + 2 * npx;
  double* y = (double*)calloc((size_t)n1 * n2, sizeof(double));
  for (int z = 0; z < Z; ++z) {
    mlf_cpu_conv2_same(X + z * npx, I1, I2, Hv + (size_t)z * IMG * IMG, IMG, IMG, full);
    for (int i2 = 0; i2 < n2; ++i2)
      for (int i1 = 0; i1 < n1; ++i1) y[i1 + (size_t)n1 * i2] += full[(c0 + N * i1) + (size_t)I1 * (c0 + N * i2)];
  }
  /* ratio at the view's pixels; big = R there, 0 elsewhere */
  memset(big, 0, npx * sizeof(float));
  for (int i2 = 0; i2 < n2; ++i2)
    for (int i1 = 0; i1 < n1; ++i1) {
      const size_t p = (u + N * i1) + (size_t)I1 * (v + N * i2);
      big[p] = ratio_of(LF[p], y[i1 + (size_t)n1 * i2]);
    }
  float* ones = (float*)calloc(npx, sizeof(float));
  for (int i2 = 0; i2 < n2; ++i2)
    for (int i1 = 0; i1 < n1; ++i1) ones[(u + N * i1) + (size_t)I1 * (v + N * i2)] = 1.f;
  for (int z = 0; z < Z; ++z) {
    const float* ht = Htv + (size_t)z * IMG * IMG;
    mlf_cpu_conv2_same(big, I1, I2, ht, IMG, IMG, B);
    mlf_cpu_conv2_same(ones, I1, I2, ht, IMG, IMG, full);
    float* xz = X + z * npx;
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < npx; ++p) xz[p] = update_of(xz[p], B[p], full[p], w);
  }
  free(ones);
  free(y);
}
