// Post-reconstruction step on a patch volume in HBM:
//   LFM_reconstruction_module.m:45-53  top-layer balance, per-layer max normalisation
//   LFM_reconstruction_module.m:60-62  global max normalisation, and the TIFF image
//                                      im2uint16(recon_stack / max(recon_stack(:)))
//   mesolf.m:101-102                   the central valid_recon_range planes
// All arithmetic is FP64 on the FP32 volume values (the reference works on the
// double array the reconstruction returns); maxima omit NaN like MATLAB's max.
//
// Volume maxima are reductions over the lenslet-blocked volume (rows of kXRow
// floats, phases < kP only); the value chain of a voxel x of plane z is
//   b = z == 0 ? (x * m[1]) / m[0] : x           (balance; m = plane maxima of X)
//   n = b / M[z]                                  (per layer; M = plane maxima of b)
//   o = n / g,  g = max over the volume of n      (global)
//   tiff = im2uint16(o / g2), g2 = max of o
// with MATLAB's left-to-right evaluation order.  Coefficients on the device:
// c[0, Z) = m, c[Z, 2Z) = M, c[2Z, 3Z) = plane maxima of n, c[3Z, 4Z) = plane
// maxima of o.
#include "internal.h"

namespace mlf {

namespace {

// max that ignores NaN (NaN only when both are NaN), as MATLAB's max
__device__ __forceinline__ double nan_max(double a, double b) {
  if (a != a) return b;
  if (b != b) return a;
  return a > b ? a : b;
}

__device__ __forceinline__ double fold_max(const double* v, int n) {
  double m = __builtin_nan("");
  for (int k = 0; k < n; ++k) m = nan_max(m, v[k]);
  return m;
}

template <int STAGE>
__device__ __forceinline__ double post_value(float x, int z, const double* __restrict__ c, int Z,
                                             double g) {
  const double v = x;
  if constexpr (STAGE == 0) return v;
  const double b = z == 0 ? (v * c[1]) / c[0] : v;
  if constexpr (STAGE == 1) return b;
  const double n = b / c[Z + z];
  if constexpr (STAGE == 2) return n;
  return n / g;
}

// im2uint16 of a double: scale by 65535 and round half away from zero,
// saturating at 0 and 65535; NaN -> 0
__device__ __forceinline__ uint16_t im2uint16(double v) {
  if (!(v > 0.0)) return 0;
  if (v >= 1.0) return 65535;
  return (uint16_t)round(v * 65535.0);
}

__device__ double block_max_f64(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v = nan_max(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double m = __builtin_nan("");
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) m = nan_max(m, sh[k]);
  return m;  // valid in thread 0
}

// Per-block maxima of the STAGE value over plane z0 + blockIdx.y.
template <int STAGE>
__global__ __launch_bounds__(256) void k_post_max(const float* __restrict__ X, int nlp, int z0,
                                                  const double* __restrict__ c, int Z,
                                                  double* __restrict__ parts) {
  __shared__ double sh[4];
  const int z = z0 + blockIdx.y;
  const double g = STAGE == 3 ? fold_max(c + 2 * Z, Z) : 0.0;
  const float* Xz = X + (size_t)z * nlp * kXRow;
  const size_t n = (size_t)nlp * kP;
  double m = __builtin_nan("");
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t l = e / kP;
    const int ph = (int)(e - l * kP);
    m = nan_max(m, post_value<STAGE>(Xz[l * kXRow + ph], z, c, Z, g));
  }
  m = block_max_f64(m, sh);
  if (threadIdx.x == 0) parts[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = m;
}

// Plane maxima from the per-block parts: one workgroup per plane z0 + blockIdx.x.
__global__ __launch_bounds__(256) void k_post_reduce(const double* __restrict__ parts, int nb,
                                                     int z0, int stage, double* __restrict__ c,
                                                     int Z) {
  __shared__ double sh[4];
  const int q = blockIdx.x, z = z0 + q;
  double m = __builtin_nan("");
  for (int k = threadIdx.x; k < nb; k += blockDim.x) m = nan_max(m, parts[(size_t)q * nb + k]);
  m = block_max_f64(m, sh);
  if (threadIdx.x != 0) return;
  if (stage == 0) {
    c[z] = m;
    c[Z + z] = m;  // the balance changes plane 0 only; stage 1 replaces M[0]
  } else {
    c[stage * Z + z] = m;
  }
}

// Output pass over MATLAB column-major [I1, I2, plane] indices: the cropped
// stack out[:, :, z - zlo] for zlo <= z < zhi and, when tiff is set, the
// uint16 image of every plane.
__global__ __launch_bounds__(256) void k_post_write(const float* __restrict__ X, int n1, int n2,
                                                    int Z, const double* __restrict__ c, int zlo,
                                                    int zhi, double* __restrict__ out,
                                                    uint16_t* __restrict__ tiff) {
  const double g = fold_max(c + 2 * Z, Z), g2 = fold_max(c + 3 * Z, Z);
  const int I1 = n1 * kN, I2 = n2 * kN;
  const size_t plane = (size_t)I1 * I2;
  const int za = tiff ? 0 : zlo, zb = tiff ? Z : zhi;
  const size_t total = (size_t)(zb - za) * plane;
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < total;
       o += (size_t)gridDim.x * blockDim.x) {
    const int z = za + (int)(o / plane);
    const size_t p = o % plane;
    const int p1 = (int)(p % I1), p2 = (int)(p / I1);
    const int L1 = p1 / kN, s1 = p1 - L1 * kN, L2 = p2 / kN, s2 = p2 - L2 * kN;
    const float x = X[(((size_t)z * n2 + L2) * n1 + L1) * kXRow + s2 * kN + s1];
    const double v = post_value<3>(x, z, c, Z, g);
    if (out && z >= zlo && z < zhi) out[(size_t)(z - zlo) * plane + p] = v;
    if (tiff) tiff[(size_t)z * plane + p] = im2uint16(v / g2);
  }
}

}  // namespace

hipError_t launch_postprocess(const float* X, int n1, int n2, int Z, double* coef, double* parts,
                              int zlo, int zhi, double* out, uint16_t* tiff, hipStream_t st) {
  const int nlp = n1 * n2;
  const int nb = kPostBlocks;
  hipLaunchKernelGGL(k_post_max<0>, dim3(nb, Z), dim3(256), 0, st, X, nlp, 0, coef, Z, parts);
  hipLaunchKernelGGL(k_post_reduce, dim3(Z), dim3(256), 0, st, parts, nb, 0, 0, coef, Z);
  hipLaunchKernelGGL(k_post_max<1>, dim3(nb, 1), dim3(256), 0, st, X, nlp, 0, coef, Z, parts);
  hipLaunchKernelGGL(k_post_reduce, dim3(1), dim3(256), 0, st, parts, nb, 0, 1, coef, Z);
  hipLaunchKernelGGL(k_post_max<2>, dim3(nb, Z), dim3(256), 0, st, X, nlp, 0, coef, Z, parts);
  hipLaunchKernelGGL(k_post_reduce, dim3(Z), dim3(256), 0, st, parts, nb, 0, 2, coef, Z);
  hipLaunchKernelGGL(k_post_max<3>, dim3(nb, Z), dim3(256), 0, st, X, nlp, 0, coef, Z, parts);
  hipLaunchKernelGGL(k_post_reduce, dim3(Z), dim3(256), 0, st, parts, nb, 0, 3, coef, Z);
  if (out || tiff) {
    const size_t total = (size_t)(tiff ? Z : zhi - zlo) * n1 * kN * n2 * kN;
    size_t grid = (total + 255) / 256;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(k_post_write, dim3((unsigned)grid), dim3(256), 0, st, X, n1, n2, Z, coef,
                       zlo, zhi, out, tiff);
  }
  return hipGetLastError();
}

}  // namespace mlf
