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
// maxima of o, c[4Z] = g, c[4Z + 1] = g2.
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
  return g == 1.0 ? n : n / g;  // x / 1 == x exactly; g is 1 unless no plane has a positive max
}

// True for a positive finite value.  Dividing by (or multiplying by) such a
// value rounds monotonically, so the max of x / m over a plane is (max x) / m
// exactly; the max passes that would only confirm this are skipped (every
// workgroup of the pass exits) and the plane maxima are computed from the
// maxima already known.  Any other divisor (0, negative, infinite, NaN) takes
// the full reduction.
__device__ __forceinline__ bool pos_finite(double v) { return v > 0.0 && v < __builtin_inf(); }

// whether the STAGE pass over plane z is implied by known maxima (see pos_finite)
template <int STAGE>
__device__ __forceinline__ bool implied(const double* __restrict__ c, int Z, int z) {
  if constexpr (STAGE == 1) return pos_finite(c[0]) && pos_finite(c[1]);
  if constexpr (STAGE == 2) return pos_finite(c[Z + z]);
  if constexpr (STAGE == 3) return pos_finite(c[4 * Z]);
  return false;
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

// Per-block maxima of the STAGE value over plane z0 + blockIdx.y.  A wave
// reads one lenslet row (kXRow floats, 16 bytes per lane) per step; lanes
// keep the phases < kP of their four.
template <int STAGE>
__global__ __launch_bounds__(256) void k_post_max(const float* __restrict__ X, int nlp, int z0,
                                                  const double* __restrict__ c, int Z,
                                                  double* __restrict__ parts) {
  __shared__ double sh[4];
  const int z = z0 + blockIdx.y;
  if (implied<STAGE>(c, Z, z)) return;  // uniform over the workgroup
  const double g = STAGE == 3 ? c[4 * Z] : 0.0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const float4* Xz = reinterpret_cast<const float4*>(X + (size_t)z * nlp * kXRow);
  double m = __builtin_nan("");
  for (int l = blockIdx.x * 4 + w; l < nlp; l += gridDim.x * 4) {
    const float4 v = Xz[(size_t)l * (kXRow / 4) + lane];
    const float xv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (4 * lane + j < kP) m = nan_max(m, post_value<STAGE>(xv[j], z, c, Z, g));
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
  // maxima implied by monotone rounding: the value chain at the known maximum
  bool imp = false;
  double m = __builtin_nan("");
  if (stage == 1 && implied<1>(c, Z, z)) {
    imp = true;
    m = (c[0] * c[1]) / c[0];
  } else if (stage == 2 && implied<2>(c, Z, z)) {
    imp = true;
    m = c[Z + z] / c[Z + z];
  } else if (stage == 3 && implied<3>(c, Z, z)) {
    imp = true;
    m = c[2 * Z + z] / c[4 * Z];
  }
  if (imp) {  // uniform over the workgroup
    if (threadIdx.x == 0) c[stage * Z + z] = m;
    return;
  }
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

// c[4Z + k] = max over the planes of c[(2 + k)Z, (3 + k)Z): g (k = 0), g2 (k = 1)
__global__ void k_post_fold(double* __restrict__ c, int Z, int k) {
  if (threadIdx.x == 0) c[4 * Z + k] = fold_max(c + (2 + k) * Z, Z);
}

// Output pass, one MATLAB column-major row [0, I1) of (p2 = blockIdx.x, plane
// z = z0 + blockIdx.y) per workgroup: the cropped stack out[:, p2, z - zlo]
// for zlo <= z < zhi and, when tiff is set, the uint16 image of every plane.
// A row reads 15 phases of each lenslet (60 bytes); the 15 rows p2 of one
// lenslet row run in neighbouring workgroups and share its lines in L2.
__global__ __launch_bounds__(256) void k_post_write(const float* __restrict__ X, int n1, int n2,
                                                    int Z, const double* __restrict__ c, int z0,
                                                    int zlo, int zhi, double* __restrict__ out,
                                                    uint16_t* __restrict__ tiff) {
  const double g = c[4 * Z], g2 = c[4 * Z + 1];
  const int p2 = blockIdx.x, z = z0 + blockIdx.y;
  const int I1 = n1 * kN, I2 = n2 * kN;
  const int L2 = p2 / kN, s2 = p2 - L2 * kN;
  const float* row = X + ((size_t)z * n2 + L2) * n1 * kXRow + s2 * kN;
  double* orow = out && z >= zlo && z < zhi ? out + ((size_t)(z - zlo) * I2 + p2) * I1 : nullptr;
  uint16_t* trow = tiff ? tiff + ((size_t)z * I2 + p2) * I1 : nullptr;
  for (int p1 = threadIdx.x; p1 < I1; p1 += blockDim.x) {
    const int L1 = p1 / kN, s1 = p1 - L1 * kN;
    const double v = post_value<3>(row[(size_t)L1 * kXRow + s1], z, c, Z, g);
    if (orow) orow[p1] = v;
    if (trow) trow[p1] = im2uint16(g2 == 1.0 ? v : v / g2);
  }
}

}  // namespace

hipError_t launch_postprocess(const float* X, int n1, int n2, int Z, double* coef, double* parts,
                              int zlo, int zhi, double* out, uint16_t* tiff, hipStream_t st) {
  const int nlp = n1 * n2;
  const int nb = kPostBlocks;
  hipLaunchKernelGGL(k_post_max<0>, dim3(nb, Z), dim3(256), 0, st, X, nlp, 0, coef, Z, parts);
  hipLaunchKernelGGL(k_post_reduce, dim3(Z), dim3(256), 0, st, parts, nb, 0, 0, coef, Z);
  // plane 0 alone: more workgroups (the parts buffer holds nb * Z)
  const int nb1 = nb * (Z < 16 ? Z : 16);
  hipLaunchKernelGGL(k_post_max<1>, dim3(nb1, 1), dim3(256), 0, st, X, nlp, 0, coef, Z, parts);
  hipLaunchKernelGGL(k_post_reduce, dim3(1), dim3(256), 0, st, parts, nb1, 0, 1, coef, Z);
  hipLaunchKernelGGL(k_post_max<2>, dim3(nb, Z), dim3(256), 0, st, X, nlp, 0, coef, Z, parts);
  hipLaunchKernelGGL(k_post_reduce, dim3(Z), dim3(256), 0, st, parts, nb, 0, 2, coef, Z);
  hipLaunchKernelGGL(k_post_fold, dim3(1), dim3(64), 0, st, coef, Z, 0);
  hipLaunchKernelGGL(k_post_max<3>, dim3(nb, Z), dim3(256), 0, st, X, nlp, 0, coef, Z, parts);
  hipLaunchKernelGGL(k_post_reduce, dim3(Z), dim3(256), 0, st, parts, nb, 0, 3, coef, Z);
  hipLaunchKernelGGL(k_post_fold, dim3(1), dim3(64), 0, st, coef, Z, 1);
  if (out || tiff) {
    const int z0 = tiff ? 0 : zlo, nz = tiff ? Z : zhi - zlo;  // planes written
    hipLaunchKernelGGL(k_post_write, dim3(n2 * kN, nz), dim3(256), 0, st, X, n1, n2, Z, coef, z0,
                       zlo, zhi, out, tiff);
  }
  return hipGetLastError();
}

}  // namespace mlf
