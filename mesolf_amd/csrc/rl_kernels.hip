// Phase-space Richardson-Lucy kernels (FP32 reference kernels).
//
// Volume layout in HBM ("lenslet-blocked"): X[z][L2][L1][s2][s1], i.e. the
// 225 phases of one lenslet are contiguous, in rows of kXRow floats whose
// pad stays zero; pixel (p1, p2) of the MATLAB array is (15*L1 + s1, 15*L2 + s2).  The compact PSF tables are
// [view][z][e][phase] with e = (d2 + r)*(2r+1) + (d1 + r).
#include "internal.h"

namespace mlf {

namespace {

__device__ __forceinline__ size_t blk_index(int p1, int p2, int z, int n1, int n2) {
  const int L1 = p1 / kN, s1 = p1 - L1 * kN;
  const int L2 = p2 / kN, s2 = p2 - L2 * kN;
  return (((size_t)z * n2 + L2) * n1 + L1) * kXRow + s2 * kN + s1;
}

__global__ void k_to_blocked(const float* __restrict__ src, float* __restrict__ dst, int n1, int n2,
                             int Z) {
  const size_t total = (size_t)n1 * n2 * Z * kP;
  const int I1 = n1 * kN, I2 = n2 * kN;
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < total;
       o += (size_t)gridDim.x * blockDim.x) {
    size_t t = o;
    const int s1 = t % kN; t /= kN;
    const int s2 = t % kN; t /= kN;
    const int L1 = t % n1; t /= n1;
    const int L2 = t % n2; t /= n2;
    const int z = (int)t;
    const int p1 = L1 * kN + s1, p2 = L2 * kN + s2;
    dst[(o / kP) * kXRow + s2 * kN + s1] = src[(size_t)p1 + (size_t)I1 * (p2 + (size_t)I2 * z)];
  }
}

template <typename T>
__global__ void k_from_blocked(const float* __restrict__ src, T* __restrict__ dst, int n1, int n2,
                               int Z) {
  const int I1 = n1 * kN, I2 = n2 * kN;
  const size_t total = (size_t)I1 * I2 * Z;
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < total;
       o += (size_t)gridDim.x * blockDim.x) {
    const int p1 = o % I1;
    const size_t r = o / I1;
    const int p2 = r % I2;
    const int z = (int)(r / I2);
    dst[o] = (T)src[blk_index(p1, p2, z, n1, n2)];
  }
}

__global__ void k_fill(float* dst, size_t n, float v) {
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < n;
       o += (size_t)gridDim.x * blockDim.x)
    dst[o] = v;
}

// Single-block FP64 sum of an image (fixed reduction order).
__global__ void k_image_sum(const float* __restrict__ img, size_t n, double* __restrict__ out) {
  __shared__ double sh[1024 / 64];
  double acc = 0.0;
  for (size_t o = threadIdx.x; o < n; o += blockDim.x) acc += (double)img[o];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) sh[wid] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
    *out = s;
  }
}

// Per-block FP64 partials of sum((a-b)^2) and sum(b^2) over a grid-stride
// range (fixed order within the block); the host adds the blocks in order.
__global__ void k_diff_norms(const float* __restrict__ a, const float* __restrict__ b, size_t n,
                             double* __restrict__ parts) {
  __shared__ double sh[2][256 / 64];
  double dd = 0.0, bb = 0.0;
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < n;
       o += (size_t)gridDim.x * blockDim.x) {
    const double x = a[o], y = b[o];
    dd += (x - y) * (x - y);
    bb += y * y;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) {
    dd += __shfl_down(dd, off, 64);
    bb += __shfl_down(bb, off, 64);
  }
  if (lane == 0) {
    sh[0][wid] = dd;
    sh[1][wid] = bb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
      s0 += sh[0][k];
      s1 += sh[1][k];
    }
    parts[2 * blockIdx.x] = s0;
    parts[2 * blockIdx.x + 1] = s1;
  }
}

// X = (1/numel) * sum(LFIMG) everywhere (recon:72-73).
__global__ void k_fill_from_sum(float* __restrict__ dst, size_t n, const double* __restrict__ sum,
                                double inv_numel) {
  const float v = (float)(inv_numel * (*sum));
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < n;
       o += (size_t)gridDim.x * blockDim.x)
    dst[o] = o % kXRow < kP ? v : 0.f;  // volume rows: the pad stays zero
}

// One thread per (output lenslet i, plane z):
//   part[z][i] = sum_{d,phase} X[z][i-d][phase] * W[z][d][phase].
__global__ void k_forward_fp32(const float* __restrict__ X, const float* __restrict__ W,
                               PlaneTables t, int n1, int n2, float* __restrict__ part) {
  const int z = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n1 * n2) return;
  const int i1 = i % n1, i2 = i / n1;
  const int r = t.r[z];
  const int K = 2 * r + 1;
  const float* Wz = W + t.off[z];
  float acc = 0.f;
  for (int d2 = -r; d2 <= r; ++d2) {
    const int L2 = i2 - d2;
    if (L2 < 0 || L2 >= n2) continue;
    for (int d1 = -r; d1 <= r; ++d1) {
      const int L1 = i1 - d1;
      if (L1 < 0 || L1 >= n1) continue;
      const float* x = X + (((size_t)z * n2 + L2) * n1 + L1) * kXRow;
      const float* w = Wz + ((size_t)(d2 + r) * K + (d1 + r)) * kP;
#pragma unroll 5
      for (int p = 0; p < kP; ++p) acc = fmaf(x[p], w[p], acc);
    }
  }
  part[(size_t)z * n1 * n2 + i] = acc;
}

__global__ void k_ratio(const float* __restrict__ part, int Z, int n1, int n2,
                        const float* __restrict__ LF, int u, int v, float* __restrict__ R,
                        float* __restrict__ y_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = n1 * n2;
  if (i >= n) return;
  double y = 0.0;
  for (int z = 0; z < Z; ++z) y += (double)part[(size_t)z * n + i];
  const float yf = (float)y;
  const int i1 = i % n1, i2 = i / n1;
  if (LF) {
    const float lf = LF[(size_t)(u + kN * i1) + (size_t)(kN * n1) * (v + kN * i2)];
    float q = lf / yf;
    if (!isfinite(q)) q = 0.f;
    R[i] = q;
  }
  if (y_out) y_out[i] = yf;
}

__global__ void k_scatter_view(const float* __restrict__ part, int Z, int n1, int n2, int u, int v,
                               float* __restrict__ sensor) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = n1 * n2;
  if (i >= n) return;
  double y = 0.0;
  for (int z = 0; z < Z; ++z) y += (double)part[(size_t)z * n + i];
  const int i1 = i % n1, i2 = i / n1;
  sensor[(size_t)(u + kN * i1) + (size_t)(kN * n1) * (v + kN * i2)] = (float)y;
}

// One block per (lenslet a, plane z), one thread per phase:
//   B  = sum_{d} R[a-d] * Wb[z][d][phase]      (backprojection of the ratio)
//   B1 = sum_{d: a-d inside} Wb[z][d][phase]   (backprojection of ones)
//   C = X*B/B1 (NaN/Inf -> 0, <0 -> 0);  X = w*C + (1-w)*X;  X<0 -> 0.
__global__ void k_backward_update_fp32(float* __restrict__ X, const float* __restrict__ Wb,
                                       PlaneTables t, const float* __restrict__ R, int n1, int n2,
                                       float w) {
  const int p = threadIdx.x;
  if (p >= kP) return;
  const int a = blockIdx.x;
  const int z = blockIdx.y;
  const int a1 = a % n1, a2 = a / n1;
  const int r = t.r[z];
  const int K = 2 * r + 1;
  const float* Wz = Wb + t.off[z];
  float B = 0.f, B1 = 0.f;
  for (int d2 = -r; d2 <= r; ++d2) {
    const int j2 = a2 - d2;
    if (j2 < 0 || j2 >= n2) continue;
    for (int d1 = -r; d1 <= r; ++d1) {
      const int j1 = a1 - d1;
      if (j1 < 0 || j1 >= n1) continue;
      const float wv = Wz[((size_t)(d2 + r) * K + (d1 + r)) * kP + p];
      B = fmaf(R[j1 + n1 * j2], wv, B);
      B1 += wv;
    }
  }
  const size_t idx = (((size_t)z * n2 + a2) * n1 + a1) * kXRow + p;
  const float x = X[idx];
  float C = x * B / B1;
  if (isnan(C) || isinf(C) || C < 0.f) C = 0.f;
  float xn = C * w + (1.f - w) * x;
  if (xn < 0.f) xn = 0.f;
  X[idx] = xn;
}

// ---------------------------------------------------------------- filter
constexpr int kFilterThreads = 256;

__device__ __forceinline__ double block_sum_f64(double v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
  __syncthreads();
  return s;
}

// Pass 1: tmp[q][p2][p1] = sum_j k1[j] * X[(p1 - n0 - j) mod I1, p2] for each
// filtered plane q; also per-block FP64 partial sums of the input plane.
// One separable pass of sample_filter (recon:222-240) along axis AXIS (1: the
// first, contiguous image axis; 2: the second), circular:
//   out[p] = sum_j k[j] * in[(p - n0 - j) mod I]
// A workgroup owns kLineGroup lenslets along the axis at one lenslet position
// of the other axis: it copies the lenslet blocks it needs (whole 225-float
// blocks, including the wrap-around halo) into LDS and each thread produces 8
// consecutive outputs of one phase row with a sliding register window (one
// LDS read per tap).  Taps are summed in the same order as the direct form.
// Pass 1 reads the volume plane and writes the temporary (same lenslet-blocked
// layout) and sums its input; pass 2 reads the temporary, writes the volume
// and sums its output; both sums are FP64 per workgroup.
constexpr int kLineGroup = 8;  // output lenslets per workgroup along the axis

__device__ __forceinline__ int floor_div(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

template <int AXIS>
__global__ __launch_bounds__(256) void k_filter_line(float* __restrict__ X, float* __restrict__ tmp,
                                                     FilterPlan fp, int n1, int n2,
                                                     double* __restrict__ part) {
  extern __shared__ float lds[];  // [nl][225]
  __shared__ double sh[4];
  const int q = blockIdx.z;
  const int z = fp.plane[q];
  const size_t plane = (size_t)n1 * n2 * kXRow;  // X and tmp share the row layout
  const float* src = AXIS == 1 ? X + (size_t)z * plane : tmp + (size_t)q * plane;
  float* dst = AXIS == 1 ? tmp + (size_t)q * plane : X + (size_t)z * plane;
  const int na = AXIS == 1 ? n1 : n2;
  const int Ia = na * kN;
  const int lo = blockIdx.y;  // lenslet on the other axis
  const int n0 = AXIS == 1 ? fp.n0_1[q] : fp.n0_2[q];
  const int len = AXIS == 1 ? fp.len_1[q] : fp.len_2[q];
  const float* __restrict__ taps = AXIS == 1 ? fp.taps_1 + fp.off_1[q] : fp.taps_2 + fp.off_2[q];
  const int P0 = blockIdx.x * kLineGroup * kN;
  const int P1 = min(P0 + kLineGroup * kN, Ia);
  const int L0 = floor_div(P0 - n0 - len + 1, kN);
  const int nl = floor_div(P1 - 1 - n0, kN) - L0 + 1;
  auto lens_of = [&](int L) -> size_t {
    L %= na;
    if (L < 0) L += na;
    return AXIS == 1 ? (size_t)lo * n1 + L : (size_t)L * n1 + lo;
  };
  for (int o = threadIdx.x; o < nl * kP; o += blockDim.x) {
    const int li = o / kP;
    lds[o] = src[lens_of(L0 + li) * kXRow + (o - li * kP)];
  }
  __syncthreads();
  // unwrapped pixel u (from lenslet L0) at other-axis phase so -> LDS
  const int so = threadIdx.x % kN, g = threadIdx.x / kN;
  const int umax = nl * kN - 1;
  auto val = [&](int u) {
    u = u < 0 ? 0 : (u > umax ? umax : u);
    const int li = u / kN, sa = u - li * kN;
    return lds[li * kP + (AXIS == 1 ? so * kN + sa : sa * kN + so)];
  };
  double own = 0.0;
  if (g < kLineGroup * kN / 8) {
    const int pb = P0 + 8 * g;
    const int ub = pb - n0 - L0 * kN;
    float w[8], acc[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      w[m] = val(ub + m);
      acc[m] = 0.f;
    }
    for (int j = 0; j < len; ++j) {
      const float kj = taps[j];
#pragma unroll
      for (int m = 0; m < 8; ++m) acc[m] = fmaf(kj, w[m], acc[m]);
#pragma unroll
      for (int m = 7; m > 0; --m) w[m] = w[m - 1];
      w[0] = val(ub - j - 1);
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int p = pb + m;
      if (p >= P1) break;
      const int L = p / kN, sa = p - L * kN;
      const size_t e = lens_of(L) * kXRow + (AXIS == 1 ? so * kN + sa : sa * kN + so);
      own += AXIS == 1 ? (double)src[e] : (double)acc[m];
      dst[e] = acc[m];
    }
  }
  const double sum = block_sum_f64(own, sh);
  if (threadIdx.x == 0) part[(size_t)q * gridDim.x * gridDim.y + blockIdx.y * gridDim.x + blockIdx.x] = sum;
}

__global__ void k_filter_scale(float* __restrict__ X, FilterPlan fp, int n1, int n2,
                               const double* __restrict__ pb, const double* __restrict__ pa,
                               int nb_before, int nb_after) {
  __shared__ double scale;
  __shared__ double sh[4];
  const int q = blockIdx.y;
  const int z = fp.plane[q];
  // plane energies: strided partial sums per thread, then a fixed-order tree
  double b = 0.0, a = 0.0;
  for (int k = threadIdx.x; k < nb_before; k += blockDim.x) b += pb[(size_t)q * nb_before + k];
  for (int k = threadIdx.x; k < nb_after; k += blockDim.x) a += pa[(size_t)q * nb_after + k];
  b = block_sum_f64(b, sh);
  a = block_sum_f64(a, sh);
  if (threadIdx.x == 0) scale = b / a;
  __syncthreads();
  const double s = scale;
  const size_t plane = (size_t)n1 * n2 * kXRow;  // rows of kXRow; the pad is left alone
  float* Xz = X + (size_t)z * plane;
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < plane;
       o += (size_t)gridDim.x * blockDim.x)
    if (o % kXRow < kP) Xz[o] = (float)((double)Xz[o] * s);
}

inline int grid_for(size_t n, int threads, int cap = 65536) {
  size_t g = (n + threads - 1) / threads;
  if (g > (size_t)cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

hipError_t launch_to_blocked(const float* src, float* dst, int n1, int n2, int Z, hipStream_t st) {
  const size_t n = (size_t)n1 * n2 * Z * kP;
  hipLaunchKernelGGL(k_to_blocked, dim3(grid_for(n, 256)), dim3(256), 0, st, src, dst, n1, n2, Z);
  return hipGetLastError();
}

hipError_t launch_from_blocked_f32(const float* src, float* dst, int n1, int n2, int Z,
                                   hipStream_t st) {
  const size_t n = (size_t)n1 * n2 * Z * kP;
  hipLaunchKernelGGL(k_from_blocked<float>, dim3(grid_for(n, 256)), dim3(256), 0, st, src, dst, n1,
                     n2, Z);
  return hipGetLastError();
}

hipError_t launch_from_blocked_f64(const float* src, double* dst, int n1, int n2, int Z,
                                   hipStream_t st) {
  const size_t n = (size_t)n1 * n2 * Z * kP;
  hipLaunchKernelGGL(k_from_blocked<double>, dim3(grid_for(n, 256)), dim3(256), 0, st, src, dst,
                     n1, n2, Z);
  return hipGetLastError();
}

hipError_t launch_init_volume(const float* LF, size_t nlf, double* sum, float* X, size_t nx,
                              double inv_numel, hipStream_t st) {
  hipLaunchKernelGGL(k_image_sum, dim3(1), dim3(1024), 0, st, LF, nlf, sum);
  hipLaunchKernelGGL(k_fill_from_sum, dim3(grid_for(nx, 256)), dim3(256), 0, st, X, nx, sum,
                     inv_numel);
  return hipGetLastError();
}

hipError_t launch_diff_norms(const float* a, const float* b, size_t n, double* parts, int nblk,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_diff_norms, dim3(nblk), dim3(256), 0, st, a, b, n, parts);
  return hipGetLastError();
}

hipError_t launch_fill(float* dst, size_t n, float value, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256)), dim3(256), 0, st, dst, n, value);
  return hipGetLastError();
}

hipError_t launch_forward_fp32(const float* X, const float* W, PlaneTables t, int n1, int n2, int Z,
                               float* part, hipStream_t st) {
  dim3 grid((n1 * n2 + 127) / 128, Z);
  hipLaunchKernelGGL(k_forward_fp32, grid, dim3(128), 0, st, X, W, t, n1, n2, part);
  return hipGetLastError();
}

hipError_t launch_ratio(const float* part, int Z, int n1, int n2, const float* LF, int u, int v,
                        float* R, float* y_out, hipStream_t st) {
  hipLaunchKernelGGL(k_ratio, dim3((n1 * n2 + 255) / 256), dim3(256), 0, st, part, Z, n1, n2, LF, u,
                     v, R, y_out);
  return hipGetLastError();
}

hipError_t launch_scatter_view(const float* part, int Z, int n1, int n2, int u, int v, float* sensor,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_scatter_view, dim3((n1 * n2 + 255) / 256), dim3(256), 0, st, part, Z, n1, n2,
                     u, v, sensor);
  return hipGetLastError();
}

hipError_t launch_backward_update_fp32(float* X, const float* Wb, PlaneTables t, const float* R,
                                       int n1, int n2, int Z, float w, hipStream_t st) {
  dim3 grid(n1 * n2, Z);
  hipLaunchKernelGGL(k_backward_update_fp32, grid, dim3(256), 0, st, X, Wb, t, R, n1, n2, w);
  return hipGetLastError();
}

static int filter_groups(int n) { return (n + kLineGroup - 1) / kLineGroup; }

int filter_partial_blocks(int n1, int n2) {
  return std::max(filter_groups(n1) * n2, filter_groups(n2) * n1);
}

static size_t filter_lds(const FilterPlan& fp, int maxlen) {
  const int nl = kLineGroup + (maxlen + 2 * kN) / kN + 1;
  return sizeof(float) * (size_t)nl * kP;
}

hipError_t launch_filter_pass1(const float* X, float* tmp, FilterPlan fp, int n1, int n2,
                               double* part_before, int maxlen, hipStream_t st) {
  hipLaunchKernelGGL(k_filter_line<1>, dim3(filter_groups(n1), n2, fp.nplanes), dim3(256),
                     filter_lds(fp, maxlen), st, const_cast<float*>(X), tmp, fp, n1, n2, part_before);
  return hipGetLastError();
}

hipError_t launch_filter_pass2(const float* tmp, float* X, FilterPlan fp, int n1, int n2,
                               double* part_after, int maxlen, hipStream_t st) {
  hipLaunchKernelGGL(k_filter_line<2>, dim3(filter_groups(n2), n1, fp.nplanes), dim3(256),
                     filter_lds(fp, maxlen), st, X, const_cast<float*>(tmp), fp, n1, n2, part_after);
  return hipGetLastError();
}

hipError_t launch_filter_scale(float* X, FilterPlan fp, int n1, int n2, const double* part_before,
                               const double* part_after, hipStream_t st) {
  hipLaunchKernelGGL(k_filter_scale, dim3(256, fp.nplanes), dim3(256), 0, st, X, fp, n1, n2,
                     part_before, part_after, filter_groups(n1) * n2, filter_groups(n2) * n1);
  return hipGetLastError();
}

}  // namespace mlf
