// Summary image of a sensor movie, 'pixelwise' detrend mode
// (summary_image_generation_module.m:62-96, :130), with the rank-1 background
// of rank_1_factorization.m:31-41 and compute_std_image.m.
//
// The movie is MATLAB column-major [npix, T] single (pixel fastest), so a
// thread per pixel walking the frames reads 64 consecutive pixels per wave
// per frame.  Arithmetic is FP64 on the FP32 samples.
//   mx = max(movie);  x = movie / mx
//   d  = x - movmean(x, window_1)            (pass 1, stored FP32)
//   y  = movmean(d, window_2) .^ poly_index  (pass 2, fused with the sums)
//   img = sqrt(sum(y.^2) / T - mean(y).^2);  summary = img / max(img)
// movmean over the window [j - a, j + b] clipped to [0, T) ('shrink'),
// a = b = (k-1)/2 for odd k, a = k/2, b = k/2 - 1 for even k, kept as a
// running FP64 sum (one sample enters and one leaves per frame; the leaving
// sample is read again, from cache).
#include "internal.h"

namespace mlf {

namespace {

__device__ __forceinline__ double nmax(double a, double b) {
  if (a != a) return b;
  if (b != b) return a;
  return a > b ? a : b;
}

template <typename F>
__device__ double block_reduce(double v, double* sh, F op) {
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();  // sh may still be read by a previous reduction
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = sh[0];
  for (int k = 1; k < (int)(blockDim.x >> 6); ++k) r = op(r, sh[k]);
  return r;  // every thread
}

constexpr int kThreads = 256;

// per-block maxima of a float array (NaN ignored)
__global__ __launch_bounds__(kThreads) void k_max_f32(const float* __restrict__ a, size_t n,
                                                      double* __restrict__ parts) {
  __shared__ double sh[kThreads / 64];
  double m = __builtin_nan("");
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    m = nmax(m, (double)a[i]);
  m = block_reduce(m, sh, nmax);
  if (threadIdx.x == 0) parts[blockIdx.x] = m;
}

// same over a 16-byte aligned array: n4 float4 and a tail of ntail floats
__global__ __launch_bounds__(kThreads) void k_max_f32x4(const float4* __restrict__ a4, size_t n4,
                                                        const float* __restrict__ tail, int ntail,
                                                        double* __restrict__ parts) {
  __shared__ double sh[kThreads / 64];
  double m = __builtin_nan("");
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
    const float4 v = a4[i];
    m = nmax(nmax(m, (double)v.x), nmax(nmax((double)v.y, (double)v.z), (double)v.w));
  }
  if (blockIdx.x == 0 && (int)threadIdx.x < ntail) m = nmax(m, (double)tail[threadIdx.x]);
  m = block_reduce(m, sh, nmax);
  if (threadIdx.x == 0) parts[blockIdx.x] = m;
}

// scalar[0] = max of parts (one workgroup)
__global__ __launch_bounds__(kThreads) void k_fold_max(const double* __restrict__ parts, int n,
                                                       double* __restrict__ out) {
  __shared__ double sh[kThreads / 64];
  double m = __builtin_nan("");
  for (int k = threadIdx.x; k < n; k += blockDim.x) m = nmax(m, parts[k]);
  m = block_reduce(m, sh, nmax);
  if (threadIdx.x == 0) *out = m;
}

// rank-1 step 1, t~ = s' * Y: parts[j * nchunk + c] = sum over chunk c of s_i Y_ij.
// A thread keeps the s of its kChunk / kThreads pixels in registers for the
// kTdotFrames frames of its workgroup.
constexpr int kChunk = kThreads * 16;  // pixels per chunk
constexpr int kTdotFrames = 16;
__global__ __launch_bounds__(kThreads) void k_r1_tdot(const float* __restrict__ Y, size_t npix, int T,
                                                      const double* __restrict__ s,
                                                      double* __restrict__ parts) {
  constexpr int kPer = kChunk / kThreads;
  const size_t i0 = (size_t)blockIdx.x * kChunk + threadIdx.x;
  const int nchunk = gridDim.x;
  double sv[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const size_t i = i0 + (size_t)q * kThreads;
    sv[q] = i < npix ? s[i] : 0.0;
  }
  // frames in groups of kG: the group's loads are in flight together, one
  // pair of barriers per group
  constexpr int kG = 4;
  __shared__ double shf[kG][kThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int j0 = blockIdx.y * kTdotFrames; j0 < T; j0 += gridDim.y * kTdotFrames) {
    for (int jg = j0; jg < j0 + kTdotFrames && jg < T; jg += kG) {
      const int nf = min(kG, T - jg);
      double acc[kG];
#pragma unroll
      for (int f = 0; f < kG; ++f) acc[f] = 0.0;
#pragma unroll
      for (int q = 0; q < kPer; ++q) {
        const size_t i = i0 + (size_t)q * kThreads;
        if (i >= npix) continue;
#pragma unroll
        for (int f = 0; f < kG; ++f)
          if (f < nf) acc[f] = fma((double)Y[i + (size_t)(jg + f) * npix], sv[q], acc[f]);
      }
#pragma unroll
      for (int f = 0; f < kG; ++f) {
        double v = acc[f];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) shf[f][wid] = v;
      }
      __syncthreads();
      if ((int)threadIdx.x < nf) {
        double r = shf[threadIdx.x][0];
        for (int k = 1; k < kThreads / 64; ++k) r += shf[threadIdx.x][k];
        parts[(size_t)(jg + threadIdx.x) * nchunk + blockIdx.x] = r;
      }
      __syncthreads();  // shf is rewritten by the next group
    }
  }
}

// rank-1 step 2a, one workgroup per frame j: t_j = (sum_c parts) / div
__global__ __launch_bounds__(kThreads) void k_r1_tsum(const double* __restrict__ parts, int nchunk, int T,
                                                      const double* __restrict__ mx,
                                                      double* __restrict__ t) {
  __shared__ double sh[kThreads / 64];
  for (int j = blockIdx.x; j < T; j += gridDim.x) {
    double v = 0.0;
    for (int c = threadIdx.x; c < nchunk; c += blockDim.x) v += parts[(size_t)j * nchunk + c];
    v = block_reduce(v, sh, [](double a, double b) { return a + b; });
    if (threadIdx.x == 0) t[j] = v / *mx;
  }
}

// rank-1 step 2b (one workgroup): t /= ||t||
__global__ __launch_bounds__(kThreads) void k_r1_tnorm(int T, double* __restrict__ t) {
  __shared__ double sh[kThreads / 64];
  double ss = 0.0;
  for (int j = threadIdx.x; j < T; j += blockDim.x) ss = fma(t[j], t[j], ss);
  const double nrm = sqrt(block_reduce(ss, sh, [](double a, double b) { return a + b; }));
  for (int j = threadIdx.x; j < T; j += blockDim.x) t[j] /= nrm;
}

// rank-1 step 3: s_i = (Y t)_i / div; per-block partial sums of s_i^2
__global__ __launch_bounds__(kThreads) void k_r1_sdot(const float* __restrict__ Y, size_t npix, int T,
                                                      const double* __restrict__ t,
                                                      const double* __restrict__ mx,
                                                      double* __restrict__ s, double* __restrict__ parts) {
  __shared__ double sh[kThreads / 64];
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < npix) {
#pragma unroll 8
    for (int j = 0; j < T; ++j) v = fma((double)Y[i + (size_t)j * npix], t[j], v);
    v /= *mx;
    s[i] = v;
  }
  const double ss = block_reduce(v * v, sh, [](double a, double b) { return a + b; });
  if (threadIdx.x == 0) parts[blockIdx.x] = ss;
}

// scalar[0] = sqrt(sum of parts) (one workgroup, fixed order per thread)
__global__ __launch_bounds__(kThreads) void k_fold_norm(const double* __restrict__ parts, int n,
                                                        double* __restrict__ out) {
  __shared__ double sh[kThreads / 64];
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) v += parts[k];
  v = block_reduce(v, sh, [](double a, double b) { return a + b; });
  if (threadIdx.x == 0) *out = sqrt(v);
}

__global__ void k_fill_f64(double* __restrict__ v, size_t n, double value) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    v[i] = value;
}

__global__ void k_scale_div(double* __restrict__ v, size_t n, const double* __restrict__ d) {
  const double dv = *d;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    v[i] /= dv;
}

__device__ __forceinline__ void window_of(int k, int* a, int* b) {
  if (k & 1) {
    *a = *b = (k - 1) / 2;
  } else {
    *a = k / 2;
    *b = k / 2 - 1;
  }
}

// Frames j of a window walk: [0, j1) and [j2, T) take the clipped window
// (cnt varies, the updates are guarded); [j1, j2) is the steady part, where
// the window has all k samples and one enters and one leaves every frame.
// The steady loop is unrolled so that its loads (independent of the running
// sum) are issued ahead; the order of the additions is the same everywhere.
__device__ __forceinline__ void walk_bounds(int a, int b, int T, int* j1, int* j2) {
  *j1 = min(a, T);
  *j2 = max(*j1, T - b - 1);
}

// pass 1: d = x - movmean(x, w1), x = Y / mx
__global__ __launch_bounds__(kThreads) void k_detrend(const float* __restrict__ Y, size_t npix, int T,
                                                      int w1, const double* __restrict__ mx,
                                                      float* __restrict__ d) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= npix) return;
  // x = Y / max as a product with the reciprocal, and the steady mean as a
  // product with 1/k: within 1 ulp of the quotients, without the FP64
  // divide sequence per sample
  const double minv = 1.0 / *mx;
  int a, b;
  window_of(w1, &a, &b);
  auto x = [&](int j) { return (double)Y[i + (size_t)j * npix] * minv; };
  double S = 0.0;
  const int h0 = min(b, T - 1);
  for (int k = 0; k <= h0; ++k) S += x(k);
  auto step = [&](int j) {
    const int lo = max(j - a, 0), hi = min(j + b, T - 1);
    d[i + (size_t)j * npix] = (float)(x(j) - S / (double)(hi - lo + 1));
    if (j + b + 1 <= T - 1) S += x(j + b + 1);
    if (j - a >= 0) S -= x(j - a);
  };
  int j1, j2;
  walk_bounds(a, b, T, &j1, &j2);
  for (int j = 0; j < j1; ++j) step(j);
  const double kinv = 1.0 / (double)(a + b + 1);
#pragma unroll 8
  for (int j = j1; j < j2; ++j) {
    const double xc = x(j), xn = x(j + b + 1), xo = x(j - a);
    d[i + (size_t)j * npix] = (float)(xc - S * kinv);
    S += xn;
    S -= xo;
  }
  for (int j = j2; j < T; ++j) step(j);
}

// pass 2: y = movmean(d, w2) .^ p; img = sqrt(sum(y^2)/T - mean(y)^2)
__global__ __launch_bounds__(kThreads) void k_smooth_std(const float* __restrict__ d, size_t npix, int T,
                                                         int w2, int p, double* __restrict__ img,
                                                         double* __restrict__ parts) {
  __shared__ double sh[kThreads / 64];
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  double v = __builtin_nan("");
  if (i < npix) {
    int a, b;
    window_of(w2, &a, &b);
    auto dv = [&](int j) { return (double)d[i + (size_t)j * npix]; };
    double S = 0.0, A = 0.0, B = 0.0;
    const int h0 = min(b, T - 1);
    for (int k = 0; k <= h0; ++k) S += dv(k);
    auto acc = [&](double sm) {
      double y = 1.0;
      for (int q = 0; q < p; ++q) y *= sm;
      A = fma(y, y, A);
      B += y;
    };
    auto step = [&](int j) {
      const int lo = max(j - a, 0), hi = min(j + b, T - 1);
      acc(S / (double)(hi - lo + 1));
      if (j + b + 1 <= T - 1) S += dv(j + b + 1);
      if (j - a >= 0) S -= dv(j - a);
    };
    int j1, j2;
    walk_bounds(a, b, T, &j1, &j2);
    for (int j = 0; j < j1; ++j) step(j);
    const double kinv = 1.0 / (double)(a + b + 1);
#pragma unroll 8
    for (int j = j1; j < j2; ++j) {
      const double dn = dv(j + b + 1), dol = dv(j - a);
      acc(S * kinv);
      S += dn;
      S -= dol;
    }
    for (int j = j2; j < T; ++j) step(j);
    const double mean = B / (double)T;
    v = sqrt(A / (double)T - mean * mean);
    img[i] = v;
  }
  v = block_reduce(v, sh, nmax);
  if (threadIdx.x == 0) parts[blockIdx.x] = v;
}

// Patch movie from a sensor movie [H, W, T] (column-major): out[r + h*c + h*w*t]
// = in[(top + r) + H*(left + c) + H*W*t], one patch column per workgroup row.
__global__ __launch_bounds__(kThreads) void k_crop_movie(const float* __restrict__ in, int64_t H,
                                                         int64_t W, int64_t top, int64_t left, int h,
                                                         int w, int T, float* __restrict__ out) {
  const int c = blockIdx.y;
  for (int t = blockIdx.z; t < T; t += gridDim.z) {
    const float* src = in + (size_t)t * H * W + (size_t)(left + c) * H + top;
    float* dst = out + (size_t)t * h * w + (size_t)c * h;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < h; r += gridDim.x * blockDim.x) dst[r] = src[r];
  }
}

// 'global' detrend: compute_std_image.m as written with the rank-1
// background (y_1 = s, y_2 = t) of the movie divided by its max:
// sqrt((A - s_i^2)/T - B^2 + 2 B s_i C - s_i^2 C^2), A = sum(x^2), B = mean(x),
// C = mean(t), x = Y / mx.  A negative argument gives NaN.
__global__ __launch_bounds__(kThreads) void k_std_bg(const float* __restrict__ Y, size_t npix, int T,
                                                     const double* __restrict__ mx,
                                                     const double* __restrict__ s,
                                                     const double* __restrict__ t, double* __restrict__ img,
                                                     double* __restrict__ parts) {
  __shared__ double sh[kThreads / 64];
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  double v = __builtin_nan("");
  if (i < npix) {
    const double m = *mx;
    double A = 0.0, Bs = 0.0, Cs = 0.0;
#pragma unroll 8
    for (int j = 0; j < T; ++j) {
      const double x = (double)Y[i + (size_t)j * npix] / m;
      A = fma(x, x, A);
      Bs += x;
      Cs += t[j];
    }
    const double B = Bs / (double)T, C = Cs / (double)T, y1 = s[i];
    v = sqrt((A - y1 * y1) / (double)T - B * B + 2.0 * B * y1 * C - y1 * y1 * (C * C));
    img[i] = v;
  }
  v = block_reduce(v, sh, nmax);
  if (threadIdx.x == 0) parts[blockIdx.x] = v;
}

// Rectification resample of frames [H, W] (column-major single) at the query
// coordinates yq [Hq] x xq [Wq] (1-based pixels): griddedInterpolant's
// 'linear' interpolation and extrapolation, i.e. the bilinear form of the
// grid cell k = clamp(floor(q), 1, size-1), which extends the boundary
// cells linearly past the image.  FP64 arithmetic.
__global__ __launch_bounds__(kThreads) void k_rectify(const float* __restrict__ in, int H, int W, int n,
                                                      const double* __restrict__ yq, int Hq,
                                                      const double* __restrict__ xq, int Wq,
                                                      float* __restrict__ out) {
  // one output row i per thread and column j per workgroup row, the cell
  // and weights computed once; frames strided over the grid's z
  const int j = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Hq) return;
  const double x = xq[j], y = yq[i];
  const int kx = min(max((int)floor(x), 1), W - 1);
  const int ky = min(max((int)floor(y), 1), H - 1);
  const double tx = x - kx, ty = y - ky;
  const size_t o0 = (size_t)(kx - 1) * H + (ky - 1);  // (row ky, column kx), 1-based
  for (int f = blockIdx.z; f < n; f += gridDim.z) {
    const float* img = in + (size_t)f * H * W + o0;
    const double v0 = (1.0 - ty) * (double)img[0] + ty * (double)img[1];
    const double v1 = (1.0 - ty) * (double)img[H] + ty * (double)img[H + 1];
    out[(size_t)f * Hq * Wq + (size_t)j * Hq + i] = (float)((1.0 - tx) * v0 + tx * v1);
  }
}

inline unsigned blocks_for(size_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

size_t summary_scratch_doubles(size_t npix, int T) {
  const size_t nchunk = (npix + kChunk - 1) / kChunk;
  const size_t nblk = (npix + kThreads - 1) / kThreads;
  // mx, norm, t[T], tdot parts[T * nchunk], per-block parts[max(nblk, 1024)]
  return 2 + (size_t)T + (size_t)T * nchunk + (nblk > 1024 ? nblk : 1024);
}

hipError_t launch_summary(const float* Y, size_t npix, int T, int w1, int w2, int p, int bg_iter,
                          int global, double* scratch, float* dbuf, double* img, double* s,
                          hipStream_t st) {
  const size_t nchunk = (npix + kChunk - 1) / kChunk;
  const unsigned nblk = blocks_for(npix);
  double* mx = scratch;
  double* nrm = scratch + 1;
  double* t = scratch + 2;
  double* tparts = t + T;
  double* parts = tparts + (size_t)T * nchunk;
  // movie max (the divisor of the background and the scale of x)
  const size_t n = npix * (size_t)T;
  const unsigned gmax = 1024;
  if (((uintptr_t)Y & 15) == 0)
    hipLaunchKernelGGL(k_max_f32x4, dim3(gmax), dim3(kThreads), 0, st, reinterpret_cast<const float4*>(Y),
                       n / 4, Y + (n & ~(size_t)3), (int)(n & 3), parts);
  else
    hipLaunchKernelGGL(k_max_f32, dim3(gmax), dim3(kThreads), 0, st, Y, n, parts);
  hipLaunchKernelGGL(k_fold_max, dim3(1), dim3(kThreads), 0, st, parts, (int)gmax, mx);
  // rank-1 background: s = ones / sqrt(npix), bg_iter rounds
  hipLaunchKernelGGL(k_fill_f64, dim3(nblk < 65535 ? nblk : 65535), dim3(kThreads), 0, st, s, npix,
                     1.0 / sqrt((double)npix));
  for (int it = 1; it <= bg_iter; ++it) {
    const int fy = (T + kTdotFrames - 1) / kTdotFrames;
    hipLaunchKernelGGL(k_r1_tdot, dim3((unsigned)nchunk, fy < 65535 ? fy : 65535), dim3(kThreads), 0, st,
                       Y, npix, T, s, tparts);
    hipLaunchKernelGGL(k_r1_tsum, dim3(T < 65535 ? T : 65535), dim3(kThreads), 0, st, tparts, (int)nchunk,
                       T, mx, t);
    hipLaunchKernelGGL(k_r1_tnorm, dim3(1), dim3(kThreads), 0, st, T, t);
    hipLaunchKernelGGL(k_r1_sdot, dim3(nblk), dim3(kThreads), 0, st, Y, npix, T, t, mx, s, parts);
    if (it < bg_iter) {
      hipLaunchKernelGGL(k_fold_norm, dim3(1), dim3(kThreads), 0, st, parts, (int)nblk, nrm);
      hipLaunchKernelGGL(k_scale_div, dim3(nblk < 65535 ? nblk : 65535), dim3(kThreads), 0, st, s, npix, nrm);
    }
  }
  // summary image
  if (global) {
    hipLaunchKernelGGL(k_std_bg, dim3(nblk), dim3(kThreads), 0, st, Y, npix, T, mx, s, t, img, parts);
  } else {
    hipLaunchKernelGGL(k_detrend, dim3(nblk), dim3(kThreads), 0, st, Y, npix, T, w1, mx, dbuf);
    hipLaunchKernelGGL(k_smooth_std, dim3(nblk), dim3(kThreads), 0, st, dbuf, npix, T, w2, p, img, parts);
  }
  hipLaunchKernelGGL(k_fold_max, dim3(1), dim3(kThreads), 0, st, parts, (int)nblk, nrm);
  hipLaunchKernelGGL(k_scale_div, dim3(nblk < 65535 ? nblk : 65535), dim3(kThreads), 0, st, img, npix, nrm);
  return hipGetLastError();
}

const double* summary_temporal(const double* scratch) { return scratch + 2; }

hipError_t launch_rectify(const float* in, int H, int W, int n, const double* yq, int Hq, const double* xq,
                          int Wq, float* out, hipStream_t st) {
  const unsigned gx = (unsigned)((Hq + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_rectify, dim3(gx, (unsigned)Wq, (unsigned)(n < 16 ? n : 16)), dim3(kThreads), 0, st,
                     in, H, W, n, yq, Hq, xq, Wq, out);
  return hipGetLastError();
}

hipError_t launch_crop_movie(const float* in, int64_t H, int64_t W, int64_t top, int64_t left, int h,
                             int w, int T, float* out, hipStream_t st) {
  const unsigned gx = (unsigned)((h + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_crop_movie, dim3(gx, (unsigned)w, (unsigned)(T < 16 ? T : 16)), dim3(kThreads), 0,
                     st, in, H, W, top, left, h, w, T, out);
  return hipGetLastError();
}

}  // namespace mlf
