// MEX operator: shift-variant correlation gathered by output-pixel phase.
//
//   out[x,y,n] = sum_{ky,kx < kw} in[x - h + kx, y - h + ky, n] * K[kx, ky, x%15, y%15]
//
// h = kw/2, zero outside the bitmap; the tap window is clipped exactly as
// mex_lfm_convolution_vec.cu:55-78 clips it (for even kw the last column/row
// tap is dropped near the right/bottom edge).
#include "internal.h"

namespace mlf {

namespace {

constexpr int kTileX = 64;   // output pixels per block along x (contiguous axis)
constexpr int kTileY = 4;    // output rows per block

// Each block computes a kTileX x kTileY output tile.  Rows of the input
// window are staged in LDS one kernel row at a time.
__global__ __launch_bounds__(kTileX* kTileY) void k_mex_conv(const float* __restrict__ in, int W,
                                                             int H, const float* __restrict__ K,
                                                             int kw, float* __restrict__ out) {
  extern __shared__ float sh[];  // (kTileY) rows x (kTileX + kw) input window
  const int tx = threadIdx.x % kTileX;
  const int ty = threadIdx.x / kTileX;
  const int x0 = blockIdx.x * kTileX;
  const int y0 = blockIdx.y * kTileY;
  const int img = blockIdx.z;
  const int x = x0 + tx, y = y0 + ty;
  const int h = kw / 2;
  const int win = kTileX + kw;
  const float* src = in + (size_t)img * W * H;

  // clipping (mex:55-78)
  const int kxs = (x - h < 0) ? h - x : 0;
  const int kxe = (x + h + 1 > W) ? kw - (x + h + 1 - W) : kw;
  const int kys = (y - h < 0) ? h - y : 0;
  const int kye = (y + h + 1 > H) ? kw - (y + h + 1 - H) : kw;
  const bool valid = (x < W) && (y < H);
  const float* Kp = K + (size_t)kw * kw * ((x % kN) + kN * (y % kN));
  float acc = 0.f;

  for (int ky = 0; ky < kw; ++ky) {
    // stage input rows y0+ty-h+ky, columns x0-h .. x0-h+win-1
    const int yy = y0 + ty - h + ky;
    for (int c = tx; c < win; c += kTileX) {
      const int xx = x0 - h + c;
      float v = 0.f;
      if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = src[(size_t)yy * W + xx];
      sh[ty * win + c] = v;
    }
    __syncthreads();
    if (valid && ky >= kys && ky < kye) {
      const float* row = sh + ty * win + tx;  // row[kx] = in[x - h + kx, yy]
      const float* kr = Kp + (size_t)ky * kw;
      for (int kx = kxs; kx < kxe; ++kx) acc = fmaf(row[kx], kr[kx], acc);
    }
    __syncthreads();
  }
  if (valid) out[(size_t)img * W * H + (size_t)y * W + x] = acc;
}

// Segment-batched form: out[s] = sum over the slices of segment s, taken in
// the order order[seg_ptr[s]] .. order[seg_ptr[s+1]-1], of conv(slice i,
// kernel of plane[i]); each slice's sum is formed separately and then added,
// as the per-plane MEX calls and additions of
// gen_seed_library_local_with_ring_bg_same_size.m:285-316 do.
__global__ __launch_bounds__(kTileX* kTileY) void k_mex_conv_segments(
    const float* __restrict__ in, int W, int H, const int32_t* __restrict__ plane,
    const int64_t* __restrict__ seg_ptr, const int32_t* __restrict__ order,
    const float* __restrict__ Kall, int kw, float* __restrict__ out) {
  extern __shared__ float sh[];
  const int tx = threadIdx.x % kTileX;
  const int ty = threadIdx.x / kTileX;
  const int x0 = blockIdx.x * kTileX;
  const int y0 = blockIdx.y * kTileY;
  const int seg = blockIdx.z;
  const int x = x0 + tx, y = y0 + ty;
  const int h = kw / 2;
  const int win = kTileX + kw;
  const int kxs = (x - h < 0) ? h - x : 0;
  const int kxe = (x + h + 1 > W) ? kw - (x + h + 1 - W) : kw;
  const int kys = (y - h < 0) ? h - y : 0;
  const int kye = (y + h + 1 > H) ? kw - (y + h + 1 - H) : kw;
  const bool valid = (x < W) && (y < H);
  const size_t kplane = (size_t)kw * kw * kP;
  const size_t koff = (size_t)kw * kw * ((x % kN) + kN * (y % kN));
  float total = 0.f;
  for (int64_t j = seg_ptr[seg]; j < seg_ptr[seg + 1]; ++j) {
    const int i = order[j];
    const float* src = in + (size_t)i * W * H;
    const float* Kp = Kall + (size_t)plane[i] * kplane + koff;
    float acc = 0.f;
    for (int ky = 0; ky < kw; ++ky) {
      const int yy = y0 + ty - h + ky;
      for (int c = tx; c < win; c += kTileX) {
        const int xx = x0 - h + c;
        float v = 0.f;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = src[(size_t)yy * W + xx];
        sh[ty * win + c] = v;
      }
      __syncthreads();
      if (valid && ky >= kys && ky < kye) {
        const float* row = sh + ty * win + tx;
        const float* kr = Kp + (size_t)ky * kw;
        for (int kx = kxs; kx < kxe; ++kx) acc = fmaf(row[kx], kr[kx], acc);
      }
      __syncthreads();
    }
    total += acc;
  }
  if (valid) out[(size_t)seg * W * H + (size_t)y * W + x] = total;
}

// ---------------------------------------------------------------- MFMA form
// For odd kw the clipping equals zero padding, and grouping outputs by phase
// turns the operator into GEMMs (FP32 MFMA, FP32 accumulation).  For output
// row y (q = y % 15), image n and x = 15 i + p:
//   out = sum_{ky, t} in[15 i + t - h, y - h + ky, n] * Kq[q][ky][t][p],
//   Kq[q][ky][t][p] = K[t - p, ky, p, q] for 0 <= t - p < kw (else 0),
// i.e. M = lenslet columns i, N = the 15 phases p (padded to 16), K = (ky, t).
// A workgroup takes one output row and kMexImg images (kMexImgW per wave) over
// 32 lenslet columns (two M-tiles).  Per chunk of kMexCK input rows it stages
// the images' rows and the chunk's Kq rows (T4 x 16 floats per row, read with
// 16-byte loads, shared by every wave) in LDS; each Kq fragment read from LDS
// feeds the MFMAs of both M-tiles.  Measured on the seed-library loop
// (396 x 396, kw = 241): 7.37 s with these settings (11.8 + 16 KB of LDS per
// workgroup at kw = 241); Kq fragments from L2 instead: 9.60 s with one image
// per wave, 8.27 s with two (three 9.52, four 10.76 s); two input rows per
// chunk, Kq from L2: 10.59 s.
typedef float mex_floatx4 __attribute__((ext_vector_type(4)));
constexpr int kMexImgW = 1;
constexpr int kMexImg = 8 * kMexImgW;
constexpr int kMexThreads = 64 * kMexImg / kMexImgW;  // one wave per kMexImgW images
constexpr int kMexCK = 1;
constexpr int kMexCols = 32;  // lenslet columns per workgroup (2 M-tiles)

__global__ void k_mex_kq(const float* __restrict__ K, int kw, int T4, float* __restrict__ Kq) {
  const size_t total = (size_t)kN * kw * T4 * 16;
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < total;
       o += (size_t)gridDim.x * blockDim.x) {
    const int p = (int)(o % 16);
    const int t = (int)((o / 16) % T4);
    const int ky = (int)((o / (16 * (size_t)T4)) % kw);
    const int q = (int)(o / (16 * (size_t)T4 * kw));
    const int kx = t - p;
    Kq[o] = (p < kN && kx >= 0 && kx < kw)
                ? K[(size_t)kx + (size_t)kw * ky + (size_t)kw * kw * (p + kN * q)]
                : 0.f;
  }
}

// List mode (seed-library batches): image j of the launch is input image
// src[j] and its projection is added to output image dst[j] (the dst of one
// launch are distinct, so no two workgroups add to the same pixel).
__global__ __launch_bounds__(kMexThreads) void k_mex_mfma(const float* __restrict__ in, int W, int H, int n,
                                                  const float* __restrict__ Kq, int kw, int T4,
                                                  float* __restrict__ out,
                                                  const int32_t* __restrict__ src,
                                                  const int32_t* __restrict__ dst) {
  extern __shared__ float xs[];  // [kMexCK][kMexImg][XL], then the chunk's Kq rows [kMexCK][T4][16]
  const int y = blockIdx.x, ng = blockIdx.y, ib = blockIdx.z;
  const int q = y % kN, h = kw / 2;
  const int i0 = kMexCols * ib;
  const int xlo = kN * i0 - h;
  const int XL = kN * kMexCols + T4;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, lr = lane & 15, lg = lane >> 4;
  const int img0 = ng * kMexImg + kMexImgW * w;
  mex_floatx4 acc[kMexImgW][2];
#pragma unroll
  for (int m = 0; m < kMexImgW; ++m) acc[m][0] = acc[m][1] = mex_floatx4{0.f, 0.f, 0.f, 0.f};
  const float* Kqq = Kq + (size_t)q * kw * T4 * 16;
  for (int c0 = 0; c0 < kw; c0 += kMexCK) {
    __syncthreads();  // previous chunk consumed
    for (int o = tid; o < kMexCK * kMexImg * XL; o += kMexThreads) {
      const int xi = o % XL;
      const int im = (o / XL) % kMexImg;
      const int kk = o / (XL * kMexImg);
      const int r = y - h + c0 + kk, x = xlo + xi, nn = ng * kMexImg + im;
      float v = 0.f;
      if (c0 + kk < kw && r >= 0 && r < H && x >= 0 && x < W && nn < n)
        v = in[(size_t)(src ? src[nn] : nn) * W * H + (size_t)r * W + x];
      xs[o] = v;
    }
    {
      float* Bs = xs + kMexCK * kMexImg * XL;
      const mex_floatx4* src4 = reinterpret_cast<const mex_floatx4*>(Kqq + (size_t)c0 * T4 * 16);
      for (int o = tid; o < kMexCK * T4 * 4; o += kMexThreads)
        reinterpret_cast<mex_floatx4*>(Bs)[o] = (c0 + o / (T4 * 4) < kw) ? src4[o] : mex_floatx4{0.f, 0.f, 0.f, 0.f};
    }
    __syncthreads();
    for (int kk = 0; kk < kMexCK && c0 + kk < kw; ++kk) {
      const float* Bq = xs + kMexCK * kMexImg * XL + (size_t)kk * T4 * 16 + lg * 16 + lr;
      const float* Ar = xs + (kk * kMexImg + kMexImgW * w) * XL + kN * lr + lg;
#pragma unroll 8
      for (int s4 = 0; s4 < T4; s4 += 4) {
        const float b = Bq[s4 * 16];
#pragma unroll
        for (int m = 0; m < kMexImgW; ++m) {
          acc[m][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ar[m * XL + s4], b, acc[m][0], 0, 0, 0);
          acc[m][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ar[m * XL + s4 + kN * 16], b, acc[m][1], 0, 0, 0);
        }
      }
    }
  }
  if (lr >= kN) return;
#pragma unroll
  for (int m = 0; m < kMexImgW; ++m) {
    const int img = img0 + m;
    if (img >= n) break;
    float* o = out + (size_t)(dst ? dst[img] : img) * W * H + (size_t)y * W;
#pragma unroll
    for (int it = 0; it < 2; ++it)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int x = kN * (i0 + 16 * it + 4 * lg + j) + lr;
        if (x < W) {
          if (dst) o[x] += acc[m][it][j];
          else o[x] = acc[m][it][j];
        }
      }
  }
}

}  // namespace

size_t mex_kq_floats(int kw) {
  const int T4 = (kw + kN - 1 + 3) / 4 * 4;
  return (size_t)kN * kw * T4 * 16;
}

hipError_t launch_mex_kq(const float* K, int kw, float* Kq, hipStream_t st) {
  const int T4 = (kw + kN - 1 + 3) / 4 * 4;
  hipLaunchKernelGGL(k_mex_kq, dim3(1024), dim3(256), 0, st, K, kw, T4, Kq);
  return hipGetLastError();
}

hipError_t launch_mex_mfma_list(const float* in, int64_t W, int64_t H, int64_t n, const float* Kq,
                                int kw, const int32_t* src, const int32_t* dst, float* out,
                                hipStream_t st) {
  const int T4 = (kw + kN - 1 + 3) / 4 * 4;
  const int XL = kN * kMexCols + T4;
  const size_t shmem = sizeof(float) * kMexCK * (kMexImg * XL + T4 * 16);
  static hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mex_mfma),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (attr != hipSuccess) return attr;
  dim3 grid((unsigned)H, (unsigned)((n + kMexImg - 1) / kMexImg),
            (unsigned)((W / kN + 1 + kMexCols - 1) / kMexCols));
  hipLaunchKernelGGL(k_mex_mfma, grid, dim3(kMexThreads), shmem, st, in, (int)W, (int)H, (int)n, Kq, kw, T4,
                     out, src, dst);
  return hipGetLastError();
}

hipError_t launch_mex_mfma(const float* in, int64_t W, int64_t H, int64_t n, const float* K, int kw,
                           float* Kq, float* out, hipStream_t st) {
  const int T4 = (kw + kN - 1 + 3) / 4 * 4;
  hipLaunchKernelGGL(k_mex_kq, dim3(1024), dim3(256), 0, st, K, kw, T4, Kq);
  const int XL = kN * kMexCols + T4;
  const size_t shmem = sizeof(float) * kMexCK * (kMexImg * XL + T4 * 16);
  static hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mex_mfma),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (attr != hipSuccess) return attr;
  dim3 grid((unsigned)H, (unsigned)((n + kMexImg - 1) / kMexImg),
            (unsigned)((W / kN + 1 + kMexCols - 1) / kMexCols));
  hipLaunchKernelGGL(k_mex_mfma, grid, dim3(kMexThreads), shmem, st, in, (int)W, (int)H, (int)n, Kq, kw, T4,
                     out, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t launch_mex_conv_segments(const float* in, int64_t W, int64_t H, const int32_t* plane,
                                    const int64_t* seg_ptr, const int32_t* order, int64_t nseg,
                                    const float* K, int kw, float* out, hipStream_t st) {
  dim3 grid((unsigned)((W + kTileX - 1) / kTileX), (unsigned)((H + kTileY - 1) / kTileY),
            (unsigned)nseg);
  const size_t shmem = sizeof(float) * kTileY * (kTileX + kw);
  hipLaunchKernelGGL(k_mex_conv_segments, grid, dim3(kTileX * kTileY), shmem, st, in, (int)W,
                     (int)H, plane, seg_ptr, order, K, kw, out);
  return hipGetLastError();
}

hipError_t launch_mex_conv(const float* in, int64_t W, int64_t H, int64_t n, const float* K, int kw,
                           float* out, hipStream_t st) {
  dim3 grid((unsigned)((W + kTileX - 1) / kTileX), (unsigned)((H + kTileY - 1) / kTileY),
            (unsigned)n);
  const size_t shmem = sizeof(float) * kTileY * (kTileX + kw);
  hipLaunchKernelGGL(k_mex_conv, grid, dim3(kTileX * kTileY), shmem, st, in, (int)W, (int)H, K, kw,
                     out);
  return hipGetLastError();
}

}  // namespace mlf
