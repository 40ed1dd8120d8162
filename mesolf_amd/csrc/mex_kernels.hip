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

// Segment-batched form: out[s] = sum over the slices i of segment s (in
// order) of conv(slice i, kernel of plane[i]); each slice's sum is formed
// separately and then added, as the per-plane MEX calls and additions of
// gen_seed_library_local_with_ring_bg_same_size.m:285-316 do.
__global__ __launch_bounds__(kTileX* kTileY) void k_mex_conv_segments(
    const float* __restrict__ in, int W, int H, const int32_t* __restrict__ plane,
    const int64_t* __restrict__ seg_ptr, const float* __restrict__ Kall, int kw,
    float* __restrict__ out) {
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
  for (int64_t i = seg_ptr[seg]; i < seg_ptr[seg + 1]; ++i) {
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

}  // namespace

hipError_t launch_mex_conv_segments(const float* in, int64_t W, int64_t H, const int32_t* plane,
                                    const int64_t* seg_ptr, int64_t nseg, const float* K, int kw,
                                    float* out, hipStream_t st) {
  dim3 grid((unsigned)((W + kTileX - 1) / kTileX), (unsigned)((H + kTileY - 1) / kTileY),
            (unsigned)nseg);
  const size_t shmem = sizeof(float) * kTileY * (kTileX + kw);
  hipLaunchKernelGGL(k_mex_conv_segments, grid, dim3(kTileX * kTileY), shmem, st, in, (int)W,
                     (int)H, plane, seg_ptr, K, kw, out);
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
