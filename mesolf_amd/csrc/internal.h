// Internal declarations shared by the host runtime and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mesolf.h"

namespace mlf {

constexpr int kN = 15;          // phases per lenslet axis (MesoLF optics, mex:35)
constexpr int kP = kN * kN;     // phases per lenslet
// Floats per lenslet row of a volume in HBM: the kP phases and a zero pad, so
// that every row (and every 80-phase group of one) starts on a 64-byte
// boundary; rows of 225 floats made the update's partial-line traffic ~25 %
// slower.  Only phases < kP are read or written as data; the pad stays zero.
constexpr int kXRow = 256;
constexpr int kC0 = kN / 2;     // centre phase, 0-based

// ------------------------------------------------------------ error state
mlf_status set_error(mlf_status s, const char* fmt, ...);
mlf_status clear_error();

#define MLF_HIP(call)                                                             \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      return ::mlf::set_error(e_ == hipErrorOutOfMemory ? MLF_ERR_OUT_OF_MEMORY   \
                                                        : MLF_ERR_DEVICE,         \
                              "%s failed: %s (%s:%d)", #call,                     \
                              hipGetErrorString(e_), __FILE__, __LINE__);         \
    }                                                                             \
  } while (0)

#define MLF_CHECK(expr)                 \
  do {                                  \
    mlf_status s_ = (expr);             \
    if (s_ != MLF_OK) return s_;        \
  } while (0)

#define MLF_LAUNCH(call)                                                                      \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return ::mlf::set_error(MLF_ERR_DEVICE, "launch %s failed: %s", #call,                  \
                              hipGetErrorString(e_));                                         \
  } while (0)

// --------------------------------------------------------- device buffer
template <typename T>
struct DevBuf {
  T* ptr = nullptr;
  size_t count = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  mlf_status alloc(size_t n) {
    release();
    if (n == 0) return MLF_OK;
    hipError_t e = hipMalloc(&ptr, n * sizeof(T));
    if (e != hipSuccess) {
      ptr = nullptr;
      return set_error(MLF_ERR_OUT_OF_MEMORY, "hipMalloc(%zu bytes) failed: %s", n * sizeof(T),
                       hipGetErrorString(e));
    }
    count = n;
    return MLF_OK;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  size_t bytes() const { return count * sizeof(T); }
};

// ----------------------------------------------------------------- launchers
// All launchers enqueue on `st` and return hipGetLastError().

// MATLAB [I1,I2,Z] (column-major) <-> lenslet-blocked [z][L2][L1][s2][s1].
hipError_t launch_to_blocked(const float* src, float* dst, int n1, int n2, int Z, hipStream_t st);
hipError_t launch_from_blocked_f32(const float* src, float* dst, int n1, int n2, int Z, hipStream_t st);
hipError_t launch_from_blocked_f64(const float* src, double* dst, int n1, int n2, int Z, hipStream_t st);
hipError_t launch_fill(float* dst, size_t n, float value, hipStream_t st);
// parts[2k], parts[2k+1] = FP64 sums of (a-b)^2 and b^2 over block k's share.
hipError_t launch_diff_norms(const float* a, const float* b, size_t n, double* parts, int nblk,
                             hipStream_t st);
hipError_t launch_init_volume(const float* LF, size_t nlf, double* sum, float* X, size_t nx,
                              double inv_numel, hipStream_t st);

// Per-plane geometry of the compact tables of one view.
struct PlaneTables {
  const int32_t* r;        // [Z] lenslet radius per plane
  const int64_t* off;      // [Z] float offset of each plane's block inside one view
};

// Forward projection of one view into per-plane partial images part[z][n1*n2].
hipError_t launch_forward_fp32(const float* X, const float* W, PlaneTables t, int n1, int n2, int Z,
                               float* part, hipStream_t st);
// y = sum_z part, R = LF(u::15, v::15) ./ y with non-finite -> 0.
hipError_t launch_ratio(const float* part, int Z, int n1, int n2, const float* LF, int u, int v,
                        float* R, float* y_out, hipStream_t st);
// B = backprojection of R, B1 = backprojection of ones; multiplicative update.
hipError_t launch_backward_update_fp32(float* X, const float* Wb, PlaneTables t, const float* R,
                                       int n1, int n2, int Z, float w, hipStream_t st);
// Scatter a view image into the sensor image at (u + 15 i1, v + 15 i2).
hipError_t launch_scatter_view(const float* part, int Z, int n1, int n2, int u, int v, float* sensor,
                               hipStream_t st);

// sample_filter: separable circular convolution per filtered plane.
struct FilterPlan {
  int nplanes = 0;                 // planes with a filter
  const int32_t* plane = nullptr;  // [nplanes] plane index z
  const int32_t* n0_1 = nullptr;   // [nplanes] first tap offset along axis 0
  const int32_t* len_1 = nullptr;  // [nplanes] number of taps along axis 0
  const int64_t* off_1 = nullptr;  // [nplanes] offset into taps_1
  const float* taps_1 = nullptr;
  const int32_t* n0_2 = nullptr;
  const int32_t* len_2 = nullptr;
  const int64_t* off_2 = nullptr;
  const float* taps_2 = nullptr;
};
// maxlen: the longest filter (taps) of the plan, sizes the LDS line buffer.
hipError_t launch_filter_pass1(const float* X, float* tmp, FilterPlan fp, int n1, int n2,
                               double* part_before, int maxlen, hipStream_t st);
hipError_t launch_filter_pass2(const float* tmp, float* X, FilterPlan fp, int n1, int n2,
                               double* part_after, int maxlen, hipStream_t st);
hipError_t launch_filter_scale(float* X, FilterPlan fp, int n1, int n2, const double* part_before,
                               const double* part_after, hipStream_t st);
// FP64 partial-sum slots per filtered plane (workgroups of the larger pass).
int filter_partial_blocks(int n1, int n2);

// Post-reconstruction normalisation (LFM_reconstruction_module.m:45-53,60-62)
// and crop (mesolf.m:101-102).  coef: 4*Z + 2 doubles, parts: kPostBlocks*Z
// doubles.  out: double [I1, I2, zhi - zlo] column-major (planes zlo..zhi-1)
// or null; tiff: uint16 [I1, I2, Z] column-major or null.  Z >= 2.
constexpr int kPostBlocks = 64;  // workgroups per plane in the max passes

// Summary image, 'pixelwise' detrend (summary_image_generation_module.m:62-96)
// and rank-1 background (rank_1_factorization.m) of a movie [npix, T] single,
// column-major.  scratch: summary_scratch_doubles(npix, T) doubles; dbuf:
// npix * T floats; img, s: npix doubles (summary image normalised by its max,
// bg_spatial); bg_temporal is left at summary_temporal(scratch) [T].
size_t summary_scratch_doubles(size_t npix, int T);
// global != 0: the 'global' detrend (std image with the rank-1 background,
// summary_image_generation_module.m:43-60); dbuf, w1, w2 and p are unused.
hipError_t launch_summary(const float* Y, size_t npix, int T, int w1, int w2, int p, int bg_iter,
                          int global, double* scratch, float* dbuf, double* img, double* s,
                          hipStream_t st);
const double* summary_temporal(const double* scratch);
// Bilinear (griddedInterpolant 'linear', extrapolating) resample of n frames
// [H, W] at yq [Hq] x xq [Wq] (device arrays, 1-based pixel coordinates).
hipError_t launch_rectify(const float* in, int H, int W, int n, const double* yq, int Hq, const double* xq,
                          int Wq, float* out, hipStream_t st);
// out [h, w, T] = in(top : top+h-1, left : left+w-1, :) of in [H, W, T], 0-based corner.
hipError_t launch_crop_movie(const float* in, int64_t H, int64_t W, int64_t top, int64_t left, int h,
                             int w, int T, float* out, hipStream_t st);
hipError_t launch_postprocess(const float* X, int n1, int n2, int Z, double* coef, double* parts,
                              int zlo, int zhi, double* out, uint16_t* tiff, hipStream_t st);

// MEX operator (mex:43-266).
hipError_t launch_mex_conv(const float* in, int64_t W, int64_t H, int64_t n, const float* K, int kw,
                           float* out, hipStream_t st);
// MFMA form of the MEX operator for odd kw (clipping == zero padding); Kq is
// a scratch of mex_kq_floats(kw) floats (the per-phase-row GEMM operand).
constexpr int kMexMfmaMaxKw = 511;
size_t mex_kq_floats(int kw);
hipError_t launch_mex_mfma(const float* in, int64_t W, int64_t H, int64_t n, const float* K, int kw,
                           float* Kq, float* out, hipStream_t st);
// Device address of the no-allocation MEX scratch on `device` (the device
// must be current), resolved once per device at context creation.
mlf_status noalloc_scratch_init(int device);
// Kq of one kernel set K [kw][kw][15][15] (for launch_mex_mfma_list).
hipError_t launch_mex_kq(const float* K, int kw, float* Kq, hipStream_t st);
// out[dst[j]] += MEX(in[src[j]], K) for j < n (dst distinct within a launch).
hipError_t launch_mex_mfma_list(const float* in, int64_t W, int64_t H, int64_t n, const float* Kq,
                                int kw, const int32_t* src, const int32_t* dst, float* out,
                                hipStream_t st);
// out[s] = sum over j in [seg_ptr[s], seg_ptr[s+1]) of MEX(in[order[j]], K[plane[order[j]]]),
// added in j order; K is [Z][15][15][kw][kw].
hipError_t launch_mex_conv_segments(const float* in, int64_t W, int64_t H, const int32_t* plane,
                                    const int64_t* seg_ptr, const int32_t* order, int64_t nseg,
                                    const float* K, int kw, float* out, hipStream_t st);

}  // namespace mlf
