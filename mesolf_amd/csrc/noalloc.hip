// MEX operator without device allocation: mlf_lfm_convolution_vec_noalloc.
//
// mlf_lfm_convolution_vec_device keeps the MFMA form's GEMM operand (Kq,
// mex_kq_floats(kw) floats, 59 MB at the production kernel width 241) in a
// buffer the library allocates on first use.  Callers that cannot allocate
// device memory (code under graph capture, tight serving loops) use this
// variant instead: Kq lives in `scratch`, a file-scope device array of
// kNoAllocFloats floats (64 MiB, odd kw <= 257) that every GPU the library
// runs on holds until the process ends.
//
// On the host the name `scratch` is a stand-in, not a device address: its
// address on a device is resolved once per device by hipGetSymbolAddress at
// context creation (mlf_context_create, with that device current) and only
// that address is passed to kernels.  Every call on one GPU shares its
// array, which keeps the last call's contents: the Kq build writes every
// element a launch reads, so nothing needs clearing.  Calls on one context
// are ordered by its stream; calls that overlap on one GPU from different
// contexts race on the array.
#include <mutex>

#include "internal.h"
#include "state.h"

namespace mlf {

namespace {
constexpr size_t kNoAllocFloats = (size_t)16 << 20;  // 64 MiB
constexpr int kNoAllocMaxKw = 257;                   // largest odd kw whose Kq fits
constexpr int kMaxDevices = 64;

static __device__ __align__(256) float scratch[kNoAllocFloats];

std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

mlf_status noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return set_error(MLF_ERR_UNSUPPORTED, "device %d beyond the %d devices with a scratch table", device,
                     kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return MLF_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_error(MLF_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                     hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<float*>(p);
  return MLF_OK;
}

}  // namespace mlf

using namespace mlf;

extern "C" mlf_status mlf_lfm_convolution_vec_noalloc(mlf_context ctx, const float* d_bitmap,
                                                      int64_t width, int64_t height, int64_t n,
                                                      const float* d_kernels, int64_t kw,
                                                      int64_t nphase, float* d_out) {
  clear_error();
  if (!ctx || !d_bitmap || !d_kernels || !d_out)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (width < 1 || height < 1 || n < 1)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bitmap dims must be >= 1");
  if (nphase != kN) return set_error(MLF_ERR_INVALID_ARGUMENT, "kernels must be [kw,kw,15,15]");
  if (kw < 1 || kw > 2048)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "kernel width %lld out of [1,2048]", (long long)kw);
  if (width > (1 << 30) || height > (1 << 30) || width * height > (int64_t)1 << 40)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bitmap too large");
  if (n > 65535) return set_error(MLF_ERR_INVALID_ARGUMENT, "batch too large (n > 65535)");
  if (kw % 2 == 1 && kw > kNoAllocMaxKw)
    return set_error(MLF_ERR_INVALID_ARGUMENT,
                     "kernel width %lld: the no-allocation MEX holds odd kernels up to %d wide "
                     "(its %zu-byte scratch); use mlf_lfm_convolution_vec_device",
                     (long long)kw, kNoAllocMaxKw, kNoAllocFloats * sizeof(float));
  MLF_CHECK(ctx_activate(ctx));
  if (kw % 2 == 1) {
    // resolved by mlf_context_create on this device
    float* kq = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_scratch_mu);
      kq = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch_ptr[ctx->device] : nullptr;
    }
    if (!kq) return set_error(MLF_ERR_INTERNAL, "scratch of device %d not resolved", ctx->device);
    if (mex_kq_floats((int)kw) > kNoAllocFloats)
      return set_error(MLF_ERR_INTERNAL, "scratch too small for kw %lld", (long long)kw);
    MLF_LAUNCH(launch_mex_mfma(d_bitmap, width, height, n, d_kernels, (int)kw, kq, d_out, ctx->stream));
    return MLF_OK;
  }
  // even kw: the scalar kernel, which uses no scratch
  MLF_LAUNCH(launch_mex_conv(d_bitmap, width, height, n, d_kernels, (int)kw, d_out, ctx->stream));
  return MLF_OK;
}
