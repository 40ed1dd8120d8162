// MATLAB MEX gateway: out = mex_lfm_convolution_vec(bitmap, kernels)
//
// Drop-in for src/cuda/mex_lfm_convolution_vec.cu:276 (mexFunction) on AMD
// GPUs: same name, arguments, output and error identifier.  bitmap is a
// single 2-D [W,H] or 3-D [W,H,n] array, kernels a single [kw,kw,15,15]
// array; the output has the bitmap's size and class.  Host arrays in and out
// (MATLAB gpuArray is CUDA-only); the device is MESOLF_DEVICE (0-based,
// default 0).  The computation is libmesolf's mlf_lfm_convolution_vec.
#include <cstdlib>

#include "mex.h"
#include "mesolf.h"

static const char* const kErrId = "parallel:gpu:mexGPUExample:InvalidInput";
static const char* const kErrMsg = "Invalid input to mex_lfm_convolution_vec()";

static int device_from_env() {
  const char* s = std::getenv("MESOLF_DEVICE");
  return s ? std::atoi(s) : 0;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  (void)nlhs;
  if (nrhs != 2) mexErrMsgIdAndTxt(kErrId, kErrMsg);  // mex:305
  const mxArray* bitmap = prhs[0];
  const mxArray* kernels = prhs[1];
  const mwSize nd = mxGetNumberOfDimensions(bitmap);
  if (nd <= 1 || nd > 3) mexErrMsgIdAndTxt(kErrId, kErrMsg);  // mex:313
  if (mxGetNumberOfDimensions(kernels) != 4) mexErrMsgIdAndTxt(kErrId, kErrMsg);  // mex:337
  if (!mxIsSingle(bitmap) || !mxIsSingle(kernels) || mxIsComplex(bitmap) || mxIsComplex(kernels))
    mexErrMsgIdAndTxt(kErrId, kErrMsg);  // mex:353
  const mwSize* bd = mxGetDimensions(bitmap);
  const mwSize* kd = mxGetDimensions(kernels);
  const int64_t W = (int64_t)bd[0], H = (int64_t)bd[1], n = nd == 3 ? (int64_t)bd[2] : 1;
  if (n < 1) mexErrMsgIdAndTxt(kErrId, kErrMsg);  // mex:332
  // square kernels of 15 x 15 phases (assumed but unchecked at mex:348)
  if (kd[0] != kd[1] || kd[2] != 15 || kd[3] != 15) mexErrMsgIdAndTxt(kErrId, kErrMsg);
  plhs[0] = mxCreateNumericArray(nd, bd, mxSINGLE_CLASS, mxREAL);
  const mlf_status st = mlf_lfm_convolution_vec(
      static_cast<const float*>(mxGetData(bitmap)), W, H, n,
      static_cast<const float*>(mxGetData(kernels)), (int64_t)kd[0], 15,
      static_cast<float*>(mxGetData(plhs[0])), device_from_env());
  if (st == MLF_ERR_INVALID_ARGUMENT) mexErrMsgIdAndTxt(kErrId, kErrMsg);
  if (st != MLF_OK) mexErrMsgIdAndTxt("mesolf:device", "%s", mlf_last_error());
}
