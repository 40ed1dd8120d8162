// MATLAB MEX gateway behind recon_theory_phase_space_LFM_with_filter.m:
//   Xguess = recon_phase_space_mex(LFIMG, H, Ht, maxIter, gpuFlag, gpu_id)
// Same arguments as src/recon/recon_theory_phase_space_LFM_with_filter.m:10.
// LFIMG, H, Ht may be single (passed through without a copy) or double
// (converted to single); gpu_id is MATLAB's 1-based device index (recon:76).
// Returns the double [I1,I2,Z] volume (recon:147 gathers a double array).
#include <cstdint>
#include <vector>

#include "mex.h"
#include "mesolf.h"

static const char* const kErrId = "mesolf:recon:InvalidInput";

// The single data of a (its own buffer when a is single, else a converted
// copy kept in `store`).
static const float* as_single(const mxArray* a, std::vector<float>* store) {
  if (mxIsSingle(a)) return static_cast<const float*>(mxGetData(a));
  const size_t n = mxGetNumberOfElements(a);
  store->resize(n);
  const double* p = static_cast<const double*>(mxGetData(a));
  for (size_t i = 0; i < n; ++i) (*store)[i] = static_cast<float>(p[i]);
  return store->data();
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  (void)nlhs;
  if (nrhs != 6) mexErrMsgIdAndTxt(kErrId, "expected (LFIMG, H, Ht, maxIter, gpuFlag, gpu_id)");
  for (int i = 0; i < 3; ++i)
    if (!(mxIsSingle(prhs[i]) || mxIsDouble(prhs[i])) || mxIsComplex(prhs[i]))
      mexErrMsgIdAndTxt(kErrId, "LFIMG, H and Ht must be real single or double arrays");
  if (mxGetNumberOfDimensions(prhs[0]) != 2) mexErrMsgIdAndTxt(kErrId, "LFIMG must be 2-D");
  const mwSize hnd = mxGetNumberOfDimensions(prhs[1]);
  if (hnd < 4 || hnd > 5 || mxGetNumberOfDimensions(prhs[2]) != hnd)
    mexErrMsgIdAndTxt(kErrId, "H and Ht must be [S,S,N,N,Z]");
  const mwSize* hd = mxGetDimensions(prhs[1]);
  const mwSize* htd = mxGetDimensions(prhs[2]);
  for (mwSize i = 0; i < hnd; ++i)
    if (hd[i] != htd[i]) mexErrMsgIdAndTxt(kErrId, "H and Ht must have the same size");
  if (hd[0] != hd[1] || hd[2] != hd[3]) mexErrMsgIdAndTxt(kErrId, "H must be [S,S,N,N,Z]");
  const int64_t Z = hnd == 5 ? (int64_t)hd[4] : 1;
  const mwSize* ld = mxGetDimensions(prhs[0]);
  const int32_t max_iter = static_cast<int32_t>(mxGetScalar(prhs[3]));
  const int32_t gpu_flag = mxGetScalar(prhs[4]) != 0.0;
  const int32_t gpu_id = static_cast<int32_t>(mxGetScalar(prhs[5]));
  std::vector<float> lf_store, H_store, Ht_store;
  const float* lf = as_single(prhs[0], &lf_store);
  const float* H = as_single(prhs[1], &H_store);
  const float* Ht = as_single(prhs[2], &Ht_store);
  const mwSize od[3] = {ld[0], ld[1], (mwSize)Z};
  plhs[0] = mxCreateNumericArray(3, od, mxDOUBLE_CLASS, mxREAL);
  const mlf_status st = mlf_recon_phase_space(
      lf, (int64_t)ld[0], (int64_t)ld[1], H, Ht, (int64_t)hd[0],
      (int64_t)hd[2], Z, max_iter, gpu_flag, gpu_id - 1, MLF_PRECISION_FP32,
      static_cast<double*>(mxGetData(plhs[0])), nullptr);
  if (st == MLF_ERR_INVALID_ARGUMENT) mexErrMsgIdAndTxt(kErrId, "%s", mlf_last_error());
  if (st != MLF_OK) mexErrMsgIdAndTxt("mesolf:recon:Failed", "%s", mlf_last_error());
}
