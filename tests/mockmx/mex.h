// Minimal stand-in for MATLAB's MEX/matrix API (the subset the mesolf
// gateways use), so the gateways can be compiled and exercised without
// MATLAB.  Arrays are column-major like MATLAB's; errors raised with
// mexErrMsgIdAndTxt become a C++ exception carrying the identifier.
#pragma once
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

typedef size_t mwSize;
typedef enum { mxUNKNOWN_CLASS = 0, mxLOGICAL_CLASS = 3, mxDOUBLE_CLASS = 6, mxSINGLE_CLASS = 7 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

struct mxArray {
  mxClassID cls = mxDOUBLE_CLASS;
  bool complex = false;
  std::vector<mwSize> dims;
  std::vector<unsigned char> data;
};

struct MexError : std::runtime_error {
  std::string id;
  MexError(const std::string& i, const std::string& m) : std::runtime_error(m), id(i) {}
};

inline size_t mx_elem_size(mxClassID c) { return c == mxSINGLE_CLASS ? 4 : c == mxDOUBLE_CLASS ? 8 : 1; }

inline mxArray* mxCreateNumericArray(mwSize nd, const mwSize* dims, mxClassID cls, mxComplexity cx) {
  mxArray* a = new mxArray;
  a->cls = cls;
  a->complex = cx == mxCOMPLEX;
  a->dims.assign(dims, dims + nd);
  while (a->dims.size() > 2 && a->dims.back() == 1) a->dims.pop_back();  // MATLAB drops trailing 1s
  if (a->dims.size() < 2) a->dims.resize(2, 1);
  size_t n = 1;
  for (mwSize d : a->dims) n *= d;
  a->data.assign(n * mx_elem_size(cls), 0);
  return a;
}
inline void mxDestroyArray(mxArray* a) { delete a; }
inline mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->dims.size(); }
inline const mwSize* mxGetDimensions(const mxArray* a) { return a->dims.data(); }
inline size_t mxGetNumberOfElements(const mxArray* a) {
  size_t n = 1;
  for (mwSize d : a->dims) n *= d;
  return n;
}
inline bool mxIsSingle(const mxArray* a) { return a->cls == mxSINGLE_CLASS; }
inline bool mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS; }
inline bool mxIsComplex(const mxArray* a) { return a->complex; }
inline mxClassID mxGetClassID(const mxArray* a) { return a->cls; }
inline void* mxGetData(const mxArray* a) { return const_cast<unsigned char*>(a->data.data()); }
inline double mxGetScalar(const mxArray* a) {
  if (a->cls == mxSINGLE_CLASS) return *reinterpret_cast<const float*>(a->data.data());
  if (a->cls == mxDOUBLE_CLASS) return *reinterpret_cast<const double*>(a->data.data());
  return a->data.empty() ? 0.0 : a->data[0];
}
inline void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw MexError(id, buf);
}
inline int mexPrintf(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  int r = vfprintf(stderr, fmt, ap);
  va_end(ap);
  return r;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
