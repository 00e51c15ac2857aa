// gateway.h -- shared argument handling of the mexct MEX gateways.
//
// The gateways keep the reference's calling convention: positional
// arguments, scalars read with mxGetScalar and converted by C rules
// (float for SAD and c, truncation to int for counts and geomFlag), arrays
// borrowed zero-copy with mxGetData, outputs created with
// mxCreateNumericArray(..., mxSINGLE_CLASS, mxREAL).  Where the reference
// would read garbage (a non-single array reinterpreted as float*, nProj
// larger than the angle vector) they raise instead.
#pragma once

#include <stdint.h>

#include "mex.h"
#include "mexct.h"

namespace mexgw {

inline const float* single_data(const mxArray* a, const char* id, const char* what) {
  if (mxGetClassID(a) != mxSINGLE_CLASS)
    mexErrMsgIdAndTxt(id, "%s must be single", what);
  return (const float*)mxGetData(a);
}

inline mwSize dim(const mxArray* a, int k) {
  return (mwSize)k < mxGetNumberOfDimensions(a) ? mxGetDimensions(a)[k] : 1;
}

inline float scalar_f(const mxArray* a) { return (float)mxGetScalar(a); }
inline int scalar_i(const mxArray* a) { return (int)mxGetScalar(a); }

// Map a C-ABI failure onto the reference's errors: "invalid geometry" via
// mexErrMsgTxt (CUDAmex_FP.cu:463), argument errors under `id`, device
// errors with the runtime's message (CHECK_CUDA_ERROR, CUDAmex_FP.cu:43-50).
inline void report(int rc, const char* id) {
  if (rc == MEXCT_OK) return;
  if (rc == MEXCT_EGEOM) mexErrMsgTxt("invalid geometry");
  if (rc == MEXCT_EINVAL) mexErrMsgIdAndTxt(id, "%s", mexct_last_error());
  mexErrMsgTxt(mexct_last_error());
}

}  // namespace mexgw
