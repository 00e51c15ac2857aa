/*
 * mex.h -- minimal stand-in for MATLAB's MEX/matrix C API, enough to build
 * and test the mexct gateways without MATLAB.  Names, argument order and
 * mxClassID numbering follow MATLAB's documented C Matrix API; build the
 * gateways against MATLAB's own mex.h for real use (the gateways only use
 * the functions declared here).
 *
 * Errors: mexErrMsgIdAndTxt / mexErrMsgTxt do not return (in MATLAB they
 * long-jump to the prompt; here they throw mexct_shim_error, which the
 * test harness entry point shim_invoke() catches).
 */
#ifndef MEXCT_SHIM_MEX_H
#define MEXCT_SHIM_MEX_H

#include <stddef.h>
#include <stdint.h>

typedef size_t mwSize;
typedef size_t mwIndex;

typedef enum {
  mxUNKNOWN_CLASS = 0,
  mxCELL_CLASS,
  mxSTRUCT_CLASS,
  mxLOGICAL_CLASS,
  mxCHAR_CLASS,
  mxVOID_CLASS,
  mxDOUBLE_CLASS,
  mxSINGLE_CLASS,
  mxINT8_CLASS,
  mxUINT8_CLASS,
  mxINT16_CLASS,
  mxUINT16_CLASS,
  mxINT32_CLASS,
  mxUINT32_CLASS,
  mxINT64_CLASS,
  mxUINT64_CLASS
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX } mxComplexity;

typedef struct mxArray_tag mxArray;

#ifdef __cplusplus
extern "C" {
#endif

void* mxGetData(const mxArray* pa);
const mwSize* mxGetDimensions(const mxArray* pa);
mwSize mxGetNumberOfDimensions(const mxArray* pa);
size_t mxGetNumberOfElements(const mxArray* pa);
double mxGetScalar(const mxArray* pa);
mxClassID mxGetClassID(const mxArray* pa);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID classid,
                              mxComplexity flag);
void mexErrMsgIdAndTxt(const char* errorid, const char* errormsg, ...);
void mexErrMsgTxt(const char* errormsg);

/* test-harness extensions (not part of MATLAB's API) */
mxArray* shim_new(int classid, int ndim, const size_t* dims, const void* data);
void shim_free(mxArray* pa);
typedef void (*shim_mex_fn)(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
int shim_invoke(shim_mex_fn fn, int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs);
const char* shim_last_id(void);
const char* shim_last_msg(void);

#ifdef __cplusplus
}
#endif

#endif
