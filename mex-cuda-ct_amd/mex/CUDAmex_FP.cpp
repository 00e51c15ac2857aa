// fp = CUDAmex_FP(vol, SAD, angles, nProjections, geomFlag)
// Gateway of projection_kernels/CUDAmex_FP.cu:350-477 over mexct_fp.
// With 4 arguments geomFlag defaults to 0 (parallel): the reference reads
// prhs[4] past the argument list there (general3D_projection_test.m:36).
#include "gateway.h"

extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  (void)nlhs;
  const char* id = "mex_conebeam_FP:invalidInput";
  if (nrhs < 4) mexErrMsgIdAndTxt(id, "4 arguments expected");
  const float* vol = mexgw::single_data(prhs[0], id, "image");
  const float* angles = mexgw::single_data(prhs[2], id, "angles");
  const mwSize W = mexgw::dim(prhs[0], 0), Wy = mexgw::dim(prhs[0], 1), H = mexgw::dim(prhs[0], 2);
  if (Wy != W) mexErrMsgIdAndTxt(id, "image must be MxMxNz");
  const float D = mexgw::scalar_f(prhs[1]);
  const int nProj = mexgw::scalar_i(prhs[3]);
  const int geomFlag = nrhs >= 5 ? mexgw::scalar_i(prhs[4]) : 0;
  if (nProj < 0) mexErrMsgIdAndTxt(id, "nProjections must be >= 0");
  const mwSize dims[] = {W, H, (mwSize)nProj};
  plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
  mexgw::report(mexct_fp(vol, (int64_t)W, (int64_t)H, D, angles,
                         (int64_t)mxGetNumberOfElements(prhs[2]), nProj, geomFlag,
                         (float*)mxGetData(plhs[0]), nullptr),
                id);
}
