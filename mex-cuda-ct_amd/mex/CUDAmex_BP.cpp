// bp = CUDAmex_BP(projections, SAD, angles, geomFlag)
// Gateway of projection_kernels/CUDAmex_BP.cu:223-324 over mexct_bp.
#include "gateway.h"

extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  (void)nlhs;
  const char* id = "mex_conebeam_fdk:invalidInput";
  if (nrhs < 4) mexErrMsgIdAndTxt(id, "4 arguments expected");
  const float* P = mexgw::single_data(prhs[0], id, "projections");
  const float* angles = mexgw::single_data(prhs[2], id, "angles");
  const mwSize W = mexgw::dim(prhs[0], 0), H = mexgw::dim(prhs[0], 1), n = mexgw::dim(prhs[0], 2);
  const float D = mexgw::scalar_f(prhs[1]);
  const int geomFlag = mexgw::scalar_i(prhs[3]);
  const mwSize dims[] = {W, W, H};
  plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
  mexgw::report(mexct_bp(P, (int64_t)W, (int64_t)H, (int64_t)n, D, angles,
                         (int64_t)mxGetNumberOfElements(prhs[2]), geomFlag,
                         (float*)mxGetData(plhs[0]), nullptr),
                id);
}
