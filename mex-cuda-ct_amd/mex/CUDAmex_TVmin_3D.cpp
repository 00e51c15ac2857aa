// image_out = CUDAmex_TVmin_3D(image_in, tvConst)
// Gateway of iterative_reconstruction/CUDAmex_TVmin_3D.cu:162-246 over
// mexct_tvmin (20 steps, as the reference).
#include "gateway.h"

extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  (void)nlhs;
  const char* id = "mex_conebeam_FP:invalidInput";
  if (nrhs != 2) mexErrMsgIdAndTxt(id, "2 arguments expected");
  const float* in = mexgw::single_data(prhs[0], id, "image_in");
  const float c = mexgw::scalar_f(prhs[1]);
  const mwSize dims[] = {mexgw::dim(prhs[0], 0), mexgw::dim(prhs[0], 1), mexgw::dim(prhs[0], 2)};
  plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
  mexgw::report(mexct_tvmin(in, (int64_t)dims[0], (int64_t)dims[1], (int64_t)dims[2], c,
                            (float*)mxGetData(plhs[0]), nullptr),
                id);
}
