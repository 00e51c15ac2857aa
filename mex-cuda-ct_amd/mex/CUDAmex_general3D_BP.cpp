// bp = CUDAmex_general3D_BP(projections, angles, pts, M, N)
// Gateway of projection_kernels/CUDAmex_general3D_BP.cu:194-282 over
// mexct_g3d_bp.
#include "gateway.h"

extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  (void)nlhs;
  const char* id = "mex_conebeam_FP:invalidInput";
  if (nrhs < 5) mexErrMsgIdAndTxt(id, "5 arguments expected");
  const float* P = mexgw::single_data(prhs[0], id, "projections");
  const float* angles = mexgw::single_data(prhs[1], id, "angles");
  const float* pts = mexgw::single_data(prhs[2], id, "pts");
  const unsigned M = (unsigned)mxGetScalar(prhs[3]);
  const unsigned N = (unsigned)mxGetScalar(prhs[4]);
  const mwSize W = mexgw::dim(prhs[0], 0), H = mexgw::dim(prhs[0], 1), n = mexgw::dim(prhs[0], 2);
  if (mexgw::dim(prhs[2], 0) != W || mexgw::dim(prhs[2], 1) != H || mexgw::dim(prhs[2], 2) != 6)
    mexErrMsgIdAndTxt(id, "pts must be MxNx6 with the projections' M and N");
  const int ppv = mxGetNumberOfDimensions(prhs[2]) >= 4 ? 1 : 0;
  if (ppv && mexgw::dim(prhs[2], 3) < n)
    mexErrMsgIdAndTxt(id, "per-view pts has fewer views than the projections");
  const mwSize dims[] = {M, M, N};
  plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
  mexgw::report(mexct_g3d_bp(P, (int64_t)W, (int64_t)H, (int64_t)n, angles,
                             (int64_t)mxGetNumberOfElements(prhs[1]), pts, ppv, (int64_t)M,
                             (int64_t)N, (float*)mxGetData(plhs[0]), nullptr),
                id);
}
