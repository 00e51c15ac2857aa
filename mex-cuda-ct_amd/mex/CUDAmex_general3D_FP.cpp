// fp = CUDAmex_general3D_FP(image, angles, pts, nProjections)
// Gateway of projection_kernels/CUDAmex_general3D_FP.cu:154-252 over
// mexct_g3d_fp.  A 4-D pts [W,H,6,nProj] gives each view its own rays.
#include "gateway.h"

extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  (void)nlhs;
  const char* id = "mex_conebeam_FP:invalidInput";
  if (nrhs < 4) mexErrMsgIdAndTxt(id, "4 arguments expected");
  const float* vol = mexgw::single_data(prhs[0], id, "image");
  const float* angles = mexgw::single_data(prhs[1], id, "angles");
  const float* pts = mexgw::single_data(prhs[2], id, "pts");
  const int nProj = mexgw::scalar_i(prhs[3]);
  const mwSize W = mexgw::dim(prhs[2], 0), H = mexgw::dim(prhs[2], 1);
  if (mexgw::dim(prhs[2], 2) != 6) mexErrMsgIdAndTxt(id, "pts must be MxNx6");
  const int ppv = mxGetNumberOfDimensions(prhs[2]) >= 4 ? 1 : 0;
  if (ppv && mexgw::dim(prhs[2], 3) < (mwSize)nProj)
    mexErrMsgIdAndTxt(id, "per-view pts has fewer views than nProjections");
  if (nProj < 0) mexErrMsgIdAndTxt(id, "nProjections must be >= 0");
  const mwSize side = mexgw::dim(prhs[0], 0), nSlices = mexgw::dim(prhs[0], 2);
  const mwSize dims[] = {W, H, (mwSize)nProj};
  plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
  mexgw::report(mexct_g3d_fp(vol, (int64_t)side, (int64_t)nSlices, angles,
                             (int64_t)mxGetNumberOfElements(prhs[1]), pts, (int64_t)W, (int64_t)H,
                             nProj, ppv, (float*)mxGetData(plhs[0]), nullptr),
                id);
}
