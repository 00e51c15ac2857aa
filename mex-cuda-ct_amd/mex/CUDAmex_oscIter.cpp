// CUDAmex_oscIter(reconVol, preScan, postScan, numSubsets, subsetPicks, scanAngles, SAD, geomFlag)
// Gateway of iterative_reconstruction/CUDAmex_oscIter.cu:362-584 over
// mexct_osc_iter.  Like the reference it writes reconVol, preScan and
// postScan IN PLACE through mxGetData and returns nothing.
#include "gateway.h"

extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  (void)nlhs;
  (void)plhs;
  const char* id = "mex_conebeam_FP:invalidInput";
  if (nrhs < 8) mexErrMsgIdAndTxt(id, "8 arguments expected");
  float* mu = (float*)mexgw::single_data(prhs[0], id, "reconVol");
  float* pre = (float*)mexgw::single_data(prhs[1], id, "preScan");
  float* post = (float*)mexgw::single_data(prhs[2], id, "postScan");
  const float* angles = mexgw::single_data(prhs[5], id, "scanAngles");
  if (mxGetClassID(prhs[4]) != mxINT32_CLASS) mexErrMsgIdAndTxt(id, "subsetPicks must be int32");
  const int32_t* picks = (const int32_t*)mxGetData(prhs[4]);
  const float D = mexgw::scalar_f(prhs[6]);
  const int geomFlag = mexgw::scalar_i(prhs[7]);
  const int numSubsets = mexgw::scalar_i(prhs[3]);
  const mwSize W = mexgw::dim(prhs[0], 0), H = mexgw::dim(prhs[0], 2);
  const mwSize n = mexgw::dim(prhs[1], 2);
  if (mexgw::dim(prhs[1], 0) != W || mexgw::dim(prhs[1], 1) != H ||
      mxGetNumberOfElements(prhs[2]) != mxGetNumberOfElements(prhs[1]))
    mexErrMsgIdAndTxt(id, "preScan/postScan must be MxNzxNproj for an MxMxNz volume");
  mexgw::report(mexct_osc_iter(mu, (int64_t)W, (int64_t)H, pre, post, (int64_t)n, numSubsets, picks,
                               (int64_t)mxGetNumberOfElements(prhs[4]), angles,
                               (int64_t)mxGetNumberOfElements(prhs[5]), D, geomFlag, nullptr),
                id);
}
