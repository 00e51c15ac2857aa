// fdk.hip -- FDK/FBP sinogram pre-filter (preFilterSinogram.m:29-152
// without Parker weighting) for gfx950.
//
// The reference zero-pads every detector row to `order` >= 2W samples,
// multiplies its FFT by the designed filter and keeps the first W outputs
// of the inverse FFT (:72-78).  With order >= 2W that is exactly the
// linear convolution y[u] = sum_u' x[u'] h[(u - u') mod order] with the
// filter's circular impulse response h, so each row is filtered here by a
// direct FP64 convolution against the 2W-1 taps g[d] = h[d mod order],
// d in [-(W-1), W-1], staged once per workgroup in LDS.  The cosine weight
// (:51-58, single * single) is applied while the row is staged, and the
// angle-step scale (:80-84) in single precision on the way out.
#include "common.h"
#include "internal.h"

namespace mexct {

constexpr int FDK_THREADS = 256;

__global__ __launch_bounds__(FDK_THREADS) void prefilter_kernel(
    const float* __restrict__ sino, float* __restrict__ out, const float* __restrict__ wgt,
    const double* __restrict__ g, const float* __restrict__ dA, int W, int H, int nrows) {
  extern __shared__ double lds[];
  double* gs = lds;              // 2W-1 taps
  double* xs = lds + (2 * W - 1);  // one row
  for (int i = threadIdx.x; i < 2 * W - 1; i += FDK_THREADS) gs[i] = g[i];
  for (int row = blockIdx.x; row < nrows; row += gridDim.x) {
    const int view = row / H, v = row - view * H;
    const size_t base = (size_t)view * W * H + (size_t)v * W;
    __syncthreads();
    for (int u = threadIdx.x; u < W; u += FDK_THREADS) {
      float xv = sino[base + u];
      if (wgt) xv = xv * wgt[(size_t)v * W + u];
      xs[u] = (double)xv;
    }
    __syncthreads();
    const float d = dA[view];
    for (int u = threadIdx.x; u < W; u += FDK_THREADS) {
      const double* gu = gs + (u + W - 1);  // gu[-u'] = g[u - u']
      double acc = 0.0;
      for (int up = 0; up < W; ++up) acc = __builtin_fma(xs[up], gu[-up], acc);
      const float y = (float)acc;
      out[base + u] = (y * d) / 2.0f;
    }
  }
}

hipError_t launch_prefilter(const float* sino, float* out, const float* wgt, const double* g,
                            const float* dA, int W, int H, int nviews, hipStream_t st) {
  const int nrows = H * nviews;
  if (nrows == 0) return hipSuccess;
  const size_t lds = sizeof(double) * (size_t)(3 * W - 1);
  int blocks = nrows < 8192 ? nrows : 8192;
  prefilter_kernel<<<blocks, FDK_THREADS, lds, st>>>(sino, out, wgt, g, dA, W, H, nrows);
  return hipGetLastError();
}

}  // namespace mexct
