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

// Register-blocked form for W <= 64 R: a wave filters one row, each lane R
// consecutive outputs.  For a block of R input samples the lane needs the
// 2R-1 taps g[b-R+1 .. b+R-1] (b = u0 - U0 + W - 1), which it loads once
// from LDS and reuses for R^2 FMAs.  Every output still sums its products in
// input order from u' = 0, the same FP64 sum as prefilter_kernel.  Rows are
// zero-padded to a multiple of R (the extra products add +0) and the taps
// by zeros on both sides, so no index leaves the LDS arrays.
template <int R>
__global__ __launch_bounds__(FDK_THREADS) void prefilter_reg_kernel(
    const float* __restrict__ sino, float* __restrict__ out, const float* __restrict__ wgt,
    const double* __restrict__ g, const float* __restrict__ dA, int W, int H, int nrows) {
  extern __shared__ double lds[];
  const int Wp = (W + R - 1) / R * R;
  const int GL = R + 64 * R + W;  // padded tap array: gs[i] = g[i - R]
  double* gs = lds;
  double* xs = lds + GL + (threadIdx.x >> 6) * Wp;
  for (int i = threadIdx.x; i < GL; i += FDK_THREADS) {
    const int k = i - R;
    gs[i] = (k >= 0 && k < 2 * W - 1) ? g[k] : 0.0;
  }
  const int lane = threadIdx.x & 63;
  const int u0 = lane * R;
  for (int row0 = blockIdx.x * 4; row0 < nrows; row0 += gridDim.x * 4) {
    const int row = row0 + (int)(threadIdx.x >> 6);
    __syncthreads();
    if (row < nrows) {
      const int view = row / H, v = row - view * H;
      const size_t base = (size_t)view * W * H + (size_t)v * W;
      for (int u = lane; u < Wp; u += 64) {
        float xv = 0.f;
        if (u < W) {
          xv = sino[base + u];
          if (wgt) xv = xv * wgt[(size_t)v * W + u];
        }
        xs[u] = (double)xv;
      }
    }
    __syncthreads();
    if (row >= nrows) continue;
    double acc[R];
#pragma unroll
    for (int k = 0; k < R; ++k) acc[k] = 0.0;
    for (int U0 = 0; U0 < Wp; U0 += R) {
      const double* t0 = gs + R + (u0 - U0 + W - 1) - (R - 1);  // t[i] = g[b - R + 1 + i]
      double t[2 * R - 1];
#pragma unroll
      for (int i = 0; i < 2 * R - 1; ++i) t[i] = t0[i];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const double x = xs[U0 + j];
#pragma unroll
        for (int k = 0; k < R; ++k) acc[k] = __builtin_fma(x, t[k - j + R - 1], acc[k]);
      }
    }
    const int view = row / H, v = row - view * H;
    const size_t base = (size_t)view * W * H + (size_t)v * W;
    const float d = dA[view];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (u0 + k < W) out[base + u0 + k] = ((float)acc[k] * d) / 2.0f;
    }
  }
}

template <int R>
static hipError_t launch_reg(const float* sino, float* out, const float* wgt, const double* g,
                             const float* dA, int W, int H, int nrows, hipStream_t st) {
  const int Wp = (W + R - 1) / R * R;
  const size_t lds = sizeof(double) * (size_t)(R + 64 * R + W + 4 * Wp);
  const int blocks = std::min((nrows + 3) / 4, 16384);
  prefilter_reg_kernel<R><<<blocks, FDK_THREADS, lds, st>>>(sino, out, wgt, g, dA, W, H, nrows);
  return hipGetLastError();
}

hipError_t launch_prefilter(const float* sino, float* out, const float* wgt, const double* g,
                            const float* dA, int W, int H, int nviews, hipStream_t st) {
  const int nrows = H * nviews;
  if (nrows == 0) return hipSuccess;
  if (W <= 64 * 2) return launch_reg<2>(sino, out, wgt, g, dA, W, H, nrows, st);
  if (W <= 64 * 4) return launch_reg<4>(sino, out, wgt, g, dA, W, H, nrows, st);
  if (W <= 64 * 6) return launch_reg<6>(sino, out, wgt, g, dA, W, H, nrows, st);
  if (W <= 64 * 8) return launch_reg<8>(sino, out, wgt, g, dA, W, H, nrows, st);
  if (W <= 64 * 12) return launch_reg<12>(sino, out, wgt, g, dA, W, H, nrows, st);
  const size_t lds = sizeof(double) * (size_t)(3 * W - 1);
  int blocks = nrows < 8192 ? nrows : 8192;
  prefilter_kernel<<<blocks, FDK_THREADS, lds, st>>>(sino, out, wgt, g, dA, W, H, nrows);
  return hipGetLastError();
}

}  // namespace mexct
