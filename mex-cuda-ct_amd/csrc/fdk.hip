// fdk.hip -- FDK/FBP sinogram pre-filter (preFilterSinogram.m:29-152), its
// Parker window (parker_weight.m) and FBP_recon.m's log transform, for gfx950.
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
//
// Staging also applies, in the reference's order, FBP_recon's
// -log(post./pre) with NaN/Inf -> 0 (FBP_recon.m:26-28) when a post array
// is given, and the Parker window after the cosine weight (:60).
#include "common.h"
#include "internal.h"

namespace mexct {

constexpr int FDK_THREADS = 256;

// Sample u of detector row v of view `view`, weighted as the filter expects.
__device__ __forceinline__ float stage_sample(const FilterIn& in, size_t base, int view, int v,
                                              int u, int W) {
  float x;
  if (in.post) {
    x = -logf(in.post[base + u] / in.sino[base + u]);
    if (!__builtin_isfinite(x)) x = 0.f;
  } else {
    x = in.sino[base + u];
  }
  if (in.wgt) x = x * in.wgt[(size_t)v * W + u];
  if (in.pw) x = x * in.pw[(size_t)view * W + u];
  return x;
}

__global__ __launch_bounds__(FDK_THREADS) void prefilter_kernel(
    FilterIn in, float* __restrict__ out, const double* __restrict__ g,
    const float* __restrict__ dA, int W, int H, int nrows) {
  extern __shared__ double lds[];
  double* gs = lds;              // 2W-1 taps
  double* xs = lds + (2 * W - 1);  // one row
  for (int i = threadIdx.x; i < 2 * W - 1; i += FDK_THREADS) gs[i] = g[i];
  for (int row = blockIdx.x; row < nrows; row += gridDim.x) {
    const int view = row / H, v = row - view * H;
    const size_t base = (size_t)view * W * H + (size_t)v * W;
    __syncthreads();
    for (int u = threadIdx.x; u < W; u += FDK_THREADS)
      xs[u] = (double)stage_sample(in, base, view, v, u, W);
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
    FilterIn in, float* __restrict__ out, const double* __restrict__ g,
    const float* __restrict__ dA, int W, int H, int nrows) {
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
      for (int u = lane; u < Wp; u += 64)
        xs[u] = u < W ? (double)stage_sample(in, base, view, v, u, W) : 0.0;
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
static hipError_t launch_reg(const FilterIn& in, float* out, const double* g, const float* dA,
                             int W, int H, int nrows, hipStream_t st) {
  const int Wp = (W + R - 1) / R * R;
  const size_t lds = sizeof(double) * (size_t)(R + 64 * R + W + 4 * Wp);
  const int blocks = std::min((nrows + 3) / 4, 16384);
  prefilter_reg_kernel<R><<<blocks, FDK_THREADS, lds, st>>>(in, out, g, dA, W, H, nrows);
  return hipGetLastError();
}

hipError_t launch_prefilter(const FilterIn& in, float* out, const double* g, const float* dA,
                            int W, int H, int nviews, hipStream_t st) {
  const int nrows = H * nviews;
  if (nrows == 0) return hipSuccess;
  if (W <= 64 * 2) return launch_reg<2>(in, out, g, dA, W, H, nrows, st);
  if (W <= 64 * 4) return launch_reg<4>(in, out, g, dA, W, H, nrows, st);
  if (W <= 64 * 6) return launch_reg<6>(in, out, g, dA, W, H, nrows, st);
  if (W <= 64 * 8) return launch_reg<8>(in, out, g, dA, W, H, nrows, st);
  if (W <= 64 * 12) return launch_reg<12>(in, out, g, dA, W, H, nrows, st);
  const size_t lds = sizeof(double) * (size_t)(3 * W - 1);
  int blocks = nrows < 8192 ? nrows : 8192;
  prefilter_kernel<<<blocks, FDK_THREADS, lds, st>>>(in, out, g, dA, W, H, nrows);
  return hipGetLastError();
}

// ---- Parker / Wesarg short-scan window (parker_weight.m:63-97)
//
// S_function (:82-89): 0 at or below -1/2, 1 at or above 1/2,
// 0.5*(1 + sin(pi*b)) between, and 0 for NaN (a 0/0 at a degenerate edge
// ray), all in single precision as MATLAB evaluates it on single data.
__device__ __forceinline__ float parker_s(float b, float pi_f) {
  if (b <= -0.5f) return 0.f;
  if (fabsf(b) < 0.5f) return 0.5f * (1.0f + sinf(pi_f * b));
  if (b >= 0.5f) return 1.f;
  return 0.f;
}

// win[i*W + u], the window of view i at detector column u.  The per-column
// terms col[u] = {2 alpha_u, b(alpha_u), b(-alpha_u)} and the per-view
// beta_i = |angle_i - angle_0| arrive already rounded to single; the four S
// terms and their sum follow :68-71 operation by operation.
__global__ __launch_bounds__(256) void parker_kernel(const float4* __restrict__ col,
                                                     const float* __restrict__ beta,
                                                     ParkerConsts k, int W, int nviews,
                                                     float* __restrict__ win) {
  const int u = (int)(blockIdx.x * 256 + threadIdx.x);
  const int i = (int)blockIdx.y;
  if (u >= W || i >= nviews) return;
  const float4 c = col[u];
  const float b = beta[i];
  const float s1 = parker_s(b / c.y - 0.5f, k.pi_f);
  const float s2 = parker_s((((b - k.d2) + c.x) - k.eps) / c.y + 0.5f, k.pi_f);
  const float s3 = parker_s(((b - k.pi_f) + c.x) / c.z - 0.5f, k.pi_f);
  const float s4 = parker_s((((b - k.pi_f) - k.d2) - k.eps) / c.z + 0.5f, k.pi_f);
  win[(size_t)i * W + u] = 0.5f * (((s1 + s2) - s3) - s4);
}

hipError_t launch_parker(const float4* col, const float* beta, const ParkerConsts& k, int W,
                         int nviews, float* win, hipStream_t st) {
  if (W <= 0 || nviews <= 0) return hipSuccess;
  const dim3 g((unsigned)((W + 255) / 256), (unsigned)nviews);
  parker_kernel<<<g, 256, 0, st>>>(col, beta, k, W, nviews, win);
  return hipGetLastError();
}

// out = sino .* window, broadcast over the detector rows (parker_weight.m:72-73)
__global__ __launch_bounds__(256) void apply_window_kernel(const float* __restrict__ sino,
                                                           const float* __restrict__ win,
                                                           float* __restrict__ out, int W, int H,
                                                           int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += stride) {
    const int64_t u = e % W, view = e / ((int64_t)W * H);
    out[e] = sino[e] * win[view * W + u];
  }
}

hipError_t launch_apply_window(const float* sino, const float* win, float* out, int W, int H,
                               int nviews, hipStream_t st) {
  const int64_t n = (int64_t)W * H * nviews;
  if (n == 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 16384);
  apply_window_kernel<<<(unsigned)blocks, 256, 0, st>>>(sino, win, out, W, H, n);
  return hipGetLastError();
}

}  // namespace mexct
