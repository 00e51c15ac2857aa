// tv.hip -- 3D total-variation gradient descent (CUDAmex_TVmin_3D.cu) and
// the device-side pieces of the OSC-TV driver (OSC_TV_recon.m).
//
// Per step (CUDAmex_TVmin_3D.cu:221-234):
//   1. tv_grad_kernel: 13-point TV gradient g on interior voxels (the
//      border of g is 0), plus per-workgroup FP64 partial sums of g^2
//      reduced with wave shuffles (a workgroup: 64 x 4 x TVZC voxels);
//   2. tv_scale_kernel: one workgroup sums the partials in a fixed order,
//      |g| = (float)sqrt(sum), a = (-c)/|g| (the reference sums in float
//      with Thrust and takes the sqrt on the host);
//   3. tv_axpy_kernel: x <- a*g + x over every voxel, border included.
// The launches stay on the stream: no host round trip per step.
#include "common.h"
#include "internal.h"

namespace mexct {

constexpr int TVBX = 64, TVBY = 4;
constexpr int TVZC = 4;  // z-slices per workgroup of the reducing kernels (one partial each)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block reduction of one double per thread (blockDim.x*y*z == 256); the
// result is valid in thread 0.
__device__ __forceinline__ double block_sum256(double v) {
  __shared__ double part[4];
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  v = wave_sum(v);
  if ((tid & 63) == 0) part[tid >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (tid == 0) s = ((part[0] + part[1]) + part[2]) + part[3];
  __syncthreads();
  return s;
}

// tvGradientKernel (CUDAmex_TVmin_3D.cu:101-158)
__global__ __launch_bounds__(256) void tv_grad_kernel(Vol3 x, float* __restrict__ g,
                                                      double* __restrict__ partials) {
  const int ix = blockIdx.x * TVBX + threadIdx.x;
  const int iy = blockIdx.y * TVBY + threadIdx.y;
  const int Nx = x.nx, Ny = x.ny, Nz = x.nz;
  double sq = 0.0;
  for (int k = 0; k < TVZC; ++k) {
    const int iz = blockIdx.z * TVZC + k;
    if (ix >= Nx || iy >= Ny || iz >= Nz) break;
    float tval = 0.f;
    if ((ix < (Nx - 1) && iy < (Ny - 1) && iz < (Nz - 1)) && (ix >= 1 && iy >= 1 && iz >= 1)) {
      const float eps = (float)1.0e-6;
      const float* in = x.p + (size_t)ix + (size_t)x.sy * iy + (size_t)x.sz * iz;
      const int64_t sy = x.sy, sz = x.sz;
      const float xyz = in[0];
      const float xmyz = in[-1];
      const float xpyz = in[1];
      const float xymz = in[-sy];
      const float xypz = in[sy];
      const float xyzm = in[-sz];
      const float xyzp = in[sz];
      const float xmypz = in[-1 + sy];
      const float xpymz = in[1 - sy];
      const float xymzp = in[-sy + sz];
      const float xypzm = in[sy - sz];
      const float xmyzp = in[-1 + sz];
      const float xpyzm = in[1 - sz];
      tval = ((xyz - xmyz) + (xyz - xymz) + (xyz - xyzm)) /
             (eps + sqrtf((xyz - xmyz) * (xyz - xmyz) + (xyz - xymz) * (xyz - xymz) +
                          (xyz - xyzm) * (xyz - xyzm)));
      tval -= ((xpyz - xyz)) / (eps + sqrtf((xpyz - xyz) * (xpyz - xyz) +
                                            (xpyz - xpymz) * (xpyz - xpymz) +
                                            (xpyz - xpyzm) * (xpyz - xpyzm)));
      tval -= ((xypz - xyz)) / (eps + sqrtf((xypz - xyz) * (xypz - xyz) +
                                            (xypz - xmypz) * (xypz - xmypz) +
                                            (xypz - xypzm) * (xypz - xypzm)));
      tval -= ((xyzp - xyz)) / (eps + sqrtf((xyzp - xyz) * (xyzp - xyz) +
                                            (xyzp - xmyzp) * (xyzp - xmyzp) +
                                            (xyzp - xymzp) * (xyzp - xymzp)));
    }
    g[(size_t)ix + (size_t)Nx * ((size_t)iy + (size_t)Ny * iz)] = tval;
    sq += (double)tval * (double)tval;
  }
  const double s = block_sum256(sq);
  if (threadIdx.x == 0 && threadIdx.y == 0)
    partials[blockIdx.x + gridDim.x * ((size_t)blockIdx.y + (size_t)gridDim.y * blockIdx.z)] = s;
}

// Fixed-order sum of n partials by one 256-thread workgroup.
__device__ __forceinline__ double sum_partials(const double* __restrict__ partials, size_t n) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;  // four loads in flight per thread
  size_t i = threadIdx.x;
  for (; i + 768 < n; i += 1024) {
    s0 += partials[i];
    s1 += partials[i + 256];
    s2 += partials[i + 512];
    s3 += partials[i + 768];
  }
  for (; i < n; i += 256) s0 += partials[i];
  return block_sum256((s0 + s1) + (s2 + s3));
}

__global__ __launch_bounds__(256) void tv_scale_kernel(const double* __restrict__ partials,
                                                       size_t n, const float* c_dev, float c,
                                                       float* a_out) {
  const double s = sum_partials(partials, n);
  if (threadIdx.x == 0) {
    const float cc = c_dev ? *c_dev : c;
    const float magnitude = (float)sqrt(s);
    *a_out = -cc / magnitude;
  }
}

// saxpy_functor (CUDAmex_TVmin_3D.cu:80-90): x <- a*g + x
__global__ __launch_bounds__(256) void tv_axpy_kernel(Vol3 x, const float* __restrict__ g,
                                                      const float* __restrict__ a_dev) {
  const int ix = blockIdx.x * TVBX + threadIdx.x;
  const int iy = blockIdx.y * TVBY + threadIdx.y;
  const int iz = blockIdx.z;
  if (ix >= x.nx || iy >= x.ny) return;
  const float a = *a_dev;
  float* p = x.p + (size_t)ix + (size_t)x.sy * iy + (size_t)x.sz * iz;
  const float gv = g[(size_t)ix + (size_t)x.nx * ((size_t)iy + (size_t)x.ny * iz)];
  *p = a * gv + *p;
}

static dim3 tv_grid(int nx, int ny, int nz) {
  return dim3((unsigned)((nx + TVBX - 1) / TVBX), (unsigned)((ny + TVBY - 1) / TVBY), (unsigned)nz);
}

// grid of the reducing kernels: TVZC slices per workgroup
static dim3 red_grid(int nx, int ny, int nz) {
  return dim3((unsigned)((nx + TVBX - 1) / TVBX), (unsigned)((ny + TVBY - 1) / TVBY),
              (unsigned)((nz + TVZC - 1) / TVZC));
}

size_t tv_partials_count(int nx, int ny, int nz) {
  dim3 g = red_grid(nx, ny, nz);
  return (size_t)g.x * g.y * g.z;
}

hipError_t launch_tvmin(Vol3 x, float* g, double* partials, float* a_dev, const float* c_dev,
                        float c, int steps, hipStream_t st) {
  const dim3 grid = tv_grid(x.nx, x.ny, x.nz);
  const dim3 block(TVBX, TVBY, 1);
  const size_t np = tv_partials_count(x.nx, x.ny, x.nz);
  for (int i = 0; i < steps; ++i) {
    tv_grad_kernel<<<red_grid(x.nx, x.ny, x.nz), block, 0, st>>>(x, g, partials);
    tv_scale_kernel<<<1, 256, 0, st>>>(partials, np, c_dev, c, a_dev);
    tv_axpy_kernel<<<grid, block, 0, st>>>(x, g, a_dev);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------- OSC-TV
__global__ __launch_bounds__(256) void diff_norm_kernel(Vol3 o, Vol3 n,
                                                        double* __restrict__ partials) {
  const int ix = blockIdx.x * TVBX + threadIdx.x;
  const int iy = blockIdx.y * TVBY + threadIdx.y;
  double sq = 0.0;
  for (int k = 0; k < TVZC; ++k) {
    const int iz = blockIdx.z * TVZC + k;
    if (ix >= o.nx || iy >= o.ny || iz >= o.nz) break;
    float ov = o.p[(size_t)ix + (size_t)o.sy * iy + (size_t)o.sz * iz];
    if (ix == 0 && iy == 0 && iz == 0) ov = 0.0f;  // mu_old(1) = 0
    const float nv = n.p[(size_t)ix + (size_t)n.sy * iy + (size_t)n.sz * iz];
    float d = ov - nv;
    if (d != d) d = 0.0f;
    sq += (double)d * (double)d;
  }
  const double s = block_sum256(sq);
  if (threadIdx.x == 0 && threadIdx.y == 0)
    partials[blockIdx.x + gridDim.x * ((size_t)blockIdx.y + (size_t)gridDim.y * blockIdx.z)] = s;
}

__global__ __launch_bounds__(256) void final_sum_kernel(const double* __restrict__ partials,
                                                        size_t n, double* out) {
  const double s = sum_partials(partials, n);
  if (threadIdx.x == 0) *out = s;
}

hipError_t launch_diff_norm(Vol3 old_v, Vol3 new_v, double* partials, double* sum,
                            hipStream_t st) {
  const dim3 grid = red_grid(old_v.nx, old_v.ny, old_v.nz);
  diff_norm_kernel<<<grid, dim3(TVBX, TVBY, 1), 0, st>>>(old_v, new_v, partials);
  final_sum_kernel<<<1, 256, 0, st>>>(partials, tv_partials_count(old_v.nx, old_v.ny, old_v.nz),
                                      sum);
  return hipGetLastError();
}

__global__ void tv_const_kernel(const double* sum, float tv, float* c_dev) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const float d_a = (float)sqrt(*sum);
    *c_dev = tv * d_a;
  }
}

hipError_t launch_tv_const(const double* sum, float tv, float* c_dev, hipStream_t st) {
  tv_const_kernel<<<1, 64, 0, st>>>(sum, tv, c_dev);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void clean_kernel(Vol3 x) {
  const int ix = blockIdx.x * TVBX + threadIdx.x;
  const int iy = blockIdx.y * TVBY + threadIdx.y;
  const int iz = blockIdx.z;
  if (ix >= x.nx || iy >= x.ny) return;
  float* p = x.p + (size_t)ix + (size_t)x.sy * iy + (size_t)x.sz * iz;
  const float v = *p;
  if (v != v || fabsf(v) == __builtin_inff()) *p = 0.0f;
}

hipError_t launch_clean(Vol3 x, hipStream_t st) {
  clean_kernel<<<tv_grid(x.nx, x.ny, x.nz), dim3(TVBX, TVBY, 1), 0, st>>>(x);
  return hipGetLastError();
}

}  // namespace mexct
