// probe.hip -- gather-throughput microbenchmark (measurement only).
//
// Issues the forward projector's load pattern -- per sample four (k, k+1)
// pair loads at the (i,j), (i,j+1), (i+1,j), (i+1,j+1) corners of a
// z-fastest haloed volume, 32-bit buffer offsets, corner strides in scalar
// offsets -- with ideal coherence: the 64 lanes of a wave hold 64
// consecutive z, each wave one (j) column, and every lane steps along x.
// The only VALU work is the offset step and a sum that keeps the loads
// alive, so the rate it reaches is the memory pipeline's ceiling for the
// projector's gathers (the "measured gather roofline").  The backprojector
// pattern uses two pair loads per update ((j0, j0+1) rows of one column).
#include "common.h"
#include "internal.h"

namespace mexct {

template <typename S>
__device__ __forceinline__ float2 pld2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(S) == 4) {
    typedef float f2v __attribute__((ext_vector_type(2)));
    f2v v = __builtin_bit_cast(f2v, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    return make_float2(v.x, v.y);
  } else {
    typedef _Float16 h2v __attribute__((ext_vector_type(2)));
    h2v v = __builtin_bit_cast(h2v, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    return make_float2((float)v.x, (float)v.y);
  }
}

template <typename S, int NL>
__global__ __launch_bounds__(256) void gather_probe_kernel(const void* vol, uint32_t bytes,
                                                           uint32_t off_j, uint32_t off_i,
                                                           uint32_t step_bytes, int PY, int PZ,
                                                           int ncols, int nz, int nsteps,
                                                           float* sink) {
  const int col = (int)(blockIdx.x * 4 + (threadIdx.x >> 6)) % ncols;  // j
  const int z = (int)(blockIdx.y * 64 + (threadIdx.x & 63)) % nz;
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc((void*)vol, 0, (int)bytes, 0x00020000);
  uint32_t voff = (uint32_t)(((1 * PY) + col + 1) * PZ + z + 1) * (uint32_t)sizeof(S);
  float s0 = 0.f, s1 = 0.f;
  for (int k = 0; k < nsteps; ++k) {
    const float2 a = pld2<S>(r, voff, 0);
    const float2 b = pld2<S>(r, voff, off_j);
    s0 = s0 + a.x;
    s1 = s1 + a.y;
    s0 = s0 + b.x;
    s1 = s1 + b.y;
    if (NL == 4) {
      const float2 c = pld2<S>(r, voff, off_i);
      const float2 d = pld2<S>(r, voff, off_i + off_j);
      s0 = s0 + c.x;
      s1 = s1 + c.y;
      s0 = s0 + d.x;
      s1 = s1 + d.y;
    }
    voff += step_bytes;
  }
  if (s0 + s1 == 1234.5f) sink[threadIdx.x] = s0;  // keeps the loads; never true for the probe data
}

hipError_t launch_gather_probe(const void* vol_z, int st_dt, int NX, int NY, int NZ, int nblocks_x,
                               int nblocks_y, int nsteps, int nloads, float* sink,
                               hipStream_t st) {
  const int PY = NY + 2 * HALO, PZ = NZ + 2 * HALO;
  const uint32_t es = st_dt == 0 ? 4u : 2u;
  const uint32_t bytes = (uint32_t)((size_t)(NX + 2 * HALO) * PY * PZ * es);
  const uint32_t off_j = (uint32_t)PZ * es, off_i = (uint32_t)PY * PZ * es;
  const int steps = nsteps < NX ? nsteps : NX;
  const dim3 g((unsigned)nblocks_x, (unsigned)nblocks_y);
  if (st_dt == 0) {
    if (nloads == 4)
      gather_probe_kernel<float, 4><<<g, 256, 0, st>>>(vol_z, bytes, off_j, off_i, off_i, PY, PZ,
                                                       NY, NZ, steps, sink);
    else
      gather_probe_kernel<float, 2><<<g, 256, 0, st>>>(vol_z, bytes, off_j, off_i, off_i, PY, PZ,
                                                       NY, NZ, steps, sink);
  } else {
    if (nloads == 4)
      gather_probe_kernel<_Float16, 4><<<g, 256, 0, st>>>(vol_z, bytes, off_j, off_i, off_i, PY, PZ,
                                                          NY, NZ, steps, sink);
    else
      gather_probe_kernel<_Float16, 2><<<g, 256, 0, st>>>(vol_z, bytes, off_j, off_i, off_i, PY, PZ,
                                                          NY, NZ, steps, sink);
  }
  return hipGetLastError();
}

}  // namespace mexct
