// probe.hip -- gather-throughput microbenchmark (measurement only).
//
// Issues the projectors' own load instructions with ideal coherence, with
// only the VALU work that keeps the loads alive, so the rate it reaches is
// the memory pipeline's ceiling for that pattern (the "measured gather
// roofline"):
//
// Each probe runs unrolled 1, 2, 4 or 8 times (up to that many samples'
// loads in flight per wave); the measured ceiling is the best of them.
//
//   FP: f32, the x-pair (DUAL) z-fastest volume, per sample two 16-byte
//       loads (rows j and j+1); f16/bf16, the corner-quad volume, one
//       16-byte load per sample.  The 64 lanes of a wave hold 64
//       consecutive z, each wave one y column, and every lane steps along x.
//   BP: the row-paired haloed sinogram; per update one 16-byte (f32) or
//       8-byte (f16/bf16) load of two adjacent pair elements.  The 64 lanes
//       of a wave hold 64 consecutive detector columns; every lane walks
//       down the rows of one view, wrapping.  The blocks of one view are
//       consecutive in the grid, so each view's slice is re-read from cache
//       W times (once per voxel on each of its rays) as the backprojector
//       re-reads it.
#include "common.h"
#include "internal.h"

namespace mexct {

template <int LB>
__device__ __forceinline__ float probe_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff) {
  if constexpr (LB == 16) {
    typedef uint32_t u4v __attribute__((ext_vector_type(4)));
    const u4v v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0);
    return __uint_as_float(v.x ^ v.y ^ v.z ^ v.w);
  } else {
    typedef uint32_t u2v __attribute__((ext_vector_type(2)));
    const u2v v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0);
    return __uint_as_float(v.x ^ v.y);
  }
}

// FP pattern.  Element (x, y, z) at ((x+2)*QY + y+2)*QZ + z+2, E bytes each.
// QUAD: one 16-byte load per sample; DUAL: two 16-byte loads, rows y and
// y+1 (QZ*E apart).
template <int QUAD, int U>
__global__ __launch_bounds__(256) void fp_probe_kernel(const void* vol, uint32_t bytes,
                                                       uint32_t E, int QY, int QZ, int ncols,
                                                       int nz, int nsteps, float* sink) {
  const int col = (int)(blockIdx.x * 4 + (threadIdx.x >> 6)) % ncols;  // y
  const int z = (int)(blockIdx.y * 64 + (threadIdx.x & 63)) % nz;
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc((void*)vol, 0, (int)bytes, 0x00020000);
  uint32_t voff = (uint32_t)((2 * QY + col + 2) * QZ + z + 2) * E;
  const uint32_t step = (uint32_t)QY * (uint32_t)QZ * E;
  const uint32_t oj = (uint32_t)QZ * E;
  float s = 0.f;
  // unrolled U times: up to U samples' loads in flight per wave
#pragma unroll U
  for (int k = 0; k < nsteps; ++k) {
    s = s + probe_ld<16>(r, voff);
    if (!QUAD) s = s + probe_ld<16>(r, voff + oj);
    voff += step;
  }
  if (s == 1234.5f) sink[threadIdx.x] = s;  // keeps the loads; never true for the probe data
}

// BP pattern.  Pair element (u, row) of view v at v*VB + (row+2)*RB + (u+2)*E2.
template <int LB, int U>
__global__ __launch_bounds__(256) void bp_probe_kernel(const void* sino, uint32_t bytes,
                                                       uint32_t E2, uint32_t RB, uint32_t VB,
                                                       int W, int H, int nviews, int nsteps,
                                                       float* sink) {
  const int u = (int)(blockIdx.x * 64 + (threadIdx.x & 63)) % W;
  const unsigned groups = gridDim.y / (unsigned)nviews;
  const int view = (int)(blockIdx.y / groups);  // blocks of one view run together
  int row = (int)(((blockIdx.y % groups) * 4 + (threadIdx.x >> 6)) % (unsigned)H);
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc((void*)sino, 0, (int)bytes, 0x00020000);
  const uint32_t base = (uint32_t)view * VB + 2u * RB + (uint32_t)(u + 2) * E2;
  float s = 0.f;
#pragma unroll U
  for (int k = 0; k < nsteps; ++k) {
    s = s + probe_ld<LB>(r, base + (uint32_t)row * RB);
    row = row + 1 == H ? 0 : row + 1;
  }
  if (s == 1234.5f) sink[threadIdx.x] = s;
}

template <int U>
static void probe_launch(const void* buf, int st_dt, int kind, int W, int H, int nviews,
                         float* sink, int64_t* samples, hipStream_t st) {
  const uint32_t es = st_dt == 0 ? 4u : 2u;
  if (kind == 0) {
    const int quad = st_dt != 0;
    const int QX = W + 2 * HALO - 1;
    const int QY = W + 2 * HALO - (quad ? 1 : 0), QZ = H + 2 * HALO;
    const uint32_t E = (quad ? 4u : 2u) * es;
    const uint32_t bytes = (uint32_t)((size_t)QX * QY * QZ * E);
    const dim3 g((unsigned)((W + 3) / 4), (unsigned)((H + 63) / 64 * nviews));
    if (quad)
      fp_probe_kernel<1, U><<<g, 256, 0, st>>>(buf, bytes, E, QY, QZ, W, H, W, sink);
    else
      fp_probe_kernel<0, U><<<g, 256, 0, st>>>(buf, bytes, E, QY, QZ, W, H, W, sink);
    *samples = (int64_t)g.x * 4 * (int64_t)g.y * 64 * W;
  } else {
    const uint32_t E2 = 2u * es, RB = (uint32_t)(W + 2 * HALO) * E2;
    const uint32_t VB = RB * (uint32_t)(H + 2 * HALO - 1);
    const uint32_t bytes = VB * (uint32_t)nviews;
    const unsigned gx = (unsigned)((W + 63) / 64);
    const unsigned groups = (unsigned)((size_t)W * W / ((size_t)4 * 64 * gx));
    const dim3 g(gx, (unsigned)nviews * (groups > 0 ? groups : 1));
    if (st_dt == 0)
      bp_probe_kernel<16, U><<<g, 256, 0, st>>>(buf, bytes, E2, RB, VB, W, H, nviews, H, sink);
    else
      bp_probe_kernel<8, U><<<g, 256, 0, st>>>(buf, bytes, E2, RB, VB, W, H, nviews, H, sink);
    *samples = (int64_t)g.x * (int64_t)g.y * 256 * H;
  }
}

// pattern bit 0: FP (0) or BP (1) loads; bits 4-5: loop unrolled 1, 2, 4 or 8 times
hipError_t launch_gather_probe(const void* buf, int st_dt, int pattern, int W, int H, int nviews,
                               float* sink, int64_t* samples, hipStream_t st) {
  const int kind = pattern & 1;
  switch ((pattern >> 4) & 3) {
    case 0: probe_launch<1>(buf, st_dt, kind, W, H, nviews, sink, samples, st); break;
    case 1: probe_launch<2>(buf, st_dt, kind, W, H, nviews, sink, samples, st); break;
    case 2: probe_launch<4>(buf, st_dt, kind, W, H, nviews, sink, samples, st); break;
    default: probe_launch<8>(buf, st_dt, kind, W, H, nviews, sink, samples, st); break;
  }
  return hipGetLastError();
}

}  // namespace mexct
