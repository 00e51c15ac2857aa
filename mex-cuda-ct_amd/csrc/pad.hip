// pad.hip -- storage conversion and halo padding of gathered operands.
//
// The reference binds its volume/projection to a texture with clamp or
// border addressing (CUDAmex_FP.cu:405-409, CUDAmex_BP.cu:270-273,
// CUDAmex_general3D_FP.cu:217-221, CUDAmex_oscIter.cu:417-462).  Here the
// operand is copied once per call into a [N+4] haloed layout whose halo
// holds what those address modes return outside the array, converted to
// the storage dtype (f32, f16 or bf16) that the gather kernels read.
#include "common.h"
#include "internal.h"

namespace mexct {

template <typename S>
__global__ __launch_bounds__(256) void pad3_kernel(const void* __restrict__ in, int in_dt,
                                                   S* __restrict__ out, int nx, int ny, int nz,
                                                   int mode) {
  const int PX = nx + 2 * HALO, PY = ny + 2 * HALO, PZ = nz + 2 * HALO;
  const size_t total = (size_t)PX * PY * PZ;
  for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
       o += (size_t)gridDim.x * blockDim.x) {
    int px = (int)(o % PX);
    size_t r = o / PX;
    int py = (int)(r % PY);
    int pz = (int)(r / PY);
    int i = px - HALO, j = py - HALO, k = pz - HALO;
    float v;
    if (mode == ADDR_CLAMP) {
      i = clampi(i, 0, nx - 1);
      j = clampi(j, 0, ny - 1);
      k = clampi(k, 0, nz - 1);
      v = load_any(in, in_dt, (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k));
    } else if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) {
      v = 0.f;
    } else {
      v = load_any(in, in_dt, (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k));
    }
    out[o] = Store<S>::from(v);
  }
}

// Row-grouped haloed sinogram for the backprojector (R = 2 or 4 rows):
// out[((view*PR + py)*PX + px)*R + r] = P(px-2, py-zh+r), r < R, py in
// [0, PR), PR = h+2*zh-1, PX = w+4, with clamp (edge copies) or border
// (zeros) outside; zh >= 2 rows of halo above and below.
// R = 2: element (py, px) and its neighbour (py, px+1) hold the four taps of
// a bilinear sample at (px-2, py-zh), one 16-byte (f32) or 8-byte load.
// R = 4 (2-byte storage): the 16-byte load of two elements holds rows
// py-zh .. py-zh+3 of two columns, the taps of two consecutive z-voxels whose
// rows differ by at most 2.
template <typename S, int R>
__global__ __launch_bounds__(256) void pad2_kernel(const void* __restrict__ in, int in_dt,
                                                   S* __restrict__ out, int w, int h, int n,
                                                   int mode, const uint32_t* scale_max, int zh) {
  const int sexp = pow2_exp(scale_max);
  const int PX = w + 2 * HALO, PR = h + 2 * zh - 1;
  const size_t total = (size_t)PX * PR * n;
  for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
       o += (size_t)gridDim.x * blockDim.x) {
    const int px = (int)(o % PX);
    const size_t r = o / PX;
    const int py = (int)(r % PR);
    const int view = (int)(r / PR);
    const size_t vbase = (size_t)w * h * view;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int i = px - HALO, j = py - zh + q;
      float v;
      if (mode == ADDR_CLAMP) {
        i = clampi(i, 0, w - 1);
        j = clampi(j, 0, h - 1);
        v = load_any(in, in_dt, (size_t)i + (size_t)w * j + vbase);
      } else if (i < 0 || i >= w || j < 0 || j >= h) {
        v = 0.f;
      } else {
        v = load_any(in, in_dt, (size_t)i + (size_t)w * j + vbase);
      }
      out[R * o + q] = Store<S>::from(ldexpf(v, -sexp));
    }
  }
}

// z-fastest haloed copy for the standard-geometry projectors:
// out[(x+2)*PY*PZ + (y+2)*PZ + (z+2)], PY = ny+4, PZ = nz+4, from an x-fastest
// input element (x,y,z) at in[x + in_sy*y + in_sz*z].  A 32x32 (z,x) tile is
// staged in LDS so both the read (along x) and the write (along z) coalesce.
template <typename S>
__global__ __launch_bounds__(256) void pad3z_kernel(const void* __restrict__ in, int in_dt,
                                                    int64_t in_sy, int64_t in_sz,
                                                    S* __restrict__ out, int nx, int ny, int nz,
                                                    int mode, const uint32_t* scale_max) {
  __shared__ float tile[32][33];
  const int sexp = pow2_exp(scale_max);
  const int PX = nx + 2 * HALO, PY = ny + 2 * HALO, PZ = nz + 2 * HALO;
  const int pz0 = blockIdx.x * 32, px0 = blockIdx.y * 32, py = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  int j = py - HALO;
  bool jin = (j >= 0 && j < ny);
  if (mode == ADDR_CLAMP) {
    j = clampi(j, 0, ny - 1);
    jin = true;
  }
  for (int r = 0; r < 4; ++r) {
    const int lz = ty + 8 * r;
    int i = px0 + tx - HALO, k = pz0 + lz - HALO;
    bool ok = jin && (i >= 0 && i < nx && k >= 0 && k < nz);
    if (mode == ADDR_CLAMP) {
      i = clampi(i, 0, nx - 1);
      k = clampi(k, 0, nz - 1);
      ok = true;
    }
    float v = 0.f;
    if (ok && px0 + tx < PX && pz0 + lz < PZ)
      v = load_any(in, in_dt, (size_t)i + (size_t)in_sy * j + (size_t)in_sz * k);
    tile[lz][tx] = ldexpf(v, -sexp);
  }
  __syncthreads();
  for (int r = 0; r < 4; ++r) {
    const int lx = ty + 8 * r;
    const int px = px0 + lx, pz = pz0 + tx;
    if (px < PX && pz < PZ) out[((size_t)px * PY + py) * PZ + pz] = Store<S>::from(tile[tx][lx]);
  }
}

// The two values of a pair element in one store: 8 bytes for f32 (the
// element's address is a multiple of 8), 4 for 2-byte storage.
template <typename S>
__device__ __forceinline__ void store_pair(S* p, float a, float b) {
  if constexpr (__is_same(S, float)) {
    typedef float f2v __attribute__((ext_vector_type(2)));
    *(f2v*)p = f2v{a, b};
  } else {
    const uint32_t lo = __builtin_bit_cast(uint16_t, Store<S>::from(a));
    const uint32_t hi = __builtin_bit_cast(uint16_t, Store<S>::from(b));
    *(uint32_t*)p = lo | (hi << 16);
  }
}

// x-pair copy (see launch_pad3d).  A block covers 32 z x 32 x of one y: the
// (z, x..x+32) tile of input row y is staged in LDS (read along x), then
// each lane writes one pair element (along z).
template <typename S>
__global__ __launch_bounds__(256) void pad3d_kernel(const void* __restrict__ in, int in_dt,
                                                    int64_t in_sy, int64_t in_sz,
                                                    S* __restrict__ out, int nx, int ny, int nz,
                                                    int mode, const uint32_t* scale_max) {
  __shared__ float tile[32][33];
  const int sexp = pow2_exp(scale_max);
  const int DX = nx + 2 * HALO - 1, DY = ny + 2 * HALO, DZ = nz + 2 * HALO;
  const int dz0 = blockIdx.x * 32, dx0 = blockIdx.y * 32, dy = blockIdx.z;
  for (int e = threadIdx.x; e < 32 * 33; e += 256) {
    const int lz = e / 33, lx = e - lz * 33;
    int i = dx0 + lx - HALO, j = dy - HALO, k = dz0 + lz - HALO;
    float v = 0.f;
    if (mode == ADDR_CLAMP) {
      i = clampi(i, 0, nx - 1);
      j = clampi(j, 0, ny - 1);
      k = clampi(k, 0, nz - 1);
      v = load_any(in, in_dt, (size_t)i + (size_t)in_sy * j + (size_t)in_sz * k);
    } else if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) {
      v = load_any(in, in_dt, (size_t)i + (size_t)in_sy * j + (size_t)in_sz * k);
    }
    tile[lz][lx] = ldexpf(v, -sexp);
  }
  __syncthreads();
  const int tz = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int p = 0; p < 4; ++p) {
    const int lx = ty + 8 * p;
    const int dx = dx0 + lx, dz = dz0 + tz;
    if (dx < DX && dz < DZ) {
      const size_t e = ((size_t)dx * DY + dy) * DZ + dz;
      store_pair<S>(out + 2 * e, tile[tz][lx], tile[tz][lx + 1]);
    }
  }
}

// Corner-quad copy (see launch_pad3q).  A block covers 32 z x 32 x of one
// y: the (z, x..x+32) tiles of input rows y and y+1 are staged in LDS (read
// along x), then each lane writes one quad element (along z).
template <typename S>
struct Quad;
template <>
struct Quad<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static void put(float* out, size_t e, float a, float b, float c, float d) {
    v4 q = {a, b, c, d};
    *(v4*)(out + 4 * e) = q;
  }
};
// f16: the difference quad {a, b - a, c - a, (d - b) - (c - a)}, each term
// computed in float and rounded to f16 once (a, b, c, d = V(x,y), V(x,y+1),
// V(x+1,y), V(x+1,y+1))
template <>
struct Quad<_Float16> {
  typedef _Float16 v4 __attribute__((ext_vector_type(4)));
  __device__ static void put(_Float16* out, size_t e, float a, float b, float c, float d) {
    const float dx = c - a;
    v4 q = {(_Float16)a, (_Float16)(b - a), (_Float16)dx, (_Float16)((d - b) - dx)};
    *(v4*)(out + 4 * e) = q;
  }
};
template <>
struct Quad<bf16_t> {
  typedef uint16_t v4 __attribute__((ext_vector_type(4)));
  __device__ static void put(bf16_t* out, size_t e, float a, float b, float c, float d) {
    v4 q = {f32_to_bf16(a), f32_to_bf16(b), f32_to_bf16(c), f32_to_bf16(d)};
    *(v4*)(out + 4 * e) = q;
  }
};

template <typename S>
__global__ __launch_bounds__(256) void pad3q_kernel(const void* __restrict__ in, int in_dt,
                                                    int64_t in_sy, int64_t in_sz,
                                                    S* __restrict__ out, int nx, int ny, int nz,
                                                    int mode, const uint32_t* scale_max) {
  __shared__ float tile[2][32][33];
  const int sexp = pow2_exp(scale_max);
  const int QX = nx + 2 * HALO - 1, QY = ny + 2 * HALO - 1, QZ = nz + 2 * HALO;
  const int qz0 = blockIdx.x * 32, qx0 = blockIdx.y * 32, qy = blockIdx.z;
  for (int e = threadIdx.x; e < 2 * 32 * 33; e += 256) {
    const int r = e / (32 * 33);
    const int rem = e - r * (32 * 33);
    const int lz = rem / 33, lx = rem - lz * 33;
    int i = qx0 + lx - HALO, j = qy + r - HALO, k = qz0 + lz - HALO;
    float v = 0.f;
    if (mode == ADDR_CLAMP) {
      i = clampi(i, 0, nx - 1);
      j = clampi(j, 0, ny - 1);
      k = clampi(k, 0, nz - 1);
      v = load_any(in, in_dt, (size_t)i + (size_t)in_sy * j + (size_t)in_sz * k);
    } else if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) {
      v = load_any(in, in_dt, (size_t)i + (size_t)in_sy * j + (size_t)in_sz * k);
    }
    tile[r][lz][lx] = ldexpf(v, -sexp);
  }
  __syncthreads();
  const int tz = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int p = 0; p < 4; ++p) {
    const int lx = ty + 8 * p;
    const int qx = qx0 + lx, qz = qz0 + tz;
    if (qx < QX && qz < QZ)
      Quad<S>::put(out, ((size_t)qx * QY + qy) * QZ + qz, tile[0][tz][lx], tile[1][tz][lx],
                   tile[0][tz][lx + 1], tile[1][tz][lx + 1]);
  }
}

__global__ __launch_bounds__(256) void unpad3_kernel(const float* __restrict__ padded,
                                                     float* __restrict__ out, int nx, int ny,
                                                     int nz) {
  const size_t total = (size_t)nx * ny * nz;
  const size_t PX = nx + 2 * HALO, PY = ny + 2 * HALO;
  for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
       o += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(o % nx);
    size_t r = o / nx;
    int j = (int)(r % ny);
    int k = (int)(r / ny);
    out[o] = padded[(size_t)(i + HALO) + PX * ((size_t)(j + HALO) + PY * (size_t)(k + HALO))];
  }
}

__global__ __launch_bounds__(256) void fill_kernel(float* p, float v, size_t n) {
  for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < n;
       o += (size_t)gridDim.x * blockDim.x)
    p[o] = v;
}

static unsigned grid_for(size_t total) {
  size_t b = (total + 255) / 256;
  if (b > 65536u) b = 65536u;
  if (b < 1) b = 1;
  return (unsigned)b;
}

hipError_t launch_pad3(const void* in, int in_dt, void* out, int st_dt, int nx, int ny, int nz,
                       int mode, hipStream_t st) {
  size_t total = (size_t)(nx + 2 * HALO) * (ny + 2 * HALO) * (nz + 2 * HALO);
  unsigned g = grid_for(total);
  if (st_dt == 0)
    pad3_kernel<float><<<g, 256, 0, st>>>(in, in_dt, (float*)out, nx, ny, nz, mode);
  else if (st_dt == 1)
    pad3_kernel<_Float16><<<g, 256, 0, st>>>(in, in_dt, (_Float16*)out, nx, ny, nz, mode);
  else
    pad3_kernel<bf16_t><<<g, 256, 0, st>>>(in, in_dt, (bf16_t*)out, nx, ny, nz, mode);
  return hipGetLastError();
}

hipError_t launch_pad2(const void* in, int in_dt, void* out, int st_dt, int w, int h, int n,
                       int mode, hipStream_t st, const uint32_t* scale_max, int rows, int zh) {
  if (zh < HALO) zh = HALO;
  size_t total = (size_t)(w + 2 * HALO) * (h + 2 * zh - 1) * n;
  unsigned g = grid_for(total);
  if (st_dt == 0)
    pad2_kernel<float, 2><<<g, 256, 0, st>>>(in, in_dt, (float*)out, w, h, n, mode, scale_max, zh);
  else if (st_dt == 1 && rows == 4)
    pad2_kernel<_Float16, 4><<<g, 256, 0, st>>>(in, in_dt, (_Float16*)out, w, h, n, mode,
                                                scale_max, zh);
  else if (st_dt == 1)
    pad2_kernel<_Float16, 2><<<g, 256, 0, st>>>(in, in_dt, (_Float16*)out, w, h, n, mode,
                                                scale_max, zh);
  else if (rows == 4)
    pad2_kernel<bf16_t, 4><<<g, 256, 0, st>>>(in, in_dt, (bf16_t*)out, w, h, n, mode, scale_max, zh);
  else
    pad2_kernel<bf16_t, 2><<<g, 256, 0, st>>>(in, in_dt, (bf16_t*)out, w, h, n, mode, scale_max, zh);
  return hipGetLastError();
}

hipError_t launch_pad3d(const void* in, int in_dt, int64_t in_sy, int64_t in_sz, void* out,
                        int st_dt, int nx, int ny, int nz, int mode, hipStream_t st,
                        const uint32_t* scale_max) {
  const dim3 g((unsigned)((nz + 2 * HALO + 31) / 32), (unsigned)((nx + 2 * HALO - 1 + 31) / 32),
               (unsigned)(ny + 2 * HALO));
  if (st_dt == 0)
    pad3d_kernel<float><<<g, 256, 0, st>>>(in, in_dt, in_sy, in_sz, (float*)out, nx, ny, nz, mode,
                                           scale_max);
  else if (st_dt == 1)
    pad3d_kernel<_Float16><<<g, 256, 0, st>>>(in, in_dt, in_sy, in_sz, (_Float16*)out, nx, ny, nz,
                                              mode, scale_max);
  else
    pad3d_kernel<bf16_t><<<g, 256, 0, st>>>(in, in_dt, in_sy, in_sz, (bf16_t*)out, nx, ny, nz,
                                            mode, scale_max);
  return hipGetLastError();
}

hipError_t launch_pad3q(const void* in, int in_dt, int64_t in_sy, int64_t in_sz, void* out,
                        int st_dt, int nx, int ny, int nz, int mode, hipStream_t st,
                        const uint32_t* scale_max) {
  const dim3 g((unsigned)((nz + 2 * HALO + 31) / 32), (unsigned)((nx + 2 * HALO - 1 + 31) / 32),
               (unsigned)(ny + 2 * HALO - 1));
  if (st_dt == 0)
    pad3q_kernel<float><<<g, 256, 0, st>>>(in, in_dt, in_sy, in_sz, (float*)out, nx, ny, nz, mode, scale_max);
  else if (st_dt == 1)
    pad3q_kernel<_Float16><<<g, 256, 0, st>>>(in, in_dt, in_sy, in_sz, (_Float16*)out, nx, ny, nz,
                                              mode, scale_max);
  else
    pad3q_kernel<bf16_t><<<g, 256, 0, st>>>(in, in_dt, in_sy, in_sz, (bf16_t*)out, nx, ny, nz,
                                            mode, scale_max);
  return hipGetLastError();
}

hipError_t launch_pad3z(const void* in, int in_dt, int64_t in_sy, int64_t in_sz, void* out,
                        int st_dt, int nx, int ny, int nz, int mode, hipStream_t st,
                        const uint32_t* scale_max) {
  const dim3 g((unsigned)((nz + 2 * HALO + 31) / 32), (unsigned)((nx + 2 * HALO + 31) / 32),
               (unsigned)(ny + 2 * HALO));
  if (st_dt == 0)
    pad3z_kernel<float><<<g, 256, 0, st>>>(in, in_dt, in_sy, in_sz, (float*)out, nx, ny, nz, mode, scale_max);
  else if (st_dt == 1)
    pad3z_kernel<_Float16><<<g, 256, 0, st>>>(in, in_dt, in_sy, in_sz, (_Float16*)out, nx, ny, nz,
                                              mode, scale_max);
  else
    pad3z_kernel<bf16_t><<<g, 256, 0, st>>>(in, in_dt, in_sy, in_sz, (bf16_t*)out, nx, ny, nz,
                                            mode, scale_max);
  return hipGetLastError();
}

hipError_t launch_unpad3(const float* padded, float* out, int nx, int ny, int nz, hipStream_t st) {
  unpad3_kernel<<<grid_for((size_t)nx * ny * nz), 256, 0, st>>>(padded, out, nx, ny, nz);
  return hipGetLastError();
}

hipError_t launch_fill(float* p, float v, size_t n, hipStream_t st) {
  fill_kernel<<<grid_for(n), 256, 0, st>>>(p, v, n);
  return hipGetLastError();
}

}  // namespace mexct
