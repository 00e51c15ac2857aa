// bp.hip -- voxel-driven backprojectors (standard geometries and the OSC
// dual backprojection with its update) and the general-3D ray-driven
// scatter, for gfx950.
//
// Voxel-driven BP (CUDAmex_BP.cu:44-221): the reference launches one kernel
// per view and read-modify-writes the whole volume each time (8 B per
// update).  Here one launch covers every view: a 64-thread workgroup owns an
// 8x8 tile of voxel columns and a run of ZR consecutive z-voxels, keeps the
// ZR sums in registers while it walks the views in order, and writes each
// voxel once.  Every voxel therefore receives its views in the reference's
// order, starting from 0, so the float sum is the reference's sum.
//
// The reference steps the detector row by a running float sum,
// proj_z += U per slice (CUDAmex_BP.cu:208-219).  For par and fan the
// running sum holds k + 0.5 exactly and is computed directly; for cone the
// thread replays the z0 additions that precede its first voxel, so every
// sample lands exactly where the reference's sample lands.
#include "common.h"
#include "internal.h"

namespace mexct {

// Four taps of a row-paired sinogram element pair: {(j0,i0), (j0+1,i0),
// (j0,i0+1), (j0+1,i0+1)} in one load.  f32 and bf16 taps are widened;
// f16 taps stay packed ({(j0,i0), (j0+1,i0)} and {(j0,i0+1), (j0+1,i0+1)}
// per dword) for the mixed-precision lerps.
typedef uint32_t u2v __attribute__((ext_vector_type(2)));
typedef uint32_t u4v __attribute__((ext_vector_type(4)));
template <typename S>
struct Taps4 {
  typedef float4 T;
};
template <>
struct Taps4<_Float16> {
  typedef u2v T;
};
template <typename S>
__device__ __forceinline__ typename Taps4<S>::T bld4(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                     uint32_t soff);
template <>
__device__ __forceinline__ float4 bld4<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                              uint32_t soff) {
  typedef float f4v __attribute__((ext_vector_type(4)));
  f4v v = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  return make_float4(v.x, v.y, v.z, v.w);
}
template <>
__device__ __forceinline__ u2v bld4<_Float16>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                              uint32_t soff) {
  return __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
}
template <>
__device__ __forceinline__ float4 bld4<bf16_t>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                               uint32_t soff) {
  const u2v u = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xffff0000u),
                     __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xffff0000u));
}

constexpr uint32_t kSuper = 4;  // default super-tile side (tiles)

struct BPKArgs {
  const void* sino;     // row-paired haloed sinogram
  uint32_t sino_bytes;  // buffer range (the launcher guarantees < 2^31)
  float* out;
  const float2* sc;
  int W, H, nviews, PW;  // PW: elements (pairs) per row
  int64_t PV;            // pairs per view
  int accumulate;        // 1: continue the sums already in out (a later chunk of views)
  const float* chain;    // cone f32: row coordinate at each z-range start (bp_chain_kernel)
  const uint32_t* sino_max;  // non-null: sino holds P * 2^-e (pow2_exp)
  float v0, SR;
  uint32_t tiles_x, nz, n_logical;
  uint32_t ncols;  // chain columns per (view, range): padded tiles x 64
  uint32_t super_;       // super-tile side in tiles (block order)
  uint32_t item_blocks;  // blocks of one batch item; n_logical = item_blocks * items
  // the z-ranges run one after another, each in a phase of phase_blocks grid
  // blocks of which the first phase_logical work
  uint32_t phase_blocks, phase_logical;
  uint64_t sino_stride;  // bytes between the items' haloed sinograms
  uint64_t out_stride;   // floats between the items' volumes
  float* const* outs;    // non-null: item i's volume at outs[i]
  int loads_only;        // launch the loads-only form
  int zh;                // rows of halo above and below in the sinogram copy
};

// Voxel column (ix, iy) of lane `lane` of tile `tile`: 8x8 tiles ordered in
// super x super super-tiles, row-major inside each.
__device__ __forceinline__ void tile_xy(uint32_t tile, uint32_t lane, uint32_t tiles_x,
                                        uint32_t sup, int& ix, int& iy) {
  const uint32_t st = tile / (sup * sup), in = tile % (sup * sup);
  const uint32_t sbx = (tiles_x + sup - 1) / sup;
  ix = (int)((st % sbx) * sup + in % sup) * 8 + (int)(lane & 7);
  iy = (int)((st / sbx) * sup + in / sup) * 8 + (int)(lane >> 3);
}

// per-view detector geometry of one voxel column
template <int GEOM>
__device__ __forceinline__ void bp_view_geom(float x, float y, float sB, float cB, float SR,
                                             float v0, uint32_t hw, float& U, float& py,
                                             float& pz, float& dz) {
  float s = x * cB + y * sB;
  float t = -x * sB + y * cB;
  if (GEOM == GEOM_PAR) {  // CUDAmex_BP.cu:92-95
    U = 1.f;
    py = t * U + (float)hw + 1.f;
    pz = -v0 * U + v0 + 0.5f;
    dz = U;
  } else if (GEOM == GEOM_FAN) {  // :149-154
    U = SR / (SR - s);
    py = t * U + (float)hw + 1.f;
    pz = 0.5f;
    dz = 1.f;
  } else {  // cone :208-212
    U = SR / (SR - s);
    py = t * U + (float)hw + 1.f;
    pz = -v0 * U + v0 + 0.5f;
    dz = U;
  }
}

template <int GEOM>
__device__ __forceinline__ float advance_rows(float pz, float dz, int z0) {
  if (GEOM == GEOM_CONE) {
    for (int m = 0; m < z0; ++m) pz = pz + dz;
    return pz;
  }
  return pz + (float)z0;  // k + 0.5 held exactly by the running sum
}

// Cone row coordinates at the z-range starts: the reference steps
// proj_z += U once per slice from z = 0 (CUDAmex_BP.cu:208-219), so the
// value at z0 is the float running sum after z0 additions.  One thread per
// (voxel column, view) runs the sum once and stores it at z0 = ZR, 2 ZR, ...,
// which saves every range from replaying the additions before it.
template <int ZR>
__global__ __launch_bounds__(256) void bp_chain_kernel(BPKArgs a, float* chain) {
  const size_t total = (size_t)a.ncols * a.nviews;
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (size_t)gridDim.x * 256) {
    const uint32_t col = (uint32_t)(t % a.ncols);
    const int v = (int)(t / a.ncols);
    int ix, iy;
    tile_xy(col >> 6, col & 63u, a.tiles_x, a.super_, ix, iy);
    if (ix >= a.W || iy >= a.W) continue;
    const uint32_t hw = (uint32_t)a.W / 2u - 1u;
    const float x = (float)ix - ((float)hw + 0.5f);
    const float y = (float)iy - ((float)hw + 0.5f);
    const float2 scv = a.sc[v];
    float U, py, pz, dz;
    bp_view_geom<GEOM_CONE>(x, y, scv.x, scv.y, a.SR, a.v0, hw, U, py, pz, dz);
    float* o = chain + (size_t)v * (a.nz - 1) * a.ncols + col;
    for (uint32_t r = 1; r < a.nz; ++r) {
      for (int m = 0; m < ZR; ++m) pz = pz + dz;
      o[(size_t)(r - 1) * a.ncols] = pz;
    }
  }
}

// SMALL: a view's haloed slice is below 2^24 bytes, so lane offsets are
// exact in float and cost one FMA.
// Bits of one loaded tap group, for the loads-only form.
__device__ __forceinline__ uint32_t load_bits(const float4& q) {
  return __float_as_uint(q.x) ^ __float_as_uint(q.y) ^ __float_as_uint(q.z) ^ __float_as_uint(q.w);
}
__device__ __forceinline__ uint32_t load_bits(const u2v& q) { return q.x ^ q.y; }
__device__ __forceinline__ uint32_t load_bits(const u4v& q) { return q.x ^ q.y ^ q.z ^ q.w; }

// LOADS: the loads-only form (the backprojector's gather ceiling): the same
// blocks, offsets and loads, each load's bits XOR-folded instead of
// interpolated; every voxel of the range receives the fold.
// ZC: clamp each detector row to the copy's rows; without it the host has
// shown that every in-circle voxel's rows lie within the zh-row halo, whose
// edge copies give the clamped taps.
template <int GEOM, int ZR, typename S, bool SMALL, int RQ, bool ZC, bool LOADS = false>
__global__ __launch_bounds__(64) void bp_kernel(BPKArgs a) {
  // Block order: the z-ranges one after another over the whole chip (a
  // phase each), and in a phase (batch item, tile), each XCD taking a
  // contiguous range.  Tiles go in super-tiles (row-major inside), so that
  // the blocks an XCD runs together cover a compact region whose view
  // footprints overlap; the voxels of one z-range project to a band of
  // detector rows whose row-paired copy stays in the memory-side cache
  // across the phase.
  const uint32_t zr = blockIdx.x / a.phase_blocks;
  const uint32_t bl = xcd_remap(blockIdx.x - zr * a.phase_blocks, a.phase_blocks);
  if (bl >= a.phase_logical) return;
  const uint32_t tiles = a.item_blocks / a.nz;
  const uint32_t item = bl / tiles;  // batch item
  const uint32_t tile = bl - item * tiles;
  const uint32_t chain_col = tile * 64u + threadIdx.x;
  int ix, iy;
  tile_xy(tile, threadIdx.x, a.tiles_x, a.super_, ix, iy);
  if (ix >= a.W || iy >= a.W) return;
  const int z0 = (int)zr * ZR;
  const int zc = min(ZR, a.H - z0);
  const size_t WW = (size_t)a.W * a.W;
  float* const ob = a.outs ? a.outs[item] : a.out + item * a.out_stride;
  float* o = ob + (size_t)ix + (size_t)a.W * iy + WW * z0;
  const uint32_t hw = (uint32_t)a.W / 2u - 1u;
  const uint32_t dx = (uint32_t)ix - hw, dy = (uint32_t)iy - hw;
  if (!(dx * dx + dy * dy < hw * hw)) {  // outside the circle of reconstruction
    for (int m = 0; m < zc; ++m) o[(size_t)m * WW] = 0.f;
    return;
  }
  const float x = (float)ix - ((float)hw + 0.5f);
  const float y = (float)iy - ((float)hw + 0.5f);
  const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
      (void*)((const char*)a.sino + item * a.sino_stride), 0, (int)a.sino_bytes, 0x00020000);
  const uint32_t es = (RQ ? 4u : 2u) * (uint32_t)sizeof(S);  // bytes of one row-group element
  const int row_bytes = a.PW * (int)es;          // one detector row of pairs
  float acc[ZR];
#pragma unroll
  for (int m = 0; m < ZR; ++m) acc[m] = 0.f;
  uint32_t fold = 0;
  const int sexp = pow2_exp(a.sino_max ? a.sino_max + item : nullptr);
  if (a.accumulate) {
    const float* q = o;
#pragma unroll
    for (int m = 0; m < ZR; ++m) {
      if (m < zc) acc[m] = ldexpf(*q, -sexp);
      q += WW;
    }
  }

  for (int v = 0; v < a.nviews; ++v) {
    const float2 scv = a.sc[v];
    float U, py, pz, dz;
    bp_view_geom<GEOM>(x, y, scv.x, scv.y, a.SR, a.v0, hw, U, py, pz, dz);
    const float UU = U * U;
    const float xb = py - 0.5f;
    const float fx = floorf(xb);
    const float ax = xb - fx;
    const float fi0 = __builtin_amdgcn_fmed3f(fx, -2.f, (float)a.W);
    // view v's slice starts at a wave-uniform scalar offset; the lane's
    // byte offset of (u = i0, v = j0) is (i0+2)*es + (j0+2)*row_bytes, an
    // integer below 2^24 and therefore exact in float
    const uint32_t soff = (uint32_t)v * (uint32_t)a.PV * es;
    const float colf = (fi0 + (float)HALO) * (float)es + (float)(a.zh * row_bytes);
    const float rbf = (float)row_bytes;
    const uint32_t colb = (uint32_t)((int)fi0 + HALO) * es;  // integer form (large slices)
    // groups of 8 z-voxels: offsets and loads of the whole group are issued
    // before its interpolation
    constexpr int G = 8;
    if constexpr (sizeof(S) == 2 && RQ) {
      // 2-byte storage, row-quad sinogram: one 16-byte load per two
      // consecutive z-voxels.  Their rows j0 <= j1 <= j0 + 2 (the row step
      // is below 2, which the host checks), so the pair (j1, j1+1) of each
      // column is the 4-row run shifted down by j1 - j0 halves.
      float zb = __builtin_fmaf((float)z0, dz, pz) - 0.5f;
      const uint32_t cb = (uint32_t)((int)fi0 + HALO) * es + (uint32_t)(a.zh * row_bytes);
      const uint32_t ax2 = pk_h2(ax);
      // Groups of GP voxel pairs, software-pipelined: the loads of group
      // g+1 are issued before group g is interpolated, so loads stay in
      // flight while the wave computes.
      constexpr int GP = 4;
      constexpr int NG = ZR / (2 * GP);
      float az0[2][GP], az1[2][GP];
      uint32_t sh[2][GP];
      u4v q[2][GP];
      auto issue = [&](int b) {
#pragma unroll
        for (int g = 0; g < GP; ++g) {
          const float zb1 = zb + dz;
          az0[b][g] = __builtin_amdgcn_fractf(zb);
          az1[b][g] = __builtin_amdgcn_fractf(zb1);
          const int j0 = ZC ? clamp_m2(cvt_flr(zb), a.H) : cvt_flr(zb);
          const int j1 = ZC ? clamp_m2(cvt_flr(zb1), a.H) : cvt_flr(zb1);
          sh[b][g] = (uint32_t)(j1 - j0) << 4;
          q[b][g] = __builtin_amdgcn_raw_buffer_load_b128(r, (uint32_t)__mul24(j0, row_bytes) + cb,
                                                          soff, 0);
          zb = zb1 + dz;
        }
      };
      auto consume = [&](int b, int m0) {
#pragma unroll
        for (int g = 0; g < GP; ++g) {
          // q.x, q.y: rows j0..j0+3 of column i0; q.z, q.w: of column i0+1
          const u4v v = q[b][g];
          if constexpr (LOADS) {
            fold ^= load_bits(v);
            continue;
          }
          if constexpr (__is_same(S, _Float16) && kPackedF16Lerp) {
            // all four rows lerped across the columns, then the second
            // voxel's pair shifted out of them (the lerp is elementwise, so
            // this equals lerping the shifted rows: one shift instead of two)
            const uint32_t h01 = pk_lerp_h(v.x, v.z, ax2);  // rows j0, j0+1
            const uint32_t h23 = pk_lerp_h(v.y, v.w, ax2);  // rows j0+2, j0+3
            const uint32_t h1 = (uint32_t)((((uint64_t)h23 << 32) | h01) >> sh[b][g]);
            acc[m0 + 2 * g] = __builtin_fmaf(UU, lerp_x(h01, az0[b][g]), acc[m0 + 2 * g]);
            acc[m0 + 2 * g + 1] = __builtin_fmaf(UU, lerp_x(h1, az1[b][g]), acc[m0 + 2 * g + 1]);
            continue;
          }
          const uint32_t A1 = (uint32_t)((((uint64_t)v.y << 32) | v.x) >> sh[b][g]);
          const uint32_t B1 = (uint32_t)((((uint64_t)v.w << 32) | v.z) >> sh[b][g]);
          float c0, c1, d0, d1;
          if constexpr (__is_same(S, _Float16)) {
            c0 = lerp_h<0>(v.x, v.z, ax);
            c1 = lerp_h<1>(v.x, v.z, ax);
            d0 = lerp_h<0>(A1, B1, ax);
            d1 = lerp_h<1>(A1, B1, ax);
          } else {
            c0 = lerp(__uint_as_float(v.x << 16), __uint_as_float(v.z << 16), ax);
            c1 = lerp(__uint_as_float(v.x & 0xffff0000u), __uint_as_float(v.z & 0xffff0000u), ax);
            d0 = lerp(__uint_as_float(A1 << 16), __uint_as_float(B1 << 16), ax);
            d1 = lerp(__uint_as_float(A1 & 0xffff0000u), __uint_as_float(B1 & 0xffff0000u), ax);
          }
          acc[m0 + 2 * g] = __builtin_fmaf(UU, lerp(c0, c1, az0[b][g]), acc[m0 + 2 * g]);
          acc[m0 + 2 * g + 1] = __builtin_fmaf(UU, lerp(d0, d1, az1[b][g]), acc[m0 + 2 * g + 1]);
        }
      };
      issue(0);
#pragma unroll
      for (int gi = 0; gi < NG; ++gi) {
        if (gi + 1 < NG) issue((gi + 1) & 1);
        __builtin_amdgcn_sched_barrier(0);
        consume(gi & 1, gi * 2 * GP);
      }
      continue;
    }
    if constexpr (sizeof(S) == 2) {
      // 2-byte storage: the row coordinate starts at z0 directly and steps in
      // texel coordinates (no replay of the reference's running sum); its
      // rounding differs by a few ulps, far below the storage rounding
      float zb = __builtin_fmaf((float)z0, dz, pz) - 0.5f;
      const uint32_t cb = colb + (uint32_t)(a.zh * row_bytes);
#pragma unroll
      for (int m0 = 0; m0 < ZR; m0 += G) {
        float az[G];
        typename Taps4<S>::T q[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          az[g] = __builtin_amdgcn_fractf(zb);
          const int j0 = ZC ? clamp_m2(cvt_flr(zb), a.H) : cvt_flr(zb);
          q[g] = bld4<S>(r, (uint32_t)__mul24(j0, row_bytes) + cb, soff);
          zb = zb + dz;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int g = 0; g < G; ++g) {
          float c0, c1;
          if constexpr (LOADS) {
            fold ^= load_bits(q[g]);
            continue;
          }
          if constexpr (__is_same(S, _Float16) && kPackedF16Lerp) {
            acc[m0 + g] =
                __builtin_fmaf(UU, lerp_x(pk_lerp_h(q[g].x, q[g].y, pk_h2(ax)), az[g]), acc[m0 + g]);
            continue;
          } else if constexpr (__is_same(S, _Float16)) {
            c0 = lerp_h<0>(q[g].x, q[g].y, ax);
            c1 = lerp_h<1>(q[g].x, q[g].y, ax);
          } else {
            c0 = lerp(q[g].x, q[g].z, ax);
            c1 = lerp(q[g].y, q[g].w, ax);
          }
          acc[m0 + g] = __builtin_fmaf(UU, lerp(c0, c1, az[g]), acc[m0 + g]);
        }
      }
      continue;
    }
    if (GEOM == GEOM_CONE && a.chain) {
      if (zr > 0) pz = a.chain[((size_t)v * (a.nz - 1) + (zr - 1)) * a.ncols + chain_col];
    } else {
      pz = advance_rows<GEOM>(pz, dz, z0);
    }
#pragma unroll
    for (int m0 = 0; m0 < ZR; m0 += G) {
      float az[G];
      typename Taps4<S>::T q[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const float zb = pz - 0.5f;
        const float fz = floorf(zb);
        az[g] = zb - fz;
        const float fj0 = ZC ? __builtin_amdgcn_fmed3f(fz, -2.f, (float)a.H) : fz;
        const uint32_t voff =
            SMALL ? (uint32_t)__builtin_fmaf(fj0, rbf, colf)
                  : (uint32_t)((int)fj0 + a.zh) * (uint32_t)row_bytes + colb;
        q[g] = bld4<S>(r, voff, soff);
        pz = pz + dz;
      }
      // keep the group's loads ahead of its interpolation (the scheduler
      // otherwise interleaves them and leaves one or two in flight)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        float c0, c1;  // rows j0 and j0+1, lerped between columns i0 and i0+1
        if constexpr (LOADS) {
          fold ^= load_bits(q[g]);
          continue;
        }
        if constexpr (__is_same(S, _Float16)) {
          c0 = lerp_h<0>(q[g].x, q[g].y, ax);
          c1 = lerp_h<1>(q[g].x, q[g].y, ax);
        } else {
          c0 = lerp(q[g].x, q[g].z, ax);
          c1 = lerp(q[g].y, q[g].w, ax);
        }
        acc[m0 + g] = acc[m0 + g] + UU * lerp(c0, c1, az[g]);
      }
    }
  }
  float* q = o;
#pragma unroll
  for (int m = 0; m < ZR; ++m) {
    if (m < zc) *q = LOADS ? __uint_as_float(fold) : ldexpf(acc[m], sexp);
    q += WW;
  }
}

static unsigned padded_grid(uint32_t n_logical) { return (n_logical + 7u) & ~7u; }

template <int ZR, typename S, bool SMALL, int RQ, bool ZC>
static void bp_dispatch_g(const BPKArgs& k0, int geom, hipStream_t st) {
  BPKArgs k = k0;  // z-ranges outermost (see bp_kernel)
  k.phase_logical = k.item_blocks / k.nz * (k.n_logical / k.item_blocks);
  k.phase_blocks = padded_grid(k.phase_logical);
  const unsigned g = k.phase_blocks * k.nz;
  if (k.loads_only) {  // cone only (the host checks)
    bp_kernel<GEOM_CONE, ZR, S, SMALL, RQ, ZC, true><<<g, 64, 0, st>>>(k);
    return;
  }
  switch (geom) {
    case GEOM_PAR: bp_kernel<GEOM_PAR, ZR, S, SMALL, RQ, ZC><<<g, 64, 0, st>>>(k); break;
    case GEOM_FAN: bp_kernel<GEOM_FAN, ZR, S, SMALL, RQ, ZC><<<g, 64, 0, st>>>(k); break;
    default: bp_kernel<GEOM_CONE, ZR, S, SMALL, RQ, ZC><<<g, 64, 0, st>>>(k); break;
  }
}

template <int ZR, typename S, bool ZC>
static void bp_dispatch(const BPKArgs& k, int geom, int rows, hipStream_t st) {
  if constexpr (sizeof(S) == 2) {
    if (rows == 4) return bp_dispatch_g<ZR, S, false, 1, ZC>(k, geom, st);
  }
  if ((uint64_t)k.PV * 2u * sizeof(S) < (1u << 24))
    bp_dispatch_g<ZR, S, true, 0, ZC>(k, geom, st);
  else
    bp_dispatch_g<ZR, S, false, 0, ZC>(k, geom, st);
}

template <typename S, bool ZC>
static void bp_dispatch_zr(const BPKArgs& k, int geom, int zr, int rows, hipStream_t st) {
  if (zr == 64)
    bp_dispatch<64, S, ZC>(k, geom, rows, st);
  else if (zr == 48)
    bp_dispatch<48, S, ZC>(k, geom, rows, st);
  else
    bp_dispatch<32, S, ZC>(k, geom, rows, st);
}

// the unclamped rows for 2-byte storage only (the host keeps f32 clamped)
template <typename S>
static void bp_dispatch_zc(const BPKArgs& k, int geom, int zr, int rows, bool zclamp,
                           hipStream_t st) {
  if (zclamp || sizeof(S) == 4)
    bp_dispatch_zr<S, true>(k, geom, zr, rows, st);
  else if constexpr (sizeof(S) == 2)
    bp_dispatch_zr<S, false>(k, geom, zr, rows, st);
}

size_t bp_chain_floats_per_view(int W, int H, int zr, int super_) {
  const uint32_t sup = super_ > 0 ? (uint32_t)super_ : kSuper;
  const uint32_t tiles_x = (uint32_t)((W + 7) / 8);
  const uint32_t tpad = (tiles_x + sup - 1) / sup * sup;
  const uint32_t nz = (uint32_t)((H + zr - 1) / zr);
  return nz > 1 ? (size_t)tpad * tpad * 64u * (nz - 1) : 0;
}

// Views go in chunks whose row-paired slices stay below 2 GiB (32-bit
// buffer offsets); each chunk continues the register sums the previous one
// stored, so the per-voxel order and the float sums are unchanged.
hipError_t launch_bp(const BPLaunch& a, hipStream_t st) {
  const int zr = a.zr == 64 || a.zr == 48 ? a.zr : 32;
  const int rows = a.st_dt != 0 && a.rows == 4 ? 4 : 2;
  const size_t es2 = (size_t)rows * (a.st_dt == 0 ? 4u : 2u);
  BPKArgs k;
  k.out = a.out;
  k.sino_max = a.sino_max;
  k.W = a.W;
  k.H = a.H;
  k.PW = a.W + 2 * HALO;
  k.zh = a.zh > HALO ? a.zh : HALO;
  k.PV = (int64_t)(a.W + 2 * HALO) * (a.H + 2 * k.zh - 1);
  k.v0 = float(a.H - 1) / 2;
  k.SR = a.sad;
  k.tiles_x = (uint32_t)((a.W + 7) / 8);
  k.nz = (uint32_t)((a.H + zr - 1) / zr);
  k.super_ = a.super_ > 0 ? (uint32_t)a.super_ : kSuper;
  const uint32_t tpad = (k.tiles_x + k.super_ - 1) / k.super_ * k.super_;
  const int items = a.items > 0 ? a.items : 1;
  if ((uint64_t)tpad * tpad * k.nz * (uint64_t)items >= (1ull << 31)) return hipErrorInvalidValue;
  k.item_blocks = tpad * tpad * k.nz;
  k.n_logical = k.item_blocks * (uint32_t)items;
  k.ncols = tpad * tpad * 64u;
  k.sino_stride = a.sino_stride;
  k.out_stride = a.out_stride;
  k.outs = a.outs;
  k.loads_only = a.loads_only ? 1 : 0;
  if (a.loads_only && a.geom != GEOM_CONE) return hipErrorInvalidValue;
  if (k.n_logical == 0) return hipSuccess;
  const size_t view_bytes = (size_t)k.PV * es2;
  int chunk = (int)std::max<size_t>(1, (((size_t)1 << 31) - 1) / view_bytes);
  // the exact (f32) cone path reads its z-range starts from the chain buffer
  const bool use_chain = a.geom == GEOM_CONE && a.st_dt == 0 && k.nz > 1 && a.chain;
  if (use_chain) chunk = std::min(chunk, a.chain_views);
  int v0 = 0;
  do {  // at least one launch, so nviews == 0 still writes the zero volume
    const int cn = std::min(chunk, a.nviews - v0);
    k.sino = (const char*)a.sino + (size_t)v0 * view_bytes;
    k.sino_bytes = (uint32_t)((size_t)cn * view_bytes);
    k.sc = a.sc + v0;
    k.nviews = cn;
    k.accumulate = v0 > 0;
    k.chain = nullptr;
    if (use_chain && cn > 0) {
      const size_t total = (size_t)k.ncols * cn;
      const unsigned cg = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
      if (zr == 64)
        bp_chain_kernel<64><<<cg, 256, 0, st>>>(k, a.chain);
      else if (zr == 48)
        bp_chain_kernel<48><<<cg, 256, 0, st>>>(k, a.chain);
      else
        bp_chain_kernel<32><<<cg, 256, 0, st>>>(k, a.chain);
      k.chain = a.chain;
    }
    if (a.st_dt == 0)
      bp_dispatch_zc<float>(k, a.geom, zr, rows, a.zclamp || a.zh < HALO, st);
    else if (a.st_dt == 1)
      bp_dispatch_zc<_Float16>(k, a.geom, zr, rows, a.zclamp || a.zh < HALO, st);
    else
      bp_dispatch_zc<bf16_t>(k, a.geom, zr, rows, a.zclamp || a.zh < HALO, st);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    v0 += cn;
  } while (v0 < a.nviews);
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// OSC dual backprojection + multiplicative update (CUDAmex_oscIter.cu:272-358
// and :113-136) for one subset: num/den sums stay in registers over the
// subset's views and the voxel is updated in place, inside the circle only.
// The ratio terms arrive row-paired (the OSC forward projection writes them
// so): pair (u, v) holds rows v and v+1 of column u, so the four taps of one
// sinogram are the pairs at u and u+1, one load.
struct OSCBPKArgs {
  const void* pre_pad;   // storage S, [2, W+4, H+3, count] with a zero halo
  const void* post_pad;
  float* mu_pad;
  const float2* sc;
  int W, H, PW, first, stride, count;  // PW: pairs per row
  int64_t PV, MPX, MPXY;               // PV: pairs per view
  float v0, SR;
  uint32_t tiles_x, nz, n_logical;
};

typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));
typedef uint32_t u2a4 __attribute__((ext_vector_type(2), aligned(4)));
// {(i0,j0), (i0,j0+1), (i0+1,j0), (i0+1,j0+1)} from the pair at p and the next
template <typename S>
__device__ __forceinline__ float4 osc_ld4(const S* p);
template <>
__device__ __forceinline__ float4 osc_ld4<float>(const float* p) {
  const f4a8 v = *(const f4a8*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
template <>
__device__ __forceinline__ float4 osc_ld4<bf16_t>(const bf16_t* p) {
  const u2a4 u = *(const u2a4*)p;
  return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xffff0000u),
                     __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xffff0000u));
}

// z-voxels per thread.  A subset's launch is small (a few views), so more
// and lighter waves win: at 256^3 x 60, 10 subsets, OSC-TV took 120 ms with
// 32 (89 VGPRs, 5 waves per SIMD), 112 ms with 16 and 110 ms with 8 (the
// cone row start replays more of the running sum per voxel, but every wave
// stays resident).
constexpr int OSC_ZR = 8;

template <int GEOM, typename S>
__global__ __launch_bounds__(64) void osc_bp_update_kernel(OSCBPKArgs a) {
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  if (b >= a.n_logical) return;
  const uint32_t zr = b % a.nz;
  const uint32_t tile = b / a.nz;
  const int ix = (int)(tile % a.tiles_x) * 8 + (int)(threadIdx.x & 7);
  const int iy = (int)(tile / a.tiles_x) * 8 + (int)(threadIdx.x >> 3);
  if (ix >= a.W || iy >= a.W) return;
  const uint32_t hw = (uint32_t)a.W / 2u - 1u;
  const uint32_t dx = (uint32_t)ix - hw, dy = (uint32_t)iy - hw;
  if (!(dx * dx + dy * dy < hw * hw)) return;
  const int z0 = (int)zr * OSC_ZR;
  const int zc = min(OSC_ZR, a.H - z0);
  const float x = (float)ix - ((float)hw + 0.5f);
  const float y = (float)iy - ((float)hw + 0.5f);
  float accN[OSC_ZR], accD[OSC_ZR];
#pragma unroll
  for (int m = 0; m < OSC_ZR; ++m) {
    accN[m] = 0.f;
    accD[m] = 0.f;
  }
  for (int j = 0; j < a.count; ++j) {
    const float2 scv = a.sc[a.first + j * a.stride];
    const float sB = scv.x, cB = scv.y;
    const float s = x * cB + y * sB;
    const float t = -x * sB + y * cB;
    float U, UU, py, pz;
    if (GEOM == GEOM_PAR) {  // :318-324
      U = 1.f;
      UU = 1.f;
      py = t * UU + (float)hw + 1.f;
      pz = -a.v0 * UU + a.v0 + 0.5f;
    } else if (GEOM == GEOM_FAN) {  // :325-330
      U = 1.f;
      UU = a.SR / (a.SR - s);
      py = t * UU + (float)hw + 1.f;
      pz = 0.5f;
    } else {  // :331-336
      U = a.SR / (a.SR - s);
      UU = a.SR / (a.SR - s);
      py = t * U + (float)hw + 1.f;
      pz = -a.v0 * UU + a.v0 + 0.5f;
    }
    const float W2 = UU * UU;
    const float xb = py - 0.5f;
    const float fx = floorf(xb);
    const float ax = xb - fx;
    const int i0 = (int)fminf(fmaxf(fx, -2.f), (float)a.W);
    const size_t coff = 2 * ((size_t)j * a.PV + (size_t)(i0 + HALO));
    const S* colN = (const S*)a.pre_pad + coff;
    const S* colD = (const S*)a.post_pad + coff;
    if (GEOM == GEOM_CONE) {  // the running sum to z0, a multiple of OSC_ZR
      for (int m = 0; m < z0; m += OSC_ZR) {
#pragma unroll
        for (int q = 0; q < OSC_ZR; ++q) pz = pz + U;
      }
    } else {
      pz = pz + (float)z0;
    }
#pragma unroll
    for (int m = 0; m < OSC_ZR; ++m) {
      const float zb = pz - 0.5f;
      const float fz = floorf(zb);
      const float az = zb - fz;
      const int j0 = (int)fminf(fmaxf(fz, -2.f), (float)a.H);
      const size_t r = 2 * (size_t)(j0 + HALO) * (size_t)a.PW;
      const float4 n = osc_ld4<S>(colN + r);
      const float4 d = osc_ld4<S>(colD + r);
      const float tn = lerp(lerp(n.x, n.z, ax), lerp(n.y, n.w, ax), az);
      const float td = lerp(lerp(d.x, d.z, ax), lerp(d.y, d.w, ax), az);
      accN[m] = accN[m] + W2 * tn;
      accD[m] = accD[m] + W2 * td;
      pz = pz + U;
    }
  }
  float* mu = a.mu_pad + (size_t)(ix + HALO) + (size_t)a.MPX * (iy + HALO) +
              (size_t)a.MPXY * (z0 + HALO);
#pragma unroll
  for (int m = 0; m < OSC_ZR; ++m) {
    if (m < zc) {
      const float fval = mu[(size_t)m * a.MPXY];
      mu[(size_t)m * a.MPXY] = fmaxf(fval + (fval * accN[m] / accD[m]), 0.f);
    }
  }
}

template <typename S>
static void osc_bp_dispatch(const OSCBPKArgs& k, int geom, unsigned g, hipStream_t st) {
  switch (geom) {
    case GEOM_PAR: osc_bp_update_kernel<GEOM_PAR, S><<<g, 64, 0, st>>>(k); break;
    case GEOM_FAN: osc_bp_update_kernel<GEOM_FAN, S><<<g, 64, 0, st>>>(k); break;
    default: osc_bp_update_kernel<GEOM_CONE, S><<<g, 64, 0, st>>>(k); break;
  }
}

hipError_t launch_osc_bp_update(const OSCBPLaunch& a, hipStream_t st) {
  OSCBPKArgs k;
  k.pre_pad = a.pre_pad;
  k.post_pad = a.post_pad;
  k.mu_pad = a.mu_pad;
  k.sc = a.sc;
  k.W = a.W;
  k.H = a.H;
  k.PW = a.W + 2 * HALO;
  k.first = a.first;
  k.stride = a.stride;
  k.count = a.count;
  k.PV = (int64_t)(a.W + 2 * HALO) * (a.H + 3);
  k.MPX = a.W + 2 * HALO;
  k.MPXY = (int64_t)(a.W + 2 * HALO) * (a.W + 2 * HALO);
  k.v0 = float(a.H - 1) / 2;
  k.SR = a.sad;
  k.tiles_x = (uint32_t)((a.W + 7) / 8);
  k.nz = (uint32_t)((a.H + OSC_ZR - 1) / OSC_ZR);
  k.n_logical = k.tiles_x * k.tiles_x * k.nz;
  if (k.n_logical == 0) return hipSuccess;
  const unsigned g = padded_grid(k.n_logical);
  if (a.st_dt == 0)
    osc_bp_dispatch<float>(k, a.geom, g, st);
  else if (a.st_dt == 2)
    osc_bp_dispatch<bf16_t>(k, a.geom, g, st);
  else
    return hipErrorInvalidValue;  // ratio terms are f32 or bf16
  return hipGetLastError();
}

namespace {
__device__ __forceinline__ uint32_t abs_bits(float v, uint32_t m) {
  const float a = fabsf(v);
  return a == a ? max(m, __float_as_uint(a)) : m;  // NaN ignored
}

// Block maxima of |x| combined with one atomicMax per block; non-negative
// floats order like their bit patterns.  The body goes in 16-byte loads from
// the first 16-byte boundary on; the few elements before it and after the
// last whole vector are read one by one.
__global__ __launch_bounds__(256) void absmax_kernel(const float* x, size_t n, uint32_t* out) {
  uint32_t m = 0;
  const size_t tid = (size_t)blockIdx.x * 256 + threadIdx.x, nth = (size_t)gridDim.x * 256;
  const size_t head = min((size_t)(((16u - ((uintptr_t)x & 15u)) & 15u) / 4u), n);
  if (tid < head) m = abs_bits(x[tid], m);
  const float4* v = (const float4*)(x + head);
  const size_t nv = (n - head) / 4;
  size_t i = tid;
  for (; i + 3 * nth < nv; i += 4 * nth) {  // four loads in flight
    const float4 q0 = v[i], q1 = v[i + nth], q2 = v[i + 2 * nth], q3 = v[i + 3 * nth];
    m = abs_bits(q0.w, abs_bits(q0.z, abs_bits(q0.y, abs_bits(q0.x, m))));
    m = abs_bits(q1.w, abs_bits(q1.z, abs_bits(q1.y, abs_bits(q1.x, m))));
    m = abs_bits(q2.w, abs_bits(q2.z, abs_bits(q2.y, abs_bits(q2.x, m))));
    m = abs_bits(q3.w, abs_bits(q3.z, abs_bits(q3.y, abs_bits(q3.x, m))));
  }
  for (; i < nv; i += nth) {
    const float4 q = v[i];
    m = abs_bits(q.w, abs_bits(q.z, abs_bits(q.y, abs_bits(q.x, m))));
  }
  const size_t t0 = head + nv * 4;
  if (tid < n - t0) m = abs_bits(x[t0 + tid], m);
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
  __shared__ uint32_t wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(out, max(max(wm[0], wm[1]), max(wm[2], wm[3])));
}
}  // namespace

hipError_t launch_absmax(const float* x, size_t n, uint32_t* out_bits, hipStream_t st) {
  if (n == 0) return hipSuccess;
  // few enough workgroups that their one atomic each does not queue up
  const unsigned g = (unsigned)std::min<size_t>((n / 4 + 255) / 256 + 1, 1024);
  absmax_kernel<<<g, 256, 0, st>>>(x, n, out_bits);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// General-3D ray-driven backprojection (CUDAmex_general3D_BP.cu:61-188):
// one thread per (ray, view), 0.5-voxel steps, two corners per step with
// the reference's weights, scattered with float atomics.
struct G3DBPKArgs {
  const float* sino;
  const float* pts;
  int64_t pts_view_stride;
  const float2* sc;
  float* out;
  int W, H, M, N;
  float bx, by, bz;
  uint32_t tiles_x, tiles_per_view, n_logical;
  int n_min;       // only rays of more samples
  int xfast;       // out is the [M,M,N] x-fastest volume
  int64_t* out64;  // fixed-point sums (x-fastest) instead of out
  const uint32_t* pmax;
  const uint32_t* gate;  // non-null and 0: nothing to do
};

__device__ __forceinline__ int floor_index(float v) {
  return (int)fminf(fmaxf(floorf(v), -2.f), 1073741824.f);
}
__device__ __forceinline__ int ceil_index(float v) {
  return (int)fminf(fmaxf(ceilf(v), -2.f), 1073741824.f);
}

// One thread per (pixel, view).  A wave holds 64 consecutive detector rows
// of one column, whose rays run near each other for a detector whose
// columns are lines in space, and scatters into a z-fastest accumulator
// acc[(i*M + j)*N + k]: the lanes' atomics at one step then land on a few
// consecutive cache lines instead of one line per lane.  g3d_unz_kernel
// turns the accumulator into the [M,M,N] x-fastest output.
__global__ __launch_bounds__(256) void g3d_bp_kernel(G3DBPKArgs a) {
  if (a.gate && *a.gate == 0u) return;  // no ray of more than n_min samples
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  if (b >= a.n_logical) return;
  const uint32_t view = b / a.tiles_per_view;
  const uint32_t t = b - view * a.tiles_per_view;
  const int ix = (int)(t % a.tiles_x) * 4 + (int)(threadIdx.x >> 6);
  const int iy = (int)(t / a.tiles_x) * 64 + (int)(threadIdx.x & 63);
  if (ix >= a.W || iy >= a.H) return;
  const float2 scv = a.sc[view];
  const float sB = scv.x, cB = scv.y;
  const size_t npts = (size_t)a.W * a.H;
  const float* p = a.pts + (size_t)view * a.pts_view_stride + (size_t)iy * a.W + ix;
  R3 st, en;
  st.x = (p[0]) * cB - (p[npts]) * sB;
  st.y = (p[0]) * sB + (p[npts]) * cB;
  st.z = (p[2 * npts]);
  en.x = (p[3 * npts]) * cB - (p[4 * npts]) * sB;
  en.y = (p[3 * npts]) * sB + (p[4 * npts]) * cB;
  en.z = (p[5 * npts]);
  // slab clip, identical to the forward projector's (:84-143)
  const float TH = parallel_threshold();
  const float qnan = __builtin_nanf("");
  float aX1, aX2, aY1, aY2, aZ1, aZ2;
  if (fabsf(en.x - st.x) >= TH) {
    aX1 = (a.bx - st.x) / (en.x - st.x);
    aX2 = (-a.bx - st.x) / (en.x - st.x);
  } else {
    aX1 = qnan;
    aX2 = qnan;
  }
  if (fabsf(en.y - st.y) >= TH) {
    aY1 = (a.by - st.y) / (en.y - st.y);
    aY2 = (-a.by - st.y) / (en.y - st.y);
  } else {
    aY1 = qnan;
    aY2 = qnan;
  }
  if (fabsf(en.z - st.z) >= TH) {
    aZ1 = (a.bz - st.z) / (en.z - st.z);
    aZ2 = (-a.bz - st.z) / (en.z - st.z);
  } else {
    aZ1 = qnan;
    aZ2 = qnan;
  }
  float aXmin = fminf(aX1, aX2), aXmax = fmaxf(aX1, aX2);
  float aYmin = fminf(aY1, aY2), aYmax = fmaxf(aY1, aY2);
  float aZmin = fminf(aZ1, aZ2), aZmax = fmaxf(aZ1, aZ2);
  float tmp = fmaxf(aXmin, fmaxf(aYmin, aZmin));
  const float aMin = fmaxf(0.f, tmp);
  tmp = fminf(aXmax, fminf(aYmax, aZmax));
  const float aMax = fminf(1.f, tmp);
  R3 cur, e2;
  cur.x = st.x + aMin * (en.x - st.x);
  cur.y = st.y + aMin * (en.y - st.y);
  cur.z = st.z + aMin * (en.z - st.z);
  e2.x = st.x + aMax * (en.x - st.x);
  e2.y = st.y + aMax * (en.y - st.y);
  e2.z = st.z + aMax * (en.z - st.z);
  const float distance = sqrtf((e2.x - cur.x) * (e2.x - cur.x) + (e2.y - cur.y) * (e2.y - cur.y) +
                               (e2.z - cur.z) * (e2.z - cur.z));
  const float stepsize = 0.5f;
  const float sX = (e2.x - cur.x) / distance * stepsize;
  const float sY = (e2.y - cur.y) / distance * stepsize;
  const float sZ = (e2.z - cur.z) / distance * stepsize;
  int n = 0;
  if (distance > 0.f) {
    const float q = ceilf(distance / stepsize);
    n = q < 16777216.f ? (int)q : 16777216;
  }
  if (n <= a.n_min) return;
  const float pval = a.sino[(size_t)view * npts + (size_t)iy * a.W + ix];
  const unsigned M = (unsigned)a.M, N = (unsigned)a.N;
  const size_t hstep = a.xfast ? (size_t)a.M * a.M : 1;  // from corner kL to kL + 1
  const float scale = a.out64 ? g3d_fixed_scale(a.pmax) : 1.f;  // the bricks' fixed point
  auto emit = [&](size_t at, float lo, float hi) {
    if (a.out64) {
      atomicAdd((unsigned long long*)(a.out64 + at), g3d_fixed(lo * scale));
      if (hi != 0.f) atomicAdd((unsigned long long*)(a.out64 + at + hstep), g3d_fixed(hi * scale));
    } else {
      atomicAdd(a.out + at, lo);
      if (hi != 0.f) atomicAdd(a.out + at + hstep, hi);
    }
  };
  // Consecutive 0.5-voxel steps often land in the same cell: their weights
  // are summed in registers and scattered once (the scatter order is not
  // fixed anyway, as with the reference's atomicAdd).  pend is the index of
  // the (kL, kL+1) pair being summed, ~0 when none.
  size_t pend = ~(size_t)0;
  float wlo = 0.f, whi = 0.f;
  for (int k = 0; k < n; ++k) {
    const float fxv = cur.x - a.bx, fyv = cur.y - a.by, fzv = cur.z - a.bz;
    const int iL = floor_index(fxv), iU = ceil_index(fxv);
    const int jL = floor_index(fyv), jU = ceil_index(fyv);
    const int kL = floor_index(fzv), kU = ceil_index(fzv);
    if (!(iL < 0 || jL < 0 || kL < 0 || (unsigned)iU >= M || (unsigned)jU >= M ||
          (unsigned)kU >= N)) {
      const float iXX = fxv - (float)iL + 0.5f;
      const float iYY = fyv - (float)jL + 0.5f;
      const float iZ = fzv - (float)kL + 0.5f;
      const float iZZ = 1 - iZ;
      float lo = pval * stepsize * iXX * iYY * iZZ;
      float hi = pval * stepsize * iXX * iYY * iZ;
      if (kU == kL) {  // both corners are the same voxel
        lo = lo + hi;
        hi = 0.f;
      }
      const size_t idx = a.xfast ? ((size_t)kL * a.M + (size_t)jL) * a.M + (size_t)iL
                                 : ((size_t)iL * a.M + (size_t)jL) * N + (size_t)kL;
      if (idx == pend) {
        wlo = wlo + lo;
        whi = whi + hi;
      } else {
        if (pend != ~(size_t)0) emit(pend, wlo, whi);
        pend = idx;
        wlo = lo;
        whi = hi;
      }
    }
    cur.x = cur.x + sX;
    cur.y = cur.y + sY;
    cur.z = cur.z + sZ;
  }
  if (pend != ~(size_t)0) emit(pend, wlo, whi);
}

hipError_t launch_g3d_bp(const G3DBPLaunch& a, hipStream_t st) {
  G3DBPKArgs k;
  k.sino = a.sino;
  k.pts = a.pts;
  k.pts_view_stride = a.pts_view_stride;
  k.sc = a.sc;
  k.out = a.out;
  k.W = a.W;
  k.H = a.H;
  k.M = a.M;
  k.N = a.N;
  k.bx = (float)(-(double)(float)a.M / 2.0);
  k.by = k.bx;
  k.bz = (float)(-(double)(float)a.N / 2.0);
  k.tiles_x = (uint32_t)((a.W + 3) / 4);
  k.tiles_per_view = k.tiles_x * (uint32_t)((a.H + 63) / 64);
  k.n_logical = k.tiles_per_view * (uint32_t)a.nviews;
  k.n_min = a.n_min;
  k.xfast = a.xfast ? 1 : 0;
  k.out64 = a.out64;
  k.pmax = a.pmax;
  k.gate = a.gate;
  if (k.n_logical == 0) return hipSuccess;
  g3d_bp_kernel<<<padded_grid(k.n_logical), 256, 0, st>>>(k);
  return hipGetLastError();
}

// acc[(i*M + j)*N + k] -> out[i + M*j + M*M*k]: 32 x 32 (k, i) tiles of one
// j through LDS, read along k and written along i.
__global__ __launch_bounds__(256) void g3d_unz_kernel(const float* __restrict__ acc,
                                                      float* __restrict__ out, int M, int N) {
  __shared__ float tile[32][33];
  const int k0 = blockIdx.x * 32, i0 = blockIdx.y * 32, j = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + ty + 8 * r, k = k0 + tx;
    tile[ty + 8 * r][tx] = (i < M && k < N) ? acc[((size_t)i * M + j) * N + k] : 0.f;
  }
  __syncthreads();
  for (int r = 0; r < 4; ++r) {
    const int k = k0 + ty + 8 * r, i = i0 + tx;
    if (i < M && k < N) out[(size_t)i + (size_t)M * j + (size_t)M * M * k] = tile[tx][ty + 8 * r];
  }
}

hipError_t launch_g3d_unz(const float* acc, float* out, int M, int N, hipStream_t st) {
  const dim3 g((unsigned)((N + 31) / 32), (unsigned)((M + 31) / 32), (unsigned)M);
  g3d_unz_kernel<<<g, 256, 0, st>>>(acc, out, M, N);
  return hipGetLastError();
}

}  // namespace mexct
