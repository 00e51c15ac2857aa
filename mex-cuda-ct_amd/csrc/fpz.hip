// fpz.hip -- standard-geometry forward projector (par, fan, cone) and the
// OSC forward projection, for gfx950.
//
// Every ray of one detector column u has the same xy path: par and fan rays
// are horizontal with identical xy endpoints down a column, and cone rays of
// a column share the source and the xy direction (only the 3D step length
// changes with the row).  So a wave marches 64 consecutive detector rows of
// one column through a z-fastest volume Vz[(x+2)*PY*PZ + (y+2)*PZ + (z+2)]:
// at each step the 64 lanes read the (k, k+1) pairs of nearly the same xy
// corners, one or two cache lines per load, where an x-fastest volume costs
// about one line per lane.  A 256-thread workgroup holds 4 adjacent
// columns.  Loads are 32-bit buffer offsets.
//
// Three z-fastest layouts, each haloed by 2 on every side:
//   PLAIN: V(x,y,z); per sample four (z, z+1) pair loads at the (j+1),
//     (i+1), (i+1,j+1) strides in scalar offsets.
//   DUAL (f32): element (x,y,z) = {V(x,y,z), V(x+1,y,z)}; the two elements
//     at z, z+1 of rows j and j+1 are two 16-byte loads per sample.
//   QUAD (f16, bf16): element (x,y,z) holds the four xy corners at z; one
//     16-byte load per sample.  bf16: {V(x,y,z), V(x,y+1,z), V(x+1,y,z),
//     V(x+1,y+1,z)}.  f16 (difference quad): {V, Dy, Dx, Dxy} with
//     Dy = V(x,y+1,z) - V, Dx = V(x+1,y,z) - V and Dxy the x difference of
//     row y+1 minus Dx, each rounded once from float, so that an x lerp of
//     both rows is one packed FMA and the y lerp one mixed FMA.
// The load unit spends about the same time on a wave's load whether its
// lanes fetch 4, 8 or 16 bytes when the data is cached, and L2 supplies
// about 64 bytes per clock per CU, so the best layout issues the fewest
// loads without drawing more cache lines: DUAL halves the loads of PLAIN
// at the same lines (the 8-byte lane stride of two consecutive z elements
// spans what the four pair loads span), QUAD quarters them for 2-byte data
// but doubles the lines for f32.  The DUAL and QUAD copies take 2x and 4x
// the plain one, which MI355X's HBM holds easily; volumes whose copy would
// pass 4 GiB use PLAIN.

#include "common.h"
#include "internal.h"
#include "ray.h"

namespace mexct {

// default views per chunk of the block order: 8 for par, fan and cone (row
// groups outermost; 16 views ran 1-2% slower at 384^3 x 512 cone f32), 16
// for general-3D rays, 1 for the OSC subset projections (a subset's few
// views padded to a chunk of 16 left most blocks empty; OSC-TV 256^3 x 60:
// f32 104.6 -> 97.8 ms, f16 99.4 -> 90.8 ms with 1, 104.0 / 117.6 / 106.2 ms
// f32 with 2 / 4 / 8)
constexpr uint32_t kViewChunkLog2 = 3, kViewChunkLog2G3D = 4, kViewChunkLog2Osc = 0;
// default ray length (samples) above which 2-byte ray sums are compensated
constexpr int kLongRay = 1024;

struct FPZArgs {
  const void* vol;     // z-fastest haloed volume
  uint32_t vol_bytes;  // buffer range
  float* out;
  float* pre;          // OSC epilogue
  float* post;
  void* pre_pad;       // storage OscPad<S>::T
  void* post_pad;
  const uint32_t* vol_max;   // non-null: the volume holds V * 2^-e (pow2_exp)
  const float2* sc;
  int W, H, rows, NX, NY, NZ;
  int PY, PZ, PYPZ;
  uint32_t off_j, off_i, off_ij;  // byte offsets of the (j+1), (i+1), (i+1,j+1) pairs
  float u0, v0, bx, by, bz, SR, RD;
  int first, stride;  // view = first + j*stride
  int PW, PV;         // OSC padded copies
  const float* pts;   // general-3D: per-pixel ray endpoints [W,H,6] (per view: + view*stride)
  int64_t pts_view_stride;
  uint32_t n_ug, n_vg, n_logical;
  uint32_t nv;  // views of the launch
  int long_ray;  // 2-byte storage: waves with a ray of more samples use compensated sums
  uint32_t vchunk_log2;  // views per chunk of the block order: 1 << vchunk_log2
  uint32_t item_blocks;  // blocks of one batch item; n_logical = item_blocks * items
  // nonzero: the row groups run one after another, each in a phase of
  // phase_blocks grid blocks of which the first phase_logical work
  uint32_t phase_blocks, phase_logical;
  uint64_t vol_stride;   // bytes between the items' volumes
  uint64_t out_stride;   // floats between the items' outputs
  float* const* outs;    // non-null: item i's output at outs[i]
};

template <typename S>
__device__ __forceinline__ float2 buf_ld2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ float2 buf_ld2<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                 uint32_t soff) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  f2v v = __builtin_bit_cast(f2v, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  return make_float2(v.x, v.y);
}

// Corner pairs and weights of one sample (f32 taps).
template <typename S>
struct Tap8 {
  float2 p00, p01, p10, p11;
  float ax, ay, az;
};
// 2-byte taps stay packed, two halves per dword, until the interpolation
// (so that the loads of the samples ahead stay in flight): the f16 first lerp
// level reads them through packed or mixed-precision FMAs, bf16 halves are
// widened by a shift or a mask.
//   plain layout: w0..w3 = (z, z+1) pairs of (i,j), (i,j+1), (i+1,j), (i+1,j+1)
//   quad layout:  w0 = {(i,j), (i,j+1)} and w1 = {(i+1,j), (i+1,j+1)} at z,
//                 w2, w3 the same at z+1 (f16: w0 = {V, Dy}, w1 = {Dx, Dxy})
//   dual layout:  w0 = {(i,j), (i+1,j)} at z, w1 the same at z+1, w2, w3
//                 the same for row j+1
template <>
struct Tap8<_Float16> {
  uint32_t w0, w1, w2, w3;
  float ax, ay, az;
};
template <>
struct Tap8<bf16_t> {
  uint32_t w0, w1, w2, w3;
  float ax, ay, az;
};


// Eight taps of one quad-layout sample: element e holds the xy corners at
// z = k, element e+1 at z = k+1.
template <typename S>
__device__ __forceinline__ void quad_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, Tap8<S>& t);
template <>
__device__ __forceinline__ void quad_ld<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                               Tap8<float>& t) {
  typedef float f4v __attribute__((ext_vector_type(4)));
  const f4v q0 = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0));
  const f4v q1 = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(r, voff + 16u, 0, 0));
  t.p00 = make_float2(q0.x, q1.x);
  t.p01 = make_float2(q0.y, q1.y);
  t.p10 = make_float2(q0.z, q1.z);
  t.p11 = make_float2(q0.w, q1.w);
}
template <>
__device__ __forceinline__ void quad_ld<_Float16>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                  Tap8<_Float16>& t) {
  typedef uint32_t u4v __attribute__((ext_vector_type(4)));
  const u4v q = __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0);
  t.w0 = q.x;
  t.w1 = q.y;
  t.w2 = q.z;
  t.w3 = q.w;
}
template <>
__device__ __forceinline__ void quad_ld<bf16_t>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                Tap8<bf16_t>& t) {
  typedef uint32_t u4v __attribute__((ext_vector_type(4)));
  const u4v q = __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0);
  // q.x = {(i,j), (i,j+1)}, q.y = {(i+1,j), (i+1,j+1)} at k; q.z, q.w at k+1
  t.w0 = q.x;
  t.w1 = q.y;
  t.w2 = q.z;
  t.w3 = q.w;
}

// Four taps {(i,j,k), (i+1,j,k), (i,j,k+1), (i+1,j,k+1)} of one f32 DUAL row.
__device__ __forceinline__ float4 dual_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  typedef float f4v __attribute__((ext_vector_type(4)));
  const f4v q = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  return make_float4(q.x, q.y, q.z, q.w);
}

// 2-byte storage marches in texel coordinates (LEAN): the position is kept
// as p - b - 0.5 and stepped by the same increment, and the weight and
// corner come from v_fract / v_cvt_flr.  Sample positions then differ from
// the reference's float rounding by a few ulps, far below the storage
// rounding, and each sample costs 9 VALU less.
// CLAMP false (2-byte storage only): the caller has shown that every sample
// lies where the 2-texel halo serves it, floor(t) in [-2, N], so the clamp
// would change nothing.
typedef float f4a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef uint32_t u4a4 __attribute__((ext_vector_type(4), aligned(4)));

template <typename S, int LAYOUT, bool CLAMP = true>
__device__ __forceinline__ Tap8<S> fetch_z(__amdgpu_buffer_rsrc_t r, const FPZArgs& a,
                                           const R3& cur, float fnx, float fny, float fnz,
                                           int base, const char* vb) {
  Tap8<S> t;
  int i, j, kk;
  if constexpr (sizeof(S) == 2) {
    t.ax = __builtin_amdgcn_fractf(cur.x);
    t.ay = __builtin_amdgcn_fractf(cur.y);
    t.az = __builtin_amdgcn_fractf(cur.z);
    i = cvt_flr(cur.x);
    j = cvt_flr(cur.y);
    kk = cvt_flr(cur.z);
    if constexpr (CLAMP) {
      i = clamp_m2(i, a.NX);
      j = clamp_m2(j, a.NY);
      kk = clamp_m2(kk, a.NZ);
    }
  } else {
    const float xb = (cur.x - a.bx) - 0.5f;
    const float yb = (cur.y - a.by) - 0.5f;
    const float zb = (cur.z - a.bz) - 0.5f;
    const float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
    t.ax = xb - fx;
    t.ay = yb - fy;
    t.az = zb - fz;
    i = (int)__builtin_amdgcn_fmed3f(fx, -2.f, fnx);
    j = (int)__builtin_amdgcn_fmed3f(fy, -2.f, fny);
    kk = (int)__builtin_amdgcn_fmed3f(fz, -2.f, fnz);
  }
  // 24-bit multiplies: |i|, |j| <= N + 2 and PZ, PY * PZ < 2^23 (checked by
  // the host).  The element index is below 2^32 (checked by the host), so
  // the sums, taken modulo 2^32, are exact.
  const uint32_t e = (uint32_t)mad_i24(i, a.PYPZ, mad_i24(j, a.PZ, kk)) + (uint32_t)base;
  if constexpr (LAYOUT == LAYOUT_DUAL64) {  // f32, elements of 8 bytes (e < 2^32)
    const char* p = vb + ((uint64_t)e << 3);
    const f4a4 q0 = *(const f4a4*)p;              // row j
    const f4a4 q1 = *(const f4a4*)(p + a.off_j);  // row j+1
    t.p00 = make_float2(q0.x, q0.z);
    t.p10 = make_float2(q0.y, q0.w);
    t.p01 = make_float2(q1.x, q1.z);
    t.p11 = make_float2(q1.y, q1.w);
  } else if constexpr (LAYOUT == LAYOUT_QUAD64) {  // 2-byte storage, elements of 8 bytes
    const u4a4 q = *(const u4a4*)(vb + ((uint64_t)e << 3));
    t.w0 = q.x;
    t.w1 = q.y;
    t.w2 = q.z;
    t.w3 = q.w;
  } else if constexpr (LAYOUT == LAYOUT_QUAD) {
    quad_ld<S>(r, e * (uint32_t)(4 * sizeof(S)), t);
  } else if constexpr (LAYOUT == LAYOUT_DUAL && __is_same(S, _Float16)) {
    typedef uint32_t u2v __attribute__((ext_vector_type(2)));
    const uint32_t voff = e * 4u;
    const u2v q0 = __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0);        // row j
    const u2v q1 = __builtin_amdgcn_raw_buffer_load_b64(r, voff, a.off_j, 0);  // row j+1
    t.w0 = q0.x;
    t.w1 = q0.y;
    t.w2 = q1.x;
    t.w3 = q1.y;
  } else if constexpr (LAYOUT == LAYOUT_DUAL && __is_same(S, float)) {
    const uint32_t voff = e * 8u;
    const float4 q0 = dual_ld(r, voff, 0);         // row j
    const float4 q1 = dual_ld(r, voff, a.off_j);   // row j+1
    t.p00 = make_float2(q0.x, q0.z);
    t.p10 = make_float2(q0.y, q0.w);
    t.p01 = make_float2(q1.x, q1.z);
    t.p11 = make_float2(q1.y, q1.w);
  } else {
    const uint32_t voff = e * (uint32_t)sizeof(S);
    if constexpr (sizeof(S) == 2) {
      t.w0 = __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0);
      t.w1 = __builtin_amdgcn_raw_buffer_load_b32(r, voff, a.off_j, 0);
      t.w2 = __builtin_amdgcn_raw_buffer_load_b32(r, voff, a.off_i, 0);
      t.w3 = __builtin_amdgcn_raw_buffer_load_b32(r, voff, a.off_ij, 0);
    } else {
      t.p00 = buf_ld2<S>(r, voff, 0);         // (i,   j  ): z k, k+1
      t.p01 = buf_ld2<S>(r, voff, a.off_j);   // (i,   j+1)
      t.p10 = buf_ld2<S>(r, voff, a.off_i);   // (i+1, j  )
      t.p11 = buf_ld2<S>(r, voff, a.off_ij);  // (i+1, j+1)
    }
  }
  return t;
}

__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

// ZPK: the f16 corner-quad volume's z lerp in packed f16 too (the plain
// projection; the OSC projection keeps its z lerp in float, since its
// multiplicative update compounds the added rounding)
template <typename S, int LAYOUT_, bool ZPK = false>
__device__ __forceinline__ float interp_z(const Tap8<S>& t) {
  constexpr int LAYOUT = base_layout(LAYOUT_);
  float c00, c10, c01, c11;  // c_{y,z}: lerps along x
  if constexpr (__is_same(S, bf16_t)) {
    if constexpr (LAYOUT == LAYOUT_QUAD) {  // w0 {(i,j), (i,j+1)}, w1 {(i+1,j), (i+1,j+1)}; w2, w3 at z+1
      c00 = lerp(bf_lo(t.w0), bf_lo(t.w1), t.ax);
      c10 = lerp(bf_hi(t.w0), bf_hi(t.w1), t.ax);
      c01 = lerp(bf_lo(t.w2), bf_lo(t.w3), t.ax);
      c11 = lerp(bf_hi(t.w2), bf_hi(t.w3), t.ax);
    } else {  // w0..w3: (z, z+1) of (i,j), (i,j+1), (i+1,j), (i+1,j+1)
      c00 = lerp(bf_lo(t.w0), bf_lo(t.w2), t.ax);
      c10 = lerp(bf_lo(t.w1), bf_lo(t.w3), t.ax);
      c01 = lerp(bf_hi(t.w0), bf_hi(t.w2), t.ax);
      c11 = lerp(bf_hi(t.w1), bf_hi(t.w3), t.ax);
    }
  } else if constexpr (__is_same(S, _Float16) && LAYOUT == LAYOUT_QUAD) {
    // difference quad: w0 = {V(i,j), V(i,j+1) - V(i,j)}, w1 = {the x
    // differences of both}, at z; w2, w3 at z+1.  One packed FMA per z gives
    // {c_j, c_j+1 - c_j} (the x lerp of row j and the y difference); then
    // (ZPK) the z lerp of both in packed f16 and the y lerp as one mixed FMA,
    // or the y lerps and the z lerp in float.
    const uint32_t axz = pk_h2(t.ax, t.az);
    const uint32_t h0 = pk_fma_b<0>(axz, t.w1, t.w0);  // at z
    const uint32_t h1 = pk_fma_b<0>(axz, t.w3, t.w2);  // at z+1
    if constexpr (ZPK) {
      uint32_t d;
      asm("v_pk_add_f16 %0, %1, %2 neg_lo:[0,1] neg_hi:[0,1]" : "=v"(d) : "v"(h1), "v"(h0));
      return fma_hl(t.ay, pk_fma_b<1>(axz, d, h0));
    }
    return lerp(fma_hl(t.ay, h0), fma_hl(t.ay, h1), t.az);
  } else if constexpr (__is_same(S, _Float16)) {
    if (LAYOUT == LAYOUT_DUAL) {
      c00 = lerp_x(t.w0, t.ax);
      c01 = lerp_x(t.w1, t.ax);
      c10 = lerp_x(t.w2, t.ax);
      c11 = lerp_x(t.w3, t.ax);
    } else {
      c00 = lerp_h<0>(t.w0, t.w2, t.ax);
      c10 = lerp_h<0>(t.w1, t.w3, t.ax);
      c01 = lerp_h<1>(t.w0, t.w2, t.ax);
      c11 = lerp_h<1>(t.w1, t.w3, t.ax);
    }
  } else {
    c00 = lerp(t.p00.x, t.p10.x, t.ax);
    c10 = lerp(t.p01.x, t.p11.x, t.ax);
    c01 = lerp(t.p00.y, t.p10.y, t.ax);
    c11 = lerp(t.p01.y, t.p11.y, t.ax);
  }
  const float c0 = lerp(c00, c10, t.ay);
  const float c1 = lerp(c01, c11, t.ay);
  return lerp(c0, c1, t.az);
}

// One ray through the z-fastest volume, software-pipelined: the loads of
// the samples ahead (k+1 in march_loop, k+1 .. k+D in march_loopd) are
// issued before sample k is interpolated.  Fetches past the last sample read
// clamped or range-checked offsets and are discarded.
// Whether every sample cur + k inc (k < n, texel coordinates, running sum)
// has floor in [-2, N] on each axis, the range the 2-texel halo serves.  The
// samples lie on the segment between the first and the last, each off by
// at most the running sum's rounding, bounded here by n half-ulps of the
// largest coordinate; the ends must then lie in [-1.5, N + 0.5].
__device__ __forceinline__ bool in_halo(const R3& c, const R3& inc, int n, float fnx, float fny,
                                        float fnz) {
  const float k = (float)(n - 1);
  const float ex = __builtin_fmaf(k, inc.x, c.x), ey = __builtin_fmaf(k, inc.y, c.y),
              ez = __builtin_fmaf(k, inc.z, c.z);
  const float big = fmaxf(fmaxf(fmaxf(fabsf(c.x), fabsf(ex)), fmaxf(fabsf(c.y), fabsf(ey))),
                          fmaxf(fabsf(c.z), fabsf(ez))) + 1.f;
  const float d = (float)n * big * 1.2e-7f + 0.01f;  // drift bound, 2^-23 per step
  return fminf(c.x, ex) - d >= -1.5f && fmaxf(c.x, ex) + d <= fnx + 0.5f &&
         fminf(c.y, ey) - d >= -1.5f && fmaxf(c.y, ey) + d <= fny + 0.5f &&
         fminf(c.z, ez) - d >= -1.5f && fmaxf(c.z, ez) + d <= fnz + 0.5f;
}

// The sample loop of the compensated waves (COMP: 2-byte storage, long
// rays): each pair of samples is added to the ray sum with Kahan's
// compensated summation, so the sum's rounding error stays near one ulp
// however many samples the ray has instead of growing with their count.
// Without COMP the samples are added one by one in ray order, as the
// reference does (march_loopd is that loop with more samples ahead).
template <typename S, int STEP2, int LAYOUT, bool COMP, bool CLAMP, bool ZPK = false>
__device__ __forceinline__ float march_loop(__amdgpu_buffer_rsrc_t r, const FPZArgs& a, R3 cur,
                                            const R3& inc, int n, float fnx, float fny, float fnz,
                                            int base, const char* vb) {
  constexpr float step = STEP2 ? 0.5f : 1.0f;
  float dval = 0.f, comp = 0.f;
  // two register sets in ping-pong: while one sample is interpolated the
  // next sample's loads are in flight
  Tap8<S> t0 = fetch_z<S, LAYOUT, CLAMP>(r, a, cur, fnx, fny, fnz, base, vb);
  int k = 0;
  for (; k + 1 < n; k += 2) {
    cur.x = cur.x + inc.x;
    cur.y = cur.y + inc.y;
    cur.z = cur.z + inc.z;
    const Tap8<S> t1 = fetch_z<S, LAYOUT, CLAMP>(r, a, cur, fnx, fny, fnz, base, vb);
    __builtin_amdgcn_sched_barrier(0);  // t1's loads stay ahead of t0's interpolation
    float x = interp_z<S, LAYOUT, ZPK>(t0) * step;
    if constexpr (!COMP) dval = dval + x;
    cur.x = cur.x + inc.x;
    cur.y = cur.y + inc.y;
    cur.z = cur.z + inc.z;
    t0 = fetch_z<S, LAYOUT, CLAMP>(r, a, cur, fnx, fny, fnz, base, vb);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (COMP) {
      x = x + interp_z<S, LAYOUT, ZPK>(t1) * step;
      const float y = x - comp;
      const float t = dval + y;
      comp = (t - dval) - y;
      dval = t;
    } else {
      dval = dval + interp_z<S, LAYOUT, ZPK>(t1) * step;
    }
  }
  if (k < n) {
    const float x = interp_z<S, LAYOUT, ZPK>(t0) * step;
    if constexpr (COMP) {
      const float y = x - comp;
      dval = dval + y;
    } else {
      dval = dval + x;
    }
  }
  return dval;
}

// The bits of every tap of a sample folded by XOR: the loads-only form of
// the projector keeps each load (all of its dwords are used) and drops the
// interpolation.
__device__ __forceinline__ uint32_t tap_bits(const Tap8<float>& t) {
  return __float_as_uint(t.p00.x) ^ __float_as_uint(t.p00.y) ^ __float_as_uint(t.p01.x) ^
         __float_as_uint(t.p01.y) ^ __float_as_uint(t.p10.x) ^ __float_as_uint(t.p10.y) ^
         __float_as_uint(t.p11.x) ^ __float_as_uint(t.p11.y);
}
template <typename S>
__device__ __forceinline__ uint32_t tap_bits(const Tap8<S>& t) {
  return t.w0 ^ t.w1 ^ t.w2 ^ t.w3;
}

// The plain sample loop with the loads of the next D samples in flight
// while one is interpolated (D + 1 register sets in rotation): the same
// samples added in the same order as march_loop's.  Fetches up to D samples
// past the end read range-checked buffer offsets (or, for the plain layout,
// clamped indices) and are discarded.  LOADS: the loads-only form (the
// projector's gather ceiling): every address and load of the real loop, the
// taps XOR-folded instead of interpolated; returns the fold's bits.
template <typename S, int STEP2, int LAYOUT, bool CLAMP, int D, bool LOADS = false, bool ZPK = false>
__device__ __forceinline__ float march_loopd(__amdgpu_buffer_rsrc_t r, const FPZArgs& a, R3 cur,
                                             const R3& inc, int n, float fnx, float fny,
                                             float fnz, int base, const char* vb) {
  constexpr float step = STEP2 ? 0.5f : 1.0f;
  float dval = 0.f;
  uint32_t fold = 0;
  Tap8<S> t[D + 1];  // t[u] holds sample k + u at the top of each trip
  t[0] = fetch_z<S, LAYOUT, CLAMP>(r, a, cur, fnx, fny, fnz, base, vb);
#pragma unroll
  for (int d = 1; d < D; ++d) {
    cur.x = cur.x + inc.x;
    cur.y = cur.y + inc.y;
    cur.z = cur.z + inc.z;
    t[d] = fetch_z<S, LAYOUT, CLAMP>(r, a, cur, fnx, fny, fnz, base, vb);
  }
  int k = 0;
  for (; k + D < n; k += D + 1) {
#pragma unroll
    for (int u = 0; u <= D; ++u) {
      cur.x = cur.x + inc.x;
      cur.y = cur.y + inc.y;
      cur.z = cur.z + inc.z;
      t[(u + D) % (D + 1)] = fetch_z<S, LAYOUT, CLAMP>(r, a, cur, fnx, fny, fnz, base, vb);
      __builtin_amdgcn_sched_barrier(0);  // sample k+u+D's loads stay ahead
      if constexpr (LOADS)
        fold ^= tap_bits(t[u]);
      else
        dval = dval + interp_z<S, LAYOUT, ZPK>(t[u]) * step;
    }
  }
#pragma unroll
  for (int u = 0; u < D; ++u)
    if (k + u < n) {
      if constexpr (LOADS)
        fold ^= tap_bits(t[u]);
      else
        dval = dval + interp_z<S, LAYOUT, ZPK>(t[u]) * step;
    }
  return LOADS ? __uint_as_float(fold) : dval;
}

// Samples in flight in the plain loop.  At 384^3 x 512 cone, f32: 32.3 ms
// with 1, 31.5 with 2, 30.9 with 3, 29.9 with 5, 29.5 with 7 and 11 (119 and
// 179 VGPRs); f16 19.8 ms with 1, 19.1-19.3 with 3 and 5; with the f16
// projections' workgroups per CU capped (fpz_dispatch_l), the cone f16 FP
// ran 15.3 ms with 3, 5 and 7 and the helical general-3D one 13.4, 12.9 and
// 12.4 ms, so 2-byte storage keeps 7 (bf16 with 7 and the caps: cone 20.0 ->
// 19.4 ms, general-3D 15.2 -> 14.0 ms, against 3 uncapped).  An OSC subset's projection
// (6 views at 256^2 x 256: 6 waves per SIMD) keeps 4, the most that lets all
// 6 stay resident (80 VGPRs f32): 307 us per subset with 2, 279 with 3, ~240
// with 4.
template <typename S, int EPI>
constexpr int fetch_ahead() {
  return EPI == 1 ? 4 : 7;
}

// EPI: 0 store the ray sum, 1 the OSC ratio terms, 2 the loads-only form.
constexpr bool loads_only(int epi) { return epi == 2; }

template <typename S, int STEP2, int LAYOUT, int EPI>
__device__ __forceinline__ float march_z(__amdgpu_buffer_rsrc_t r, const FPZArgs& a, R3 cur,
                                         const R3& inc, int n, const char* vb) {
  const float fnx = (float)a.NX, fny = (float)a.NY, fnz = (float)a.NZ;
  const int base = 2 * a.PYPZ + 2 * a.PZ + 2;
  if (n <= 0) return 0.f;
  if constexpr (sizeof(S) == 2) {  // texel coordinates (see fetch_z)
    cur.x = (cur.x - a.bx) - 0.5f;
    cur.y = (cur.y - a.by) - 0.5f;
    cur.z = (cur.z - a.bz) - 0.5f;
    // one loop per wave: compensated when any of its rays is long; without
    // the index clamps when every ray stays inside the halo
    const bool comp = __builtin_amdgcn_ballot_w64(n > a.long_ray) != 0;
    // the plain layout reads its neighbours through scalar offsets, which the
    // buffer range check does not cover, and the 64-bit layouts have no range
    // check at all (loads run past the ray's end): their indices are always
    // clamped
    const bool clamp = LAYOUT == LAYOUT_PLAIN || wide64(LAYOUT) ||
                       __builtin_amdgcn_ballot_w64(!in_halo(cur, inc, n, fnx, fny, fnz)) != 0;
    if (comp && !loads_only(EPI))
      return clamp ? march_loop<S, STEP2, LAYOUT, true, true, EPI == 0>(r, a, cur, inc, n, fnx, fny, fnz, base, vb)
                   : march_loop<S, STEP2, LAYOUT, true, false, EPI == 0>(r, a, cur, inc, n, fnx, fny, fnz, base, vb);
    return clamp ? march_loopd<S, STEP2, LAYOUT, true, fetch_ahead<S, EPI>(), loads_only(EPI), EPI == 0>(
                       r, a, cur, inc, n, fnx, fny, fnz, base, vb)
                 : march_loopd<S, STEP2, LAYOUT, false, fetch_ahead<S, EPI>(), loads_only(EPI), EPI == 0>(
                       r, a, cur, inc, n, fnx, fny, fnz, base, vb);
  }
  return march_loopd<S, STEP2, LAYOUT, true, fetch_ahead<S, EPI>(), loads_only(EPI)>(
      r, a, cur, inc, n, fnx, fny, fnz, base, vb);
}

// Storage of the OSC ratio terms for a volume storage S.  They span the
// counts times the line integral (up to ~1e5 and down to ~1e-5 and below),
// beyond FP16's range, so an FP16 volume copy pairs with BF16 ratio terms.
template <typename S>
struct OscPad {
  typedef S T;
};
template <>
struct OscPad<_Float16> {
  typedef bf16_t T;
};

// Value v of detector pixel (u, row) into a row-paired copy: pair q (that
// of (u, row)) holds it first, the pair of (u, row - 1) second.
template <typename P>
__device__ __forceinline__ void osc_pair_store(void* pad, size_t q, int PW, float v) {
  P* p = (P*)pad;
  const P s = Store<P>::from(v);
  p[2 * q] = s;
  p[2 * (q - (size_t)PW) + 1] = s;
}

// EPI 0: store the ray sum.  EPI 1: OSC ratio terms of CUDAmex_oscIter.cu:262-268
// written in place, plus zero-haloed row-paired copies ([2, W+4, H+3, views])
// for the dual backprojection.  EPI 2: the loads-only form, storing the
// fold of each ray's taps (a measurement, not a projection).
template <int GEOM, typename S, int EPI, int LAYOUT>
__global__ __launch_bounds__(256) void fpz_kernel(FPZArgs a) {
  // Block order: (batch item, chunk of 2^vchunk_log2 views, column group,
  // view in chunk), each XCD taking a contiguous range.  The blocks an XCD
  // runs together then march the same 4 detector columns of consecutive
  // views, whose rays lie within a few voxels of each other, so the volume
  // lines they gather are shared in that XCD's L2 (view-major order:
  // 47.8 ms -> 32.4 ms at 384^3 x 512 cone, f32).  The row group (64
  // detector rows) is either innermost, or, for the par, fan and cone
  // projections (phase_blocks), outermost in dispatch order: the whole chip then
  // projects one row group, whose rays cross a slab of the volume about 64
  // voxels thick whose copy stays in the memory-side cache across views
  // (384^3 x 512 cone: f32 29.6 -> 26.0 ms, f16 18.9 -> 17.4 ms).
  uint32_t item, vg, b;
  if (a.phase_blocks) {
    vg = blockIdx.x / a.phase_blocks;
    const uint32_t bl = xcd_remap(blockIdx.x - vg * a.phase_blocks, a.phase_blocks);
    if (bl >= a.phase_logical) return;
    const uint32_t per = a.item_blocks / a.n_vg;
    item = bl / per;
    b = bl - item * per;
  } else {
    const uint32_t b0 = xcd_remap(blockIdx.x, gridDim.x);
    if (b0 >= a.n_logical) return;
    item = b0 / a.item_blocks;
    const uint32_t bi = b0 - item * a.item_blocks;
    vg = bi % a.n_vg;
    b = bi / a.n_vg;
  }
  uint32_t rest = b;
  const uint32_t vw = rest & ((1u << a.vchunk_log2) - 1u);
  rest >>= a.vchunk_log2;
  const uint32_t ug = rest % a.n_ug;
  const uint32_t jv = ((rest / a.n_ug) << a.vchunk_log2) + vw;
  if (jv >= a.nv) return;
  const int ix = (int)ug * 4 + (int)(threadIdx.x >> 6);
  const int iy = (int)vg * 64 + (int)(threadIdx.x & 63);
  if (ix >= a.W || iy >= a.H) return;
  const int view = a.first + (int)jv * a.stride;
  const size_t p = (size_t)view * a.W * a.H + (size_t)iy * a.W + ix;
  float* const out = a.outs ? a.outs[item] : a.out + item * a.out_stride;
  float preval = 0.f, postval = 0.f;
  size_t q = 0;
  typedef typename OscPad<S>::T P;
  if (EPI != 1) {
    if (iy >= a.rows) {  // the reference's grid never writes these rows
      out[p] = 0.f;
      return;
    }
  } else {
    preval = a.pre[p];
    postval = a.post[p];
    q = (size_t)jv * a.PV + (size_t)(iy + HALO) * a.PW + (ix + HALO);
    if (iy >= a.rows) {  // unprocessed rows keep their values
      osc_pair_store<P>(a.pre_pad, q, a.PW, preval);
      osc_pair_store<P>(a.post_pad, q, a.PW, postval);
      return;
    }
  }
  const float2 scv = a.sc[view];
  R3 st, en;
  if (GEOM == GEOM_G3D) {
    g3d_ray(a.pts + (size_t)view * a.pts_view_stride + (size_t)iy * a.W + ix, (size_t)a.W * a.H,
            scv.x, scv.y, st, en);
  } else {
    std_ray<GEOM>(ix, iy, a.u0, a.v0, scv.x, scv.y, a.SR, a.RD, st, en);
  }
  // CUDAmex_FP marches cone rays with step 1.0 (CUDAmex_FP.cu:231), par and
  // fan with 0.5; CUDAmex_oscIter uses 0.5 for every geometry (:244).
  constexpr int STEP2 = (EPI == 1 || GEOM != GEOM_CONE) ? 1 : 0;
  if constexpr (loads_only(EPI)) {
    R3 cur, inc;
    const int n = clip_ray(st, en, a.bx, a.by, a.bz, STEP2 ? 0.5f : 1.0f, cur, inc);
    const char* vb = (const char*)a.vol + item * a.vol_stride;
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc((void*)vb, 0, (int)a.vol_bytes, 0x00020000);
    out[p] = march_z<S, STEP2, LAYOUT, EPI>(r, a, cur, inc, n, vb);
    return;
  }
  R3 cur, inc;
  const int n = clip_ray(st, en, a.bx, a.by, a.bz, STEP2 ? 0.5f : 1.0f, cur, inc);
  const char* vb = (const char*)a.vol + item * a.vol_stride;
  // below 4 GiB (the 64-bit layouts read through vb instead)
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc((void*)vb, 0, (int)a.vol_bytes, 0x00020000);
  const float dval = ldexpf(march_z<S, STEP2, LAYOUT, EPI>(r, a, cur, inc, n, vb),
                            pow2_exp(a.vol_max ? a.vol_max + item : nullptr));
  if (EPI != 1) {
    out[p] = dval;
  } else {
    const float np = (preval * expf(-dval) - postval);
    const float npo = preval * expf(-dval) * dval;
    a.pre[p] = np;
    a.post[p] = npo;
    osc_pair_store<P>(a.pre_pad, q, a.PW, np);
    osc_pair_store<P>(a.post_pad, q, a.PW, npo);
  }
}

// The block order's view chunk and the number of batch items.
static void set_view_chunk(FPZArgs& k, uint32_t log2, uint32_t items) {
  k.vchunk_log2 = log2;
  const uint32_t c = 1u << log2;
  const uint32_t nvp = (k.nv + c - 1) / c * c;
  k.item_blocks = k.n_ug * k.n_vg * nvp;
  k.n_logical = k.item_blocks * items;
}

// W x H detector, NX x NX x NZ volume (the same sizes in the standard
// geometries; general-3D rays have their own detector).
static FPZArgs make_zargs(const void* vol, int st_dt, int layout_, const float2* sc, int W, int H,
                          int nviews, float sad, int NX = 0, int NZ = 0) {
  const int layout = base_layout(layout_);
  if (NX == 0) NX = W;
  if (NZ == 0) NZ = H;
  FPZArgs k{};
  k.vol = vol;
  k.sc = sc;
  k.W = W;
  k.H = H;
  const int nb = (W + 31) / 32;  // CUDAmex_FP.cu:434-440 grid extent
  k.rows = H < 32 * nb ? H : 32 * nb;
  k.NX = NX;
  k.NY = NX;
  k.NZ = NZ;
  // element counts: x and y in [-2, N+1] (N+4), except the element of the
  // corner pair or quad, whose base runs over [-2, N] (N+3)
  const int PX = layout == LAYOUT_PLAIN ? NX + 2 * HALO : NX + 2 * HALO - 1;
  k.PY = layout == LAYOUT_QUAD ? NX + 2 * HALO - 1 : NX + 2 * HALO;
  k.PZ = NZ + 2 * HALO;
  k.PYPZ = k.PY * k.PZ;
  const uint32_t es = (st_dt == 0 ? 4u : 2u) *
                      (layout == LAYOUT_QUAD ? 4u : layout == LAYOUT_DUAL ? 2u : 1u);
  k.vol_bytes = wide64(layout_) ? 0u : (uint32_t)((size_t)PX * k.PY * k.PZ * es);
  k.off_j = (uint32_t)k.PZ * es;
  k.off_i = (uint32_t)k.PY * (uint32_t)k.PZ * es;
  k.off_ij = k.off_i + k.off_j;
  k.u0 = float(W - 1) / 2;
  k.v0 = float(H - 1) / 2;
  k.bx = (float)(-(double)(float)NX / 2.0);
  k.by = k.bx;
  k.bz = (float)(-(double)(float)NZ / 2.0);
  k.SR = sad;
  k.RD = sad;
  k.first = 0;
  k.stride = 1;
  k.n_ug = (uint32_t)((W + 3) / 4);
  k.n_vg = (uint32_t)((H + 63) / 64);
  k.nv = (uint32_t)nviews;
  k.long_ray = kLongRay;
  set_view_chunk(k, kViewChunkLog2, 1);
  return k;
}

// The 2-byte par, fan and cone projections (and their loads-only form) run
// at most 5 workgroups (20 waves) per CU: each reserves 32 KB of LDS, which
// the kernel does not use.  Their VGPRs would otherwise let 8 run; at 384^3
// x 512 cone the median of ten f16 launches went 15.4 -> 15.0 ms (6 per CU:
// 15.2 ms, 4: 15.0 ms).  f32 (119 VGPRs) runs 4 per CU anyway.
constexpr unsigned kF16OccLds = 32768;
// The general-3D projection likewise, with fewer workgroups still: 41000 B
// of LDS per workgroup for 2-byte storage, 54000 B for f32, the fastest of 0, 23000,
// 32000, 41000, 54000 and 81000 B at the helical 320^3 x 360 views x 256^2
// workload (f16 15.4 -> 14.0 ms, f32 23.6 -> 22.8 ms).  The cone f32 FP ran
// the same with any of them (26.6 ms).
constexpr unsigned kG3DOccLdsF16 = 41000, kG3DOccLdsF32 = 54000;

template <int EPI, typename S, int LAYOUT>
static void fpz_dispatch_l(const FPZArgs& k0, int geom, hipStream_t st) {
  FPZArgs k = k0;
  unsigned g = (k.n_logical + 7u) & ~7u;
  const unsigned lds = EPI != 1 && sizeof(S) == 2 ? kF16OccLds : 0u;
  // row groups outermost (see fpz_kernel), except for general-3D rays and
  // the OSC subset projections: a few views, all of whose blocks are
  // resident at once (at 256^3 x 60 views, 10 subsets, the OSC-TV run took
  // 119 ms with the phases against 103 ms without)
  if (geom != GEOM_G3D && EPI != 1) {
    k.phase_logical = k.item_blocks / k.n_vg * (k.n_logical / k.item_blocks);
    k.phase_blocks = (k.phase_logical + 7u) & ~7u;
    g = k.phase_blocks * k.n_vg;
  }
  switch (geom) {
    case GEOM_PAR: fpz_kernel<GEOM_PAR, S, EPI, LAYOUT><<<g, 256, lds, st>>>(k); break;
    case GEOM_FAN: fpz_kernel<GEOM_FAN, S, EPI, LAYOUT><<<g, 256, lds, st>>>(k); break;
    case GEOM_CONE: fpz_kernel<GEOM_CONE, S, EPI, LAYOUT><<<g, 256, lds, st>>>(k); break;
    default:
      if constexpr (EPI == 0) {
        const unsigned l3 = sizeof(S) == 2 ? kG3DOccLdsF16 : kG3DOccLdsF32;
        fpz_kernel<GEOM_G3D, S, 0, LAYOUT><<<g, 256, l3, st>>>(k);
      }
      break;
  }
}

// Layouts built: PLAIN for every storage, DUAL and DUAL64 for f32, QUAD and
// QUAD64 for 2-byte storage (the loads-only form, EPI 2, below).
template <int EPI, typename S>
static void fpz_dispatch(const FPZArgs& k, int geom, int layout, hipStream_t st) {
  if constexpr (sizeof(S) == 2) {
    if (layout == LAYOUT_QUAD) return fpz_dispatch_l<EPI, S, LAYOUT_QUAD>(k, geom, st);
    if (layout == LAYOUT_QUAD64) return fpz_dispatch_l<EPI, S, LAYOUT_QUAD64>(k, geom, st);
  }
  if constexpr (!__is_same(S, bf16_t)) {
    if (layout == LAYOUT_DUAL) return fpz_dispatch_l<EPI, S, LAYOUT_DUAL>(k, geom, st);
  }
  if constexpr (__is_same(S, float))
    if (layout == LAYOUT_DUAL64) return fpz_dispatch_l<EPI, S, LAYOUT_DUAL64>(k, geom, st);
  fpz_dispatch_l<EPI, S, LAYOUT_PLAIN>(k, geom, st);
}

// The loads-only form exists for the layouts the projector picks below
// 4 GiB: DUAL for f32, QUAD for 2-byte storage.
template <typename S>
static hipError_t fpz_loads_dispatch(const FPZArgs& k, int geom, int layout, hipStream_t st) {
  constexpr int L = sizeof(S) == 2 ? LAYOUT_QUAD : LAYOUT_DUAL;
  if (layout != L || geom == GEOM_G3D) return hipErrorInvalidValue;
  fpz_dispatch_l<2, S, L>(k, geom, st);
  return hipSuccess;
}

hipError_t launch_fpz(const FPZLaunch& a, hipStream_t st) {
  FPZArgs k = make_zargs(a.vol, a.st_dt, a.layout, a.sc, a.W, a.H, a.nviews, a.sad, a.NX, a.NZ);
  k.out = a.out;
  k.vol_max = a.vol_max;
  k.pts = a.pts;
  k.pts_view_stride = a.pts_view_stride;
  if (a.long_ray != 0) k.long_ray = a.long_ray;
  if (a.geom == GEOM_G3D) k.rows = a.H;  // every detector row is a ray
  const int items = a.items > 0 ? a.items : 1;
  set_view_chunk(k,
                 a.vchunk_log2 >= 0          ? (uint32_t)a.vchunk_log2
                 : a.geom == GEOM_G3D ? kViewChunkLog2G3D
                                      : kViewChunkLog2,
                 (uint32_t)items);
  k.vol_stride = a.vol_stride;
  k.out_stride = a.out_stride;
  k.outs = a.outs;
  if ((uint64_t)k.item_blocks * (uint64_t)items >= (1ull << 31)) return hipErrorInvalidValue;
  if (k.n_logical == 0) return hipSuccess;
  if (a.loads_only) {
    const hipError_t e = a.st_dt == 0   ? fpz_loads_dispatch<float>(k, a.geom, a.layout, st)
                         : a.st_dt == 1 ? fpz_loads_dispatch<_Float16>(k, a.geom, a.layout, st)
                                        : fpz_loads_dispatch<bf16_t>(k, a.geom, a.layout, st);
    return e != hipSuccess ? e : hipGetLastError();
  }
  if (a.st_dt == 0)
    fpz_dispatch<0, float>(k, a.geom, a.layout, st);
  else if (a.st_dt == 1)
    fpz_dispatch<0, _Float16>(k, a.geom, a.layout, st);
  else
    fpz_dispatch<0, bf16_t>(k, a.geom, a.layout, st);
  return hipGetLastError();
}

hipError_t launch_osc_fp(const OSCFPLaunch& a, hipStream_t st) {
  FPZArgs k = make_zargs(a.mu_z, a.st_dt, a.layout, a.sc, a.W, a.H, a.count, a.sad);
  set_view_chunk(k, kViewChunkLog2Osc, 1);
  k.pre = a.pre;
  k.post = a.post;
  k.pre_pad = a.pre_pad;
  k.post_pad = a.post_pad;
  k.vol_max = a.vol_max;
  k.first = a.first;
  k.stride = a.stride;
  k.PW = a.W + 2 * HALO;        // pairs per row
  k.PV = (a.W + 2 * HALO) * (a.H + 3);  // pairs per view: rows -2 .. H
  if (a.long_ray != 0) k.long_ray = a.long_ray;
  if (k.n_logical == 0) return hipSuccess;
  if (a.st_dt == 0)
    fpz_dispatch<1, float>(k, a.geom, a.layout, st);
  else if (a.st_dt == 1)
    fpz_dispatch<1, _Float16>(k, a.geom, a.layout, st);
  else
    fpz_dispatch<1, bf16_t>(k, a.geom, a.layout, st);
  return hipGetLastError();
}

}  // namespace mexct
