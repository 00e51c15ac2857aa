// common.h -- device helpers shared by the gfx950 kernels.
//
// Every kernel is compiled with -ffp-contract=off: a float expression is
// rounded once per operation in source order, so the sample positions,
// weights and sums reproduce the reference kernels' single-precision
// arithmetic operation by operation.  FMAs appear only where written
// explicitly (the texture lerp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mexct {

enum Geom { GEOM_PAR = 0, GEOM_FAN = 1, GEOM_CONE = 2, GEOM_G3D = 3 };
enum Mode { ADDR_CLAMP = 0, ADDR_BORDER = 1 };

// Every gathered operand (volume for FP, sinogram for BP) is stored with a
// 2-texel halo on each side of each axis: edge copies for clamp
// addressing, zeros for border addressing.  A corner index clamped to
// [-2, N] then always reads the texels the reference's address mode
// selects, and the x-neighbour pair (i, i+1) is one 8-byte load.
constexpr int HALO = 2;

// storage tags
struct bf16_t {
  uint16_t v;
};
typedef float f2a4 __attribute__((ext_vector_type(2), aligned(4)));
typedef _Float16 h2a2 __attribute__((ext_vector_type(2), aligned(2)));
typedef uint16_t u2a2 __attribute__((ext_vector_type(2), aligned(2)));

template <typename S>
struct Store;

template <>
struct Store<float> {
  static constexpr int bytes = 4;
  __device__ __forceinline__ static float2 ld2(const float* p) {
    f2a4 v = *(const f2a4*)p;
    return make_float2(v.x, v.y);
  }
  __device__ __forceinline__ static float ld(const float* p) { return *p; }
  __device__ __forceinline__ static float from(float f) { return f; }
};

template <>
struct Store<_Float16> {
  static constexpr int bytes = 2;
  __device__ __forceinline__ static float2 ld2(const _Float16* p) {
    h2a2 v = *(const h2a2*)p;
    return make_float2((float)v.x, (float)v.y);
  }
  __device__ __forceinline__ static float ld(const _Float16* p) { return (float)*p; }
  __device__ __forceinline__ static _Float16 from(float f) { return (_Float16)f; }
};

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

template <>
struct Store<bf16_t> {
  static constexpr int bytes = 2;
  __device__ __forceinline__ static float2 ld2(const bf16_t* p) {
    u2a2 v = *(const u2a2*)p;
    return make_float2(bf16_to_f32(v.x), bf16_to_f32(v.y));
  }
  __device__ __forceinline__ static float ld(const bf16_t* p) { return bf16_to_f32(p->v); }
  __device__ __forceinline__ static bf16_t from(float f) {
    bf16_t b;
    b.v = f32_to_bf16(f);
    return b;
  }
};

// Load element i of an input array of dtype 0 f32, 1 f16, 2 bf16.
__device__ __forceinline__ float load_any(const void* p, int dtype, size_t i) {
  if (dtype == 0) return ((const float*)p)[i];
  if (dtype == 1) return (float)((const _Float16*)p)[i];
  return bf16_to_f32(((const uint16_t*)p)[i]);
}

// lerp(lo, hi, t) of the two f16 halves of one dword A, in float (the same
// identity as lerp_h).
__device__ __forceinline__ float lerp_x(uint32_t A, float t) {
  float d, c;
  asm("v_fma_mix_f32 %0, %1, 1.0, -%1 op_sel:[1,0,0] op_sel_hi:[1,0,1]" : "=v"(d) : "v"(A));
  asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,0,1]" : "=v"(c) : "v"(t), "v"(d), "v"(A));
  return c;
}

// floor(x) as an int in one instruction
__device__ __forceinline__ int cvt_flr(float x) {
  int i;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(i) : "v"(x));
  return i;
}
// a * b + c with a, b signed 24-bit and c a full 32-bit addend, one
// full-rate instruction (the compiler otherwise forms mul24 + add into a
// 64-bit v_mad_u64_u32, a quarter-rate one); b is wave-uniform
__device__ __forceinline__ int mad_i24(int a, int b, int c) {
  int r;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
  return r;
}
// median(i, -2, hi): the corner index clamped to [-2, hi]
__device__ __forceinline__ int clamp_m2(int i, int hi) {
  int r;
  asm("v_med3_i32 %0, %1, -2, %2" : "=v"(r) : "v"(i), "s"(hi));
  return r;
}

// FP16 copies of f32 data are stored as v * 2^-e, e the binary exponent of
// max|v| (max = m * 2^e, m in [0.5, 1)), so that they use FP16's range
// (subnormals below 6.1e-5, overflow above 65504) however the data is
// scaled.  A power of two commutes with float rounding, so results computed
// from the scaled copy and multiplied by 2^e equal those from an unscaled
// copy wherever that one neither overflows nor underflows.
__device__ __forceinline__ int pow2_exp(const uint32_t* maxbits) {
  int e = 0;
  if (maxbits) {
    const float m = __uint_as_float(*maxbits);
    if (m > 0.f && m < 3.0e38f) (void)frexpf(m, &e);
  }
  return e;
}

// texture lerp (1-t)*a + t*b evaluated as fma(t, b - a, a)
__device__ __forceinline__ float lerp(float a, float b, float t) { return __builtin_fmaf(t, b - a, a); }

// lerp(a, b, t) of the f16 halves HI of A (a) and B (b), in float:
// fma(b, 1, -a) is exactly the difference of the widened halves and
// fma(t, d, a) the lerp, so the result equals lerp((float)a, (float)b, t).
template <int HI>
__device__ __forceinline__ float lerp_h(uint32_t A, uint32_t B, float t) {
  float d, c;
  if (HI) {
    asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel:[1,0,1] op_sel_hi:[1,0,1]" : "=v"(d) : "v"(B), "v"(A));
    asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "=v"(c) : "v"(t), "v"(d), "v"(A));
  } else {
    asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[1,0,1]" : "=v"(d) : "v"(B), "v"(A));
    asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,0,1]" : "=v"(c) : "v"(t), "v"(d), "v"(A));
  }
  return c;
}

// The first lerp level of 2-byte f16 taps in packed f16 arithmetic: both
// halves of A and B at once, {lerp(A.lo, B.lo, t), lerp(A.hi, B.hi, t)} as
// fma(t, B - A, A) with one f16 rounding per step.  t2 holds t rounded to
// f16 in both halves (pk_h2).  Two instructions where two lerp_h and a
// float lerp take six, at a rounding of the same size as the f16 storage's
// own (2^-11 relative); a hardware texture unit's 8-bit weights round the
// lerp far more coarsely.  The f16 backprojector uses it for its column
// lerps (1.06x at 384^3 x 512 cone), the f16 forward projector for the x
// lerps of both y rows and then the z lerp of both (its y lerp stays in
// float).
constexpr bool kPackedF16Lerp = true;
__device__ __forceinline__ uint32_t pk_h2(float t) {
  return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(t, t));
}
__device__ __forceinline__ uint32_t pk_lerp_h(uint32_t A, uint32_t B, uint32_t t2) {
  uint32_t d, c;
  asm("v_pk_add_f16 %0, %1, %2 neg_lo:[0,1] neg_hi:[0,1]" : "=v"(d) : "v"(B), "v"(A));
  asm("v_pk_fma_f16 %0, %1, %2, %3" : "=v"(c) : "v"(t2), "v"(d), "v"(A));
  return c;
}
// {a, b} rounded to f16 (toward zero) in one dword
__device__ __forceinline__ uint32_t pk_h2(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(a, b));
}
// fma(t, D, A) on both f16 halves of D and A, t the low (HI = 0) or the high
// (HI = 1) half of t2 in both
template <int HI>
__device__ __forceinline__ uint32_t pk_fma_b(uint32_t t2, uint32_t D, uint32_t A) {
  uint32_t c;
  if (HI)
    asm("v_pk_fma_f16 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[1,1,1]" : "=v"(c) : "v"(t2), "v"(D), "v"(A));
  else
    asm("v_pk_fma_f16 %0, %1, %2, %3 op_sel_hi:[0,1,1]" : "=v"(c) : "v"(t2), "v"(D), "v"(A));
  return c;
}
// H.lo + t * H.hi of the two f16 halves of H, in float (one rounding)
__device__ __forceinline__ float fma_hl(float t, uint32_t H) {
  float c;
  asm("v_fma_mix_f32 %0, %1, %2, %2 op_sel:[0,1,0] op_sel_hi:[0,1,1]" : "=v"(c) : "v"(t), "v"(H));
  return c;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// Fixed point of the general-3D BP's deterministic sums.  A sample corner is
// at most |P| * 0.5 * 1.5^3 < 2^(e+1) for max|P| < 2^e (pmax: the bits of
// max|P|), and a run of samples merged in one cell holds at most 4 of them
// (a unit cell is crossed in under 4 half-voxel steps), so with scale
// 2^(27-e) a merged corner rounds to an int32, and 2^32 of them sum in an
// int64 without overflow.
__device__ __forceinline__ float g3d_fixed_scale(const uint32_t* pmax) {
  int e = 0;
  (void)frexpf(__uint_as_float(*pmax), &e);
  return ldexpf(1.f, 27 - e);
}
// A scaled corner v to the nearest integer, ties upward (floor(v + 0.5), one
// v_cvt_rpi_i32_f32), sign-extended to 64 bits.
__device__ __forceinline__ unsigned long long g3d_fixed(float v) {
  int i;
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(i) : "v"(v));
  return (unsigned long long)(long long)i;
}

// float threshold equivalent to the reference's `fabsf(d) >= 1e-5`
// (a float compared with the double literal): the least float >= 1e-5.
__device__ __forceinline__ float parallel_threshold() { return __uint_as_float(0x3727c5adu); }

struct R3 {
  float x, y, z;
};

// Flat block index -> logical tile index so that the 8 XCDs (which receive
// workgroups round-robin) each walk a contiguous range of tiles: rays of one
// view that run at the same time share one L2.  `nblk` is padded to a
// multiple of 8 by the launcher; ids >= n_logical exit.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nblk_padded) {
  uint32_t per = nblk_padded >> 3;
  return (b & 7u) * per + (b >> 3);
}

}  // namespace mexct
