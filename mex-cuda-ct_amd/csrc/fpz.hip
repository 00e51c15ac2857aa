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
// columns.  Loads are 32-bit buffer offsets with the (j+1), (i+1) and
// (i+1, j+1) corner strides in scalar offsets.
//
// Arithmetic follows the oracle operation by operation (ray setup and clip
// in ray.h; sample coordinate p - b, then - 0.5; lerps x, then y, then z),
// so the projections are the reference's float results.
#include "common.h"
#include "internal.h"
#include "ray.h"

namespace mexct {

struct FPZArgs {
  const void* vol;     // z-fastest haloed volume
  uint32_t vol_bytes;  // buffer range
  float* out;
  float* pre;          // OSC epilogue
  float* post;
  void* pre_pad;       // storage type S, FP16 scaled by 2^-e (e: exponent of *sino_max)
  void* post_pad;
  const uint32_t* sino_max;  // bits of max(|pre|, |post|), read when S is FP16
  const float2* sc;
  int W, H, rows, NX, NY, NZ;
  int PY, PZ;
  uint32_t off_j, off_i, off_ij;  // byte offsets of the (j+1), (i+1), (i+1,j+1) pairs
  float u0, v0, bx, by, bz, SR, RD;
  int first, stride;  // view = first + j*stride
  int PW, PV;         // OSC padded copies
  uint32_t n_ug, n_vg, n_logical;
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
template <>
__device__ __forceinline__ float2 buf_ld2<_Float16>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                    uint32_t soff) {
  typedef _Float16 h2v __attribute__((ext_vector_type(2)));
  h2v v = __builtin_bit_cast(h2v, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  return make_float2((float)v.x, (float)v.y);
}
template <>
__device__ __forceinline__ float2 buf_ld2<bf16_t>(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                  uint32_t soff) {
  const uint32_t u = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
  return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u));
}

// Corner pairs and weights of one sample.
struct Tap8 {
  float2 p00, p01, p10, p11;
  float ax, ay, az;
};

template <typename S>
__device__ __forceinline__ Tap8 fetch_z(__amdgpu_buffer_rsrc_t r, const FPZArgs& a, const R3& cur,
                                        float fnx, float fny, float fnz, int base) {
  Tap8 t;
  const float xb = (cur.x - a.bx) - 0.5f;
  const float yb = (cur.y - a.by) - 0.5f;
  const float zb = (cur.z - a.bz) - 0.5f;
  const float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
  t.ax = xb - fx;
  t.ay = yb - fy;
  t.az = zb - fz;
  const int i = (int)__builtin_amdgcn_fmed3f(fx, -2.f, fnx);
  const int j = (int)__builtin_amdgcn_fmed3f(fy, -2.f, fny);
  const int kk = (int)__builtin_amdgcn_fmed3f(fz, -2.f, fnz);
  // 24-bit multiplies: |i|, |j| <= N + 2 and PY, PZ < 2^23 (checked by the host)
  const uint32_t e = (uint32_t)(__mul24(__mul24(i, a.PY) + j, a.PZ) + kk + base);
  const uint32_t voff = e * (uint32_t)sizeof(S);
  t.p00 = buf_ld2<S>(r, voff, 0);         // (i,   j  ): z k, k+1
  t.p01 = buf_ld2<S>(r, voff, a.off_j);   // (i,   j+1)
  t.p10 = buf_ld2<S>(r, voff, a.off_i);   // (i+1, j  )
  t.p11 = buf_ld2<S>(r, voff, a.off_ij);  // (i+1, j+1)
  return t;
}

__device__ __forceinline__ float interp_z(const Tap8& t) {
  const float c00 = lerp(t.p00.x, t.p10.x, t.ax);
  const float c10 = lerp(t.p01.x, t.p11.x, t.ax);
  const float c01 = lerp(t.p00.y, t.p10.y, t.ax);
  const float c11 = lerp(t.p01.y, t.p11.y, t.ax);
  const float c0 = lerp(c00, c10, t.ay);
  const float c1 = lerp(c01, c11, t.ay);
  return lerp(c0, c1, t.az);
}

// One ray through the z-fastest volume, software-pipelined: the loads of
// sample k+1 are issued before sample k is interpolated.  A fetch past the
// last sample reads a clamped, in-range corner and is discarded.
template <typename S, int STEP2>
__device__ __forceinline__ float march_z(__amdgpu_buffer_rsrc_t r, const FPZArgs& a, R3 cur,
                                         const R3& inc, int n) {
  constexpr float step = STEP2 ? 0.5f : 1.0f;
  const float fnx = (float)a.NX, fny = (float)a.NY, fnz = (float)a.NZ;
  const int base = 2 * a.PY * a.PZ + 2 * a.PZ + 2;
  float dval = 0.f;
  if (n <= 0) return dval;
  // two register sets in ping-pong: while one sample is interpolated the
  // next sample's four loads are in flight
  Tap8 t0 = fetch_z<S>(r, a, cur, fnx, fny, fnz, base);
  int k = 0;
  for (; k + 1 < n; k += 2) {
    cur.x = cur.x + inc.x;
    cur.y = cur.y + inc.y;
    cur.z = cur.z + inc.z;
    const Tap8 t1 = fetch_z<S>(r, a, cur, fnx, fny, fnz, base);
    dval = dval + interp_z(t0) * step;
    cur.x = cur.x + inc.x;
    cur.y = cur.y + inc.y;
    cur.z = cur.z + inc.z;
    t0 = fetch_z<S>(r, a, cur, fnx, fny, fnz, base);
    dval = dval + interp_z(t1) * step;
  }
  if (k < n) dval = dval + interp_z(t0) * step;
  return dval;
}

// The FP16 copies of the OSC ratio terms are scaled by an exact power of two
// (counts reach 1e5, beyond FP16's 65504); the backprojection only forms the
// ratio of two sums of them, which the common scale leaves unchanged.
template <typename S>
__device__ __forceinline__ float scale_pow2(float v, int e) {
  if (__is_same(S, _Float16)) return ldexpf(v, -e);
  return v;
}

// EPI 0: store the ray sum.  EPI 1: OSC ratio terms of CUDAmex_oscIter.cu:262-268
// written in place, plus zero-haloed copies for the dual backprojection.
template <int GEOM, typename S, int EPI>
__global__ __launch_bounds__(256) void fpz_kernel(FPZArgs a) {
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  if (b >= a.n_logical) return;
  const uint32_t vg = b % a.n_vg;
  const uint32_t rest = b / a.n_vg;
  const uint32_t ug = rest % a.n_ug;
  const uint32_t jv = rest / a.n_ug;
  const int ix = (int)ug * 4 + (int)(threadIdx.x >> 6);
  const int iy = (int)vg * 64 + (int)(threadIdx.x & 63);
  if (ix >= a.W || iy >= a.H) return;
  const int view = a.first + (int)jv * a.stride;
  const size_t p = (size_t)view * a.W * a.H + (size_t)iy * a.W + ix;
  float preval = 0.f, postval = 0.f;
  size_t q = 0;
  int sexp = 0;  // FP16 copies hold value * 2^-sexp, so that |value| <= 1
  if (EPI == 0) {
    if (iy >= a.rows) {  // the reference's grid never writes these rows
      a.out[p] = 0.f;
      return;
    }
  } else {
    preval = a.pre[p];
    postval = a.post[p];
    q = (size_t)jv * a.PV + (size_t)(iy + HALO) * a.PW + (ix + HALO);
    if (__is_same(S, _Float16)) (void)frexpf(__uint_as_float(*a.sino_max), &sexp);
    if (iy >= a.rows) {  // unprocessed rows keep their values
      ((S*)a.pre_pad)[q] = Store<S>::from(scale_pow2<S>(preval, sexp));
      ((S*)a.post_pad)[q] = Store<S>::from(scale_pow2<S>(postval, sexp));
      return;
    }
  }
  const float2 scv = a.sc[view];
  R3 st, en;
  std_ray<GEOM>(ix, iy, a.u0, a.v0, scv.x, scv.y, a.SR, a.RD, st, en);
  // CUDAmex_FP marches cone rays with step 1.0 (CUDAmex_FP.cu:231), par and
  // fan with 0.5; CUDAmex_oscIter uses 0.5 for every geometry (:244).
  constexpr int STEP2 = (EPI == 1 || GEOM != GEOM_CONE) ? 1 : 0;
  R3 cur, inc;
  const int n = clip_ray(st, en, a.bx, a.by, a.bz, STEP2 ? 0.5f : 1.0f, cur, inc);
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc((void*)a.vol, 0, (int)a.vol_bytes, 0x00020000);
  const float dval = march_z<S, STEP2>(r, a, cur, inc, n);
  if (EPI == 0) {
    a.out[p] = dval;
  } else {
    const float np = (preval * expf(-dval) - postval);
    const float npo = preval * expf(-dval) * dval;
    a.pre[p] = np;
    a.post[p] = npo;
    ((S*)a.pre_pad)[q] = Store<S>::from(scale_pow2<S>(np, sexp));
    ((S*)a.post_pad)[q] = Store<S>::from(scale_pow2<S>(npo, sexp));
  }
}

static FPZArgs make_zargs(const void* vol, int st_dt, const float2* sc, int W, int H, int nviews,
                          float sad) {
  FPZArgs k{};
  k.vol = vol;
  k.sc = sc;
  k.W = W;
  k.H = H;
  const int nb = (W + 31) / 32;  // CUDAmex_FP.cu:434-440 grid extent
  k.rows = H < 32 * nb ? H : 32 * nb;
  k.NX = W;
  k.NY = W;
  k.NZ = H;
  k.PY = W + 2 * HALO;
  k.PZ = H + 2 * HALO;
  const uint32_t es = st_dt == 0 ? 4u : 2u;
  k.vol_bytes = (uint32_t)((size_t)(W + 2 * HALO) * k.PY * k.PZ * es);
  k.off_j = (uint32_t)k.PZ * es;
  k.off_i = (uint32_t)k.PY * (uint32_t)k.PZ * es;
  k.off_ij = k.off_i + k.off_j;
  k.u0 = float(W - 1) / 2;
  k.v0 = float(H - 1) / 2;
  k.bx = (float)(-(double)(float)W / 2.0);
  k.by = k.bx;
  k.bz = (float)(-(double)(float)H / 2.0);
  k.SR = sad;
  k.RD = sad;
  k.first = 0;
  k.stride = 1;
  k.n_ug = (uint32_t)((W + 3) / 4);
  k.n_vg = (uint32_t)((H + 63) / 64);
  k.n_logical = k.n_ug * k.n_vg * (uint32_t)nviews;
  return k;
}

template <int EPI, typename S>
static void fpz_dispatch(const FPZArgs& k, int geom, hipStream_t st) {
  const unsigned g = (k.n_logical + 7u) & ~7u;
  switch (geom) {
    case GEOM_PAR: fpz_kernel<GEOM_PAR, S, EPI><<<g, 256, 0, st>>>(k); break;
    case GEOM_FAN: fpz_kernel<GEOM_FAN, S, EPI><<<g, 256, 0, st>>>(k); break;
    default: fpz_kernel<GEOM_CONE, S, EPI><<<g, 256, 0, st>>>(k); break;
  }
}

hipError_t launch_fpz(const FPZLaunch& a, hipStream_t st) {
  FPZArgs k = make_zargs(a.vol, a.st_dt, a.sc, a.W, a.H, a.nviews, a.sad);
  k.out = a.out;
  if (k.n_logical == 0) return hipSuccess;
  if (a.st_dt == 0)
    fpz_dispatch<0, float>(k, a.geom, st);
  else if (a.st_dt == 1)
    fpz_dispatch<0, _Float16>(k, a.geom, st);
  else
    fpz_dispatch<0, bf16_t>(k, a.geom, st);
  return hipGetLastError();
}

hipError_t launch_osc_fp(const OSCFPLaunch& a, hipStream_t st) {
  FPZArgs k = make_zargs(a.mu_z, a.st_dt, a.sc, a.W, a.H, a.count, a.sad);
  k.pre = a.pre;
  k.post = a.post;
  k.pre_pad = a.pre_pad;
  k.post_pad = a.post_pad;
  k.sino_max = a.sino_max;
  k.first = a.first;
  k.stride = a.stride;
  k.PW = a.W + 2 * HALO;
  k.PV = (a.W + 2 * HALO) * (a.H + 2 * HALO);
  if (k.n_logical == 0) return hipSuccess;
  if (a.st_dt == 0)
    fpz_dispatch<1, float>(k, a.geom, st);
  else if (a.st_dt == 1)
    fpz_dispatch<1, _Float16>(k, a.geom, st);
  else
    fpz_dispatch<1, bf16_t>(k, a.geom, st);
  return hipGetLastError();
}

}  // namespace mexct
