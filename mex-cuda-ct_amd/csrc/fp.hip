// fp.hip -- ray-marching forward projectors for gfx950.
//
// One launch covers every view (the reference launches once per view with a
// sync and a D2H copy in between, CUDAmex_FP.cu:445-472).  A 256-thread
// workgroup owns a 16x16 patch of detector pixels of one view; the flat
// workgroup id is remapped so each XCD walks a contiguous run of patches
// and the rays in flight on one XCD sample neighbouring voxels through one
// L2.  Each thread marches one ray:
//
//   ray setup    standard geometries CUDAmex_FP.cu:64-77 (par), :162-175
//                (cone), :259-272 (fan); general-3D
//                CUDAmex_general3D_FP.cu:69-77
//   slab clip    CUDAmex_FP.cu:85-127 (same text in every projector)
//   march        CUDAmex_FP.cu:129-148: `while (travelled < distance)`
//                with travelled += step is the exact count
//                n = ceil(distance / step) because travelled = k*step is
//                exact for step 0.5 or 1; the position update stays a
//                running float sum, as in the reference, so every sample
//                sits where the reference's sample sits.
//   sample       8-tap trilinear on the haloed volume: four pair loads
//                of (i, i+1) along x, lerps x, then y, then z.
#include "common.h"
#include "internal.h"

namespace mexct {

struct FPKArgs {
  const void* vol;
  float* out;
  const float2* sc;
  const float* pts;
  int64_t pts_view_stride;
  int W, H, rows;
  int NX, NY, NZ, PX, PXY;
  float u0, v0, bx, by, bz, SR, RD, step;
  uint32_t tiles_x, tiles_per_view, n_logical;
};

// standard-geometry ray of detector pixel (ix, iy)
template <int GEOM>
__device__ __forceinline__ void std_ray(int ix, int iy, float u0, float v0, float sB, float cB,
                                        float SR, float RD, R3& st, R3& en) {
  if (GEOM == GEOM_PAR) {
    float u = (float)ix - u0;
    float v = (float)iy - v0;
    st.x = (SR * cB) - u * sB;
    st.y = (SR * sB) + u * cB;
    st.z = v;
    en.x = -RD * cB - u * sB;
    en.y = -RD * sB + u * cB;
    en.z = v;
  } else if (GEOM == GEOM_FAN) {
    float u = ((float)ix - u0) * (SR + RD) / SR;
    float v = ((float)iy - v0);
    st.x = (SR * cB);
    st.y = (SR * sB);
    st.z = v;
    en.x = -RD * cB - u * sB;
    en.y = -RD * sB + u * cB;
    en.z = v;
  } else {
    float u = ((float)ix - u0) * (SR + RD) / SR;
    float v = ((float)iy - v0) * (SR + RD) / SR;
    st.x = (SR * cB);
    st.y = (SR * sB);
    st.z = 0.0f;
    en.x = -RD * cB - u * sB;
    en.y = -RD * sB + u * cB;
    en.z = v;
  }
}

// general-3D ray: endpoints rotated about z (CUDAmex_general3D_FP.cu:69-77)
__device__ __forceinline__ void g3d_ray(const float* __restrict__ p, size_t npts, float sB,
                                        float cB, R3& st, R3& en) {
  float p0 = p[0], p1 = p[npts], p2 = p[2 * npts];
  float p3 = p[3 * npts], p4 = p[4 * npts], p5 = p[5 * npts];
  st.x = (p0)*cB - (p1)*sB;
  st.y = (p0)*sB + (p1)*cB;
  st.z = (p2);
  en.x = (p3)*cB - (p4)*sB;
  en.y = (p3)*sB + (p4)*cB;
  en.z = (p5);
}

// Largest sample count marched per ray.  travelled = k*step stays exact
// below 2^24 steps; a ray needing more would never terminate in the
// reference, so the count is capped instead of hanging the device.
constexpr int kMaxSamples = 1 << 24;

// Slab clip (CUDAmex_FP.cu:85-137): first sample position, per-sample
// increment and sample count.
__device__ __forceinline__ int clip_ray(const R3& st, const R3& en, float bx, float by, float bz,
                                        float stepsize, R3& cur, R3& inc) {
  const float TH = parallel_threshold();
  const float qnan = __builtin_nanf("");
  float aX1, aX2, aY1, aY2, aZ1, aZ2;
  if (fabsf(en.x - st.x) >= TH) {
    aX1 = (bx - st.x) / (en.x - st.x);
    aX2 = (-bx - st.x) / (en.x - st.x);
  } else {
    aX1 = qnan;
    aX2 = qnan;
  }
  if (fabsf(en.y - st.y) >= TH) {
    aY1 = (by - st.y) / (en.y - st.y);
    aY2 = (-by - st.y) / (en.y - st.y);
  } else {
    aY1 = qnan;
    aY2 = qnan;
  }
  if (fabsf(en.z - st.z) >= TH) {
    aZ1 = (bz - st.z) / (en.z - st.z);
    aZ2 = (-bz - st.z) / (en.z - st.z);
  } else {
    aZ1 = qnan;
    aZ2 = qnan;
  }
  float aXmin = fminf(aX1, aX2), aXmax = fmaxf(aX1, aX2);
  float aYmin = fminf(aY1, aY2), aYmax = fmaxf(aY1, aY2);
  float aZmin = fminf(aZ1, aZ2), aZmax = fmaxf(aZ1, aZ2);
  float tmp = fmaxf(aXmin, fmaxf(aYmin, aZmin));
  float aMin = fmaxf(0.f, tmp);
  tmp = fminf(aXmax, fminf(aYmax, aZmax));
  float aMax = fminf(1.f, tmp);
  cur.x = st.x + aMin * (en.x - st.x);
  cur.y = st.y + aMin * (en.y - st.y);
  cur.z = st.z + aMin * (en.z - st.z);
  R3 e2;
  e2.x = st.x + aMax * (en.x - st.x);
  e2.y = st.y + aMax * (en.y - st.y);
  e2.z = st.z + aMax * (en.z - st.z);
  float distance = sqrtf((e2.x - cur.x) * (e2.x - cur.x) + (e2.y - cur.y) * (e2.y - cur.y) +
                         (e2.z - cur.z) * (e2.z - cur.z));
  inc.x = (e2.x - cur.x) / distance * stepsize;
  inc.y = (e2.y - cur.y) / distance * stepsize;
  inc.z = (e2.z - cur.z) / distance * stepsize;
  if (!(distance > 0.f)) return 0;
  float q = ceilf(distance / stepsize);
  return q < (float)kMaxSamples ? (int)q : kMaxSamples;
}

// 8-tap trilinear sample of a haloed volume at texture coordinate (x,y,z).
template <typename S>
__device__ __forceinline__ float tex3_halo(const S* __restrict__ vol, int PX, int PXY, int NX,
                                           int NY, int NZ, float x, float y, float z) {
  float xb = x - 0.5f, yb = y - 0.5f, zb = z - 0.5f;
  float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
  float ax = xb - fx, ay = yb - fy, az = zb - fz;
  int i = (int)fminf(fmaxf(fx, -2.f), (float)NX);
  int j = (int)fminf(fmaxf(fy, -2.f), (float)NY);
  int k = (int)fminf(fmaxf(fz, -2.f), (float)NZ);
  uint32_t base = (uint32_t)(k + HALO) * (uint32_t)PXY + (uint32_t)(j + HALO) * (uint32_t)PX +
                  (uint32_t)(i + HALO);
  const S* p = vol + base;
  float2 p00 = Store<S>::ld2(p);
  float2 p10 = Store<S>::ld2(p + PX);
  float2 p01 = Store<S>::ld2(p + PXY);
  float2 p11 = Store<S>::ld2(p + PXY + PX);
  float c00 = lerp(p00.x, p00.y, ax);
  float c10 = lerp(p10.x, p10.y, ax);
  float c01 = lerp(p01.x, p01.y, ax);
  float c11 = lerp(p11.x, p11.y, ax);
  float c0 = lerp(c00, c10, ay);
  float c1 = lerp(c01, c11, ay);
  return lerp(c0, c1, az);
}

template <typename S>
__device__ __forceinline__ float march(const S* __restrict__ vol, const FPKArgs& a, R3 cur,
                                       const R3& inc, int n) {
  float dval = 0.f;
  for (int k = 0; k < n; ++k) {
    float v = tex3_halo<S>(vol, a.PX, a.PXY, a.NX, a.NY, a.NZ, cur.x - a.bx, cur.y - a.by,
                           cur.z - a.bz);
    dval = dval + v * a.step;
    cur.x = cur.x + inc.x;
    cur.y = cur.y + inc.y;
    cur.z = cur.z + inc.z;
  }
  return dval;
}

template <int GEOM, typename S>
__global__ __launch_bounds__(256) void fp_kernel(FPKArgs a) {
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  if (b >= a.n_logical) return;
  const uint32_t view = b / a.tiles_per_view;
  const uint32_t t = b - view * a.tiles_per_view;
  const int ix = (int)(t % a.tiles_x) * 16 + (int)(threadIdx.x & 15);
  const int iy = (int)(t / a.tiles_x) * 16 + (int)(threadIdx.x >> 4);
  if (ix >= a.W || iy >= a.H) return;
  float* o = a.out + (size_t)view * a.W * a.H + (size_t)iy * a.W + ix;
  if (GEOM != GEOM_G3D && iy >= a.rows) {
    *o = 0.f;
    return;
  }
  const float2 scv = a.sc[view];
  R3 st, en;
  if (GEOM == GEOM_G3D) {
    g3d_ray(a.pts + (size_t)view * a.pts_view_stride + (size_t)iy * a.W + ix, (size_t)a.W * a.H,
            scv.x, scv.y, st, en);
  } else {
    std_ray<GEOM>(ix, iy, a.u0, a.v0, scv.x, scv.y, a.SR, a.RD, st, en);
  }
  R3 cur, inc;
  const int n = clip_ray(st, en, a.bx, a.by, a.bz, a.step, cur, inc);
  *o = march<S>((const S*)a.vol, a, cur, inc, n);
}

// OSC forward projection over the views first + j*stride (j < count) of one
// subset with the ratio terms of CUDAmex_oscIter.cu:262-268 written in
// place, plus zero-haloed copies for the dual backprojection.
struct OSCFPKArgs {
  FPKArgs f;
  float* pre;
  float* post;
  float* pre_pad;
  float* post_pad;
  int first, stride;
  int PW, PV;
};

template <int GEOM>
__global__ __launch_bounds__(256) void osc_fp_kernel(OSCFPKArgs a) {
  const FPKArgs& f = a.f;
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  if (b >= f.n_logical) return;
  const uint32_t j = b / f.tiles_per_view;
  const uint32_t t = b - j * f.tiles_per_view;
  const int ix = (int)(t % f.tiles_x) * 16 + (int)(threadIdx.x & 15);
  const int iy = (int)(t / f.tiles_x) * 16 + (int)(threadIdx.x >> 4);
  if (ix >= f.W || iy >= f.H) return;
  const int view = a.first + (int)j * a.stride;
  const size_t p = (size_t)view * f.W * f.H + (size_t)iy * f.W + ix;
  const size_t q = (size_t)j * a.PV + (size_t)(iy + HALO) * a.PW + (ix + HALO);
  const float preval = a.pre[p], postval = a.post[p];
  if (iy >= f.rows) {  // rows the reference's grid never reaches keep their values
    a.pre_pad[q] = preval;
    a.post_pad[q] = postval;
    return;
  }
  const float2 scv = f.sc[view];
  R3 st, en;
  std_ray<GEOM>(ix, iy, f.u0, f.v0, scv.x, scv.y, f.SR, f.RD, st, en);
  R3 cur, inc;
  const int n = clip_ray(st, en, f.bx, f.by, f.bz, f.step, cur, inc);
  const float dval = march<float>((const float*)f.vol, f, cur, inc, n);
  const float np = (preval * expf(-dval) - postval);
  const float npo = preval * expf(-dval) * dval;
  a.pre[p] = np;
  a.post[p] = npo;
  a.pre_pad[q] = np;
  a.post_pad[q] = npo;
}

static FPKArgs make_args(const void* vol, float* out, const float2* sc, int geom, int W, int H,
                         int nviews, int NX, int NY, int NZ, float sad, float step) {
  FPKArgs k;
  k.vol = vol;
  k.out = out;
  k.sc = sc;
  k.pts = nullptr;
  k.pts_view_stride = 0;
  k.W = W;
  k.H = H;
  const int nb = (W + 31) / 32;  // CUDAmex_FP.cu:434-440 grid extent
  k.rows = H < 32 * nb ? H : 32 * nb;
  k.NX = NX;
  k.NY = NY;
  k.NZ = NZ;
  k.PX = NX + 2 * HALO;
  k.PXY = (NX + 2 * HALO) * (NY + 2 * HALO);
  k.u0 = float(W - 1) / 2;
  k.v0 = float(H - 1) / 2;
  k.bx = (float)(-(double)(float)NX / 2.0);
  k.by = (float)(-(double)(float)NY / 2.0);
  k.bz = (float)(-(double)(float)NZ / 2.0);
  k.SR = sad;
  k.RD = sad;
  k.step = step;
  k.tiles_x = (uint32_t)((W + 15) / 16);
  k.tiles_per_view = k.tiles_x * (uint32_t)((H + 15) / 16);
  k.n_logical = k.tiles_per_view * (uint32_t)nviews;
  (void)geom;
  return k;
}

static unsigned padded_grid(uint32_t n_logical) { return (n_logical + 7u) & ~7u; }

template <typename S>
static void fp_dispatch(const FPKArgs& k, int geom, hipStream_t st) {
  const unsigned g = padded_grid(k.n_logical);
  switch (geom) {
    case GEOM_PAR: fp_kernel<GEOM_PAR, S><<<g, 256, 0, st>>>(k); break;
    case GEOM_FAN: fp_kernel<GEOM_FAN, S><<<g, 256, 0, st>>>(k); break;
    case GEOM_CONE: fp_kernel<GEOM_CONE, S><<<g, 256, 0, st>>>(k); break;
    default: fp_kernel<GEOM_G3D, S><<<g, 256, 0, st>>>(k); break;
  }
}

hipError_t launch_fp(const FPLaunch& a, hipStream_t st) {
  FPKArgs k = make_args(a.vol, a.out, a.sc, a.geom, a.W, a.H, a.nviews, a.NX, a.NY, a.NZ, a.sad,
                        a.step);
  k.pts = a.pts;
  k.pts_view_stride = a.pts_view_stride;
  if (k.n_logical == 0) return hipSuccess;
  if (a.st_dt == 0)
    fp_dispatch<float>(k, a.geom, st);
  else if (a.st_dt == 1)
    fp_dispatch<_Float16>(k, a.geom, st);
  else
    fp_dispatch<bf16_t>(k, a.geom, st);
  return hipGetLastError();
}

hipError_t launch_osc_fp(const OSCFPLaunch& a, hipStream_t st) {
  OSCFPKArgs k;
  k.f = make_args(a.mu_pad, nullptr, a.sc, a.geom, a.W, a.H, a.count, a.W, a.W, a.H, a.sad, 0.5f);
  k.pre = a.pre;
  k.post = a.post;
  k.pre_pad = a.pre_pad;
  k.post_pad = a.post_pad;
  k.first = a.first;
  k.stride = a.stride;
  k.PW = a.W + 2 * HALO;
  k.PV = (a.W + 2 * HALO) * (a.H + 2 * HALO);
  if (k.f.n_logical == 0) return hipSuccess;
  const unsigned g = padded_grid(k.f.n_logical);
  switch (a.geom) {
    case GEOM_PAR: osc_fp_kernel<GEOM_PAR><<<g, 256, 0, st>>>(k); break;
    case GEOM_FAN: osc_fp_kernel<GEOM_FAN><<<g, 256, 0, st>>>(k); break;
    default: osc_fp_kernel<GEOM_CONE><<<g, 256, 0, st>>>(k); break;
  }
  return hipGetLastError();
}

}  // namespace mexct
