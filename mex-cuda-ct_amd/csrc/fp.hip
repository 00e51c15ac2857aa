// fp.hip -- ray-marching forward projector on an x-fastest haloed volume:
// the general-3D projector (CUDAmex_general3D_FP.cu:59-150), whose rays have
// no shared structure, and the standard geometries when the volume is too
// large for the z-fastest kernel's 32-bit offsets (fpz.hip).
//
// One launch covers every view (the reference launches once per view with a
// sync and a D2H copy in between, CUDAmex_FP.cu:445-472).  A 256-thread
// workgroup owns a 16x16 patch of detector pixels of one view; the flat
// workgroup id is remapped so each XCD walks a contiguous run of patches.
// Each sample is an 8-tap trilinear fetch: four pair loads of (i, i+1)
// along x, lerps x, then y, then z (ray setup and clip in ray.h).
#include "common.h"
#include "internal.h"
#include "ray.h"

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

}  // namespace mexct
