// ray.h -- ray setup and slab clipping shared by the forward projectors.
//
//   ray setup    standard geometries CUDAmex_FP.cu:64-77 (par), :162-175
//                (cone), :259-272 (fan); general-3D
//                CUDAmex_general3D_FP.cu:69-77
//   slab clip    CUDAmex_FP.cu:85-137 (same text in every projector)
// The march that follows (CUDAmex_FP.cu:139-148) loops `while (travelled <
// distance)` with travelled += step; travelled = k*step is exact for step
// 0.5 or 1, so the loop count is n = ceil(distance / step).  Positions stay
// a running float sum as in the reference.
#pragma once

#include "common.h"

namespace mexct {

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

}  // namespace mexct
