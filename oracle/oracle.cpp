// oracle.cpp -- CPU restatement of the mex-cuda-CT operators.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Build: g++ -O2 -std=c++17 -ffp-contract=off -fno-fast-math -pthread
// (oracle/Makefile).  FP contraction must stay off: every float
// expression below is evaluated with one rounding per operation, in the
// order the reference's source text gives.
//
// Each function names the reference lines it restates.  The texture
// sampler is the public CUDA texture-fetch formula (unnormalised
// coordinates, x_B = x - 0.5, i = floor(x_B), a = x_B - i) with exact
// FP32 weights, lerp(a, b, t) = fmaf(t, b - a, a), applied x, then y,
// then z.  Clamp uses the edge texel, border uses 0.  A second weight
// mode (orc_set_texture_weight_bits(8)) rounds each fraction to 8 bits, the
// texture unit's documented weight format, to measure how far exact
// weights sit from what NVIDIA hardware computes.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <thread>
#include <vector>

namespace {

int g_threads = 0;
// Fraction bits of the texture weights: 0 exact FP32, else rounded to
// multiples of 2^-bits (see weight_of).
int g_weight_bits = 0;

int nthreads() {
  if (g_threads > 0) return g_threads;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

// Dynamic parallel-for over [0, n): work items are claimed in chunks.
void parallel_for(int64_t n, const std::function<void(int64_t)>& fn, int64_t chunk = 1) {
  int nt = nthreads();
  if (nt <= 1 || n <= chunk) {
    for (int64_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<int64_t> next(0);
  auto worker = [&]() {
    for (;;) {
      int64_t s = next.fetch_add(chunk);
      if (s >= n) break;
      int64_t e = std::min(n, s + chunk);
      for (int64_t i = s; i < e; ++i) fn(i);
    }
  };
  std::vector<std::thread> pool;
  int nw = (int)std::min<int64_t>(nt, (n + chunk - 1) / chunk);
  for (int t = 1; t < nw; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
}

template <typename R>
inline R nan_of() { return std::numeric_limits<R>::quiet_NaN(); }

template <typename R>
inline R lerp_(R a, R b, R t) { return std::fma(t, b - a, a); }

// Integral-valued float -> int with saturation (the coordinate is finite
// for every ray the reference builds; saturation only guards pathology).
template <typename R>
inline int to_index(R f) {
  if (!(f > R(-1073741824.0))) return -1073741824;
  if (f > R(1073741824.0)) return 1073741824;
  return (int)f;
}

template <typename R, typename E>
inline R fetch3(const E* T, int nx, int ny, int nz, int i, int j, int k, int mode) {
  if (mode == ORC_CLAMP) {
    i = std::min(std::max(i, 0), nx - 1);
    j = std::min(std::max(j, 0), ny - 1);
    k = std::min(std::max(k, 0), nz - 1);
  } else if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) {
    return R(0);
  }
  return R(T[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k)]);
}

template <typename R, typename E>
inline R fetch2(const E* T, int nx, int ny, int i, int j, int mode) {
  if (mode == ORC_CLAMP) {
    i = std::min(std::max(i, 0), nx - 1);
    j = std::min(std::max(j, 0), ny - 1);
  } else if (i < 0 || i >= nx || j < 0 || j >= ny) {
    return R(0);
  }
  return R(T[(size_t)i + (size_t)nx * (size_t)j]);
}

// Filter weight of fraction a.  The public CUDA Programming Guide
// (texture fetching appendix, linear filtering) stores it in 9-bit fixed
// point with 8 fractional bits; with g_weight_bits = 8 the fraction is
// rounded to the nearest multiple of 2^-8 (ties to even).  The rounding
// rule of the hardware is not documented; nearest is the model here.
template <typename R>
inline R weight_of(R a) {
  if (g_weight_bits <= 0) return a;
  const R s = (R)(1 << g_weight_bits);
  return std::nearbyint(a * s) / s;
}

// tex3D(reconTex, x, y, z) with cudaFilterModeLinear, unnormalised
// coordinates (CUDAmex_FP.cu:405-409, CUDAmex_general3D_FP.cu:217-221,
// CUDAmex_oscIter.cu:458-462).
template <typename R, typename E>
R tex3(const E* T, int nx, int ny, int nz, R x, R y, R z, int mode) {
  const R h = R(0.5);
  R xb = x - h, yb = y - h, zb = z - h;
  R fx = std::floor(xb), fy = std::floor(yb), fz = std::floor(zb);
  R ax = weight_of(xb - fx), ay = weight_of(yb - fy), az = weight_of(zb - fz);
  int i = to_index(fx), j = to_index(fy), k = to_index(fz);
  R t000 = fetch3<R, E>(T, nx, ny, nz, i, j, k, mode);
  R t100 = fetch3<R, E>(T, nx, ny, nz, i + 1, j, k, mode);
  R t010 = fetch3<R, E>(T, nx, ny, nz, i, j + 1, k, mode);
  R t110 = fetch3<R, E>(T, nx, ny, nz, i + 1, j + 1, k, mode);
  R t001 = fetch3<R, E>(T, nx, ny, nz, i, j, k + 1, mode);
  R t101 = fetch3<R, E>(T, nx, ny, nz, i + 1, j, k + 1, mode);
  R t011 = fetch3<R, E>(T, nx, ny, nz, i, j + 1, k + 1, mode);
  R t111 = fetch3<R, E>(T, nx, ny, nz, i + 1, j + 1, k + 1, mode);
  R c00 = lerp_(t000, t100, ax);
  R c10 = lerp_(t010, t110, ax);
  R c01 = lerp_(t001, t101, ax);
  R c11 = lerp_(t011, t111, ax);
  R c0 = lerp_(c00, c10, ay);
  R c1 = lerp_(c01, c11, ay);
  return lerp_(c0, c1, az);
}

// tex2D(projtex, x, y) (CUDAmex_BP.cu:270-273, CUDAmex_oscIter.cu:417-425).
template <typename R, typename E>
R tex2(const E* T, int nx, int ny, R x, R y, int mode) {
  const R h = R(0.5);
  R xb = x - h, yb = y - h;
  R fx = std::floor(xb), fy = std::floor(yb);
  R ax = weight_of(xb - fx), ay = weight_of(yb - fy);
  int i = to_index(fx), j = to_index(fy);
  R t00 = fetch2<R, E>(T, nx, ny, i, j, mode);
  R t10 = fetch2<R, E>(T, nx, ny, i + 1, j, mode);
  R t01 = fetch2<R, E>(T, nx, ny, i, j + 1, mode);
  R t11 = fetch2<R, E>(T, nx, ny, i + 1, j + 1, mode);
  R c0 = lerp_(t00, t10, ax);
  R c1 = lerp_(t01, t11, ax);
  return lerp_(c0, c1, ay);
}

template <typename R>
struct V3 {
  R x, y, z;
};

// The ray march shared by every forward projector: slab clip against the
// box [b, -b], then march from alphaMin with a fixed step while
// travelled < distance (CUDAmex_FP.cu:85-148; identical text at :183-245,
// :280-342, CUDAmex_general3D_FP.cu:84-146, CUDAmex_oscIter.cu:196-259).
template <typename R, typename E>
R march(const E* T, int nx, int ny, int nz, R bx, R by, R bz, V3<R> st, V3<R> en, R stepsize,
        int mode, int64_t* nsamp) {
  R aX1, aX2, aY1, aY2, aZ1, aZ2;
  // `fabsf(d) >= 1e-5` compares the float against the double literal.
  if ((double)std::fabs(en.x - st.x) >= 1e-5) {
    aX1 = (bx - st.x) / (en.x - st.x);
    aX2 = (-bx - st.x) / (en.x - st.x);
  } else {
    aX1 = nan_of<R>();
    aX2 = nan_of<R>();
  }
  if ((double)std::fabs(en.y - st.y) >= 1e-5) {
    aY1 = (by - st.y) / (en.y - st.y);
    aY2 = (-by - st.y) / (en.y - st.y);
  } else {
    aY1 = nan_of<R>();
    aY2 = nan_of<R>();
  }
  if ((double)std::fabs(en.z - st.z) >= 1e-5) {
    aZ1 = (bz - st.z) / (en.z - st.z);
    aZ2 = (-bz - st.z) / (en.z - st.z);
  } else {
    aZ1 = nan_of<R>();
    aZ2 = nan_of<R>();
  }
  R aXmin = std::fmin(aX1, aX2), aXmax = std::fmax(aX1, aX2);
  R aYmin = std::fmin(aY1, aY2), aYmax = std::fmax(aY1, aY2);
  R aZmin = std::fmin(aZ1, aZ2), aZmax = std::fmax(aZ1, aZ2);
  R tmp = std::fmax(aXmin, std::fmax(aYmin, aZmin));
  R aMin = std::fmax(R(0), tmp);
  tmp = std::fmin(aXmax, std::fmin(aYmax, aZmax));
  R aMax = std::fmin(R(1), tmp);

  V3<R> cur, e2;
  cur.x = st.x + aMin * (en.x - st.x);
  cur.y = st.y + aMin * (en.y - st.y);
  cur.z = st.z + aMin * (en.z - st.z);
  e2.x = st.x + aMax * (en.x - st.x);
  e2.y = st.y + aMax * (en.y - st.y);
  e2.z = st.z + aMax * (en.z - st.z);
  R distance = std::sqrt((e2.x - cur.x) * (e2.x - cur.x) + (e2.y - cur.y) * (e2.y - cur.y) +
                         (e2.z - cur.z) * (e2.z - cur.z));
  R travelled = R(0);
  R sX = (e2.x - cur.x) / distance * stepsize;
  R sY = (e2.y - cur.y) / distance * stepsize;
  R sZ = (e2.z - cur.z) / distance * stepsize;
  R dval = R(0);
  int64_t n = 0;
  while (travelled < distance) {
    dval += tex3<R, E>(T, nx, ny, nz, cur.x - bx, cur.y - by, cur.z - bz, mode) * stepsize;
    cur.x += sX;
    cur.y += sY;
    cur.z += sZ;
    travelled += stepsize;
    ++n;
  }
  if (nsamp) *nsamp += n;
  return dval;
}

template <typename R>
inline void sincos_of(float beta, R* s, R* c);
template <>
inline void sincos_of<float>(float beta, float* s, float* c) {
  // host sincosf on the float angle (CUDAmex_FP.cu:449)
  sincosf(beta, s, c);
}
template <>
inline void sincos_of<double>(float beta, double* s, double* c) {
  *s = std::sin((double)beta);
  *c = std::cos((double)beta);
}

// Ray endpoints and step of the standard geometries for detector pixel
// (ix, iy): par CUDAmex_FP.cu:64-77,133; cone :162-175,231; fan :259-272,328.
// `coneStep` selects 1.0 for cone (CUDAmex_FP) or 0.5 (CUDAmex_oscIter.cu:244).
template <typename R>
void std_ray(int geom, int ix, int iy, R u0, R v0, R sinB, R cosB, R SR, R RD, V3<R>* st, V3<R>* en,
             R* step, R coneStep) {
  R u, v;
  switch (geom) {
    case 0:
      u = (R)ix - u0;
      v = (R)iy - v0;
      st->x = (SR * cosB) - u * sinB;
      st->y = (SR * sinB) + u * cosB;
      st->z = v;
      en->x = -RD * cosB - u * sinB;
      en->y = -RD * sinB + u * cosB;
      en->z = v;
      *step = R(1) / R(2);
      break;
    case 1:
      u = ((R)ix - u0) * (SR + RD) / SR;
      v = ((R)iy - v0);
      st->x = (SR * cosB);
      st->y = (SR * sinB);
      st->z = v;
      en->x = -RD * cosB - u * sinB;
      en->y = -RD * sinB + u * cosB;
      en->z = v;
      *step = R(1) / R(2);
      break;
    default:
      u = ((R)ix - u0) * (SR + RD) / SR;
      v = ((R)iy - v0) * (SR + RD) / SR;
      st->x = (SR * cosB);
      st->y = (SR * sinB);
      st->z = R(0);
      en->x = -RD * cosB - u * sinB;
      en->y = -RD * sinB + u * cosB;
      en->z = v;
      *step = coneStep;
      break;
  }
}

// CUDAmex_FP: per-view launch over a (ceil(W/32))^2 grid of 32x32 blocks,
// so detector rows iy >= 32*ceil(W/32) are never written
// (CUDAmex_FP.cu:432-440).
template <typename R, typename O>
int64_t fp_std(const float* vol, int W, int H, R sad, const float* angles, int nProj, int geom,
               O* out) {
  const R u0 = R(W - 1) / 2;
  const R v0 = R(H - 1) / 2;
  const R bx = -(R)W / R(2), by = -(R)W / R(2), bz = -(R)H / R(2);
  const int nb = (W + 31) / 32;
  const int rows = std::min(H, 32 * nb);
  std::memset(out, 0, sizeof(O) * (size_t)W * H * nProj);
  std::vector<int64_t> counts((size_t)nProj * rows, 0);
  parallel_for((int64_t)nProj * rows, [&](int64_t item) {
    int view = (int)(item / rows), iy = (int)(item % rows);
    R s, c;
    sincos_of<R>(angles[view], &s, &c);
    int64_t n = 0;
    for (int ix = 0; ix < W; ++ix) {
      V3<R> st, en;
      R step;
      std_ray<R>(geom, ix, iy, u0, v0, s, c, sad, sad, &st, &en, &step, R(1));
      out[(size_t)ix + (size_t)W * iy + (size_t)W * H * view] =
          (O)march<R, float>(vol, W, W, H, bx, by, bz, st, en, step, ORC_CLAMP, &n);
    }
    counts[item] = n;
  });
  int64_t tot = 0;
  for (auto v : counts) tot += v;
  return tot;
}

// Voxel-driven backprojection of one column for all views, accumulating
// in view order (CUDAmex_BP.cu:69-103 par, :127-161 fan, :189-220 cone).
template <typename R, typename O>
void bp_column(const float* sino, int W, int H, int nProj, R SR, const float* angles, int geom,
               int ix, int iy, O* col /* H values, stride cstride */, size_t cstride) {
  const uint32_t hw = (uint32_t)W / 2u - 1u;
  uint32_t dx = (uint32_t)ix - hw, dy = (uint32_t)iy - hw;
  if (!(dx * dx + dy * dy < hw * hw)) return;
  const R v0 = R(H - 1) / 2;
  R x = (R)ix - ((R)hw + R(0.5));
  R y = (R)iy - ((R)hw + R(0.5));
  std::vector<R> acc(H, R(0));
  for (int view = 0; view < nProj; ++view) {
    R sinB, cosB;
    sincos_of<R>(angles[view], &sinB, &cosB);
    R s = x * cosB + y * sinB;
    R t = -x * sinB + y * cosB;
    R U, py, pz, dz;
    if (geom == 0) {
      U = R(1);
      py = t * U + (R)hw + R(1);
      pz = -v0 * U + v0 + R(0.5);
      dz = U;
    } else if (geom == 1) {
      U = SR / (SR - s);
      py = t * U + (R)hw + R(1);
      pz = R(0.5);
      dz = R(1);
    } else {
      U = SR / (SR - s);
      py = t * U + (R)hw + R(1);
      pz = -v0 * U + v0 + R(0.5);
      dz = U;
    }
    const float* P = sino + (size_t)W * H * view;
    for (int iz = 0; iz < H; ++iz) {
      acc[iz] += U * U * tex2<R, float>(P, W, H, py, pz, ORC_CLAMP);
      pz += dz;
    }
  }
  for (int iz = 0; iz < H; ++iz) col[(size_t)iz * cstride] = (O)acc[iz];
}

template <typename R, typename O>
void bp_std(const float* sino, int W, int H, int nProj, R sad, const float* angles, int geom,
            O* out) {
  std::memset(out, 0, sizeof(O) * (size_t)W * W * H);
  parallel_for((int64_t)W * W, [&](int64_t item) {
    int iy = (int)(item / W), ix = (int)(item % W);
    bp_column<R, O>(sino, W, H, nProj, sad, angles, geom, ix, iy,
                    out + (size_t)ix + (size_t)W * iy, (size_t)W * W);
  }, 16);
}

// General-3D ray of pixel (ix,iy): pts rotated about z by beta
// (CUDAmex_general3D_FP.cu:69-77; CUDAmex_general3D_BP.cu:71-79).
template <typename R>
void g3d_ray(const float* pts, int W, int H, int ix, int iy, R sinB, R cosB, V3<R>* st, V3<R>* en) {
  const size_t npts = (size_t)W * H;
  const size_t p = (size_t)iy * W + ix;
  R p0 = pts[p], p1 = pts[p + npts], p2 = pts[p + 2 * npts];
  R p3 = pts[p + 3 * npts], p4 = pts[p + 4 * npts], p5 = pts[p + 5 * npts];
  st->x = p0 * cosB - p1 * sinB;
  st->y = p0 * sinB + p1 * cosB;
  st->z = p2;
  en->x = p3 * cosB - p4 * sinB;
  en->y = p3 * sinB + p4 * cosB;
  en->z = p5;
}

template <typename R, typename O>
int64_t g3d_fp(const float* vol, int side, int nSlices, const float* angles, const float* pts, int W,
               int H, int nProj, int ppv, O* out) {
  const R bx = -(R)side / R(2), by = bx, bz = -(R)nSlices / R(2);
  std::vector<int64_t> counts((size_t)nProj * H, 0);
  parallel_for((int64_t)nProj * H, [&](int64_t item) {
    int view = (int)(item / H), iy = (int)(item % H);
    R s, c;
    sincos_of<R>(angles[view], &s, &c);
    const float* vp = pts + (ppv ? (size_t)view * 6 * W * H : 0);
    int64_t n = 0;
    for (int ix = 0; ix < W; ++ix) {
      V3<R> st, en;
      g3d_ray<R>(vp, W, H, ix, iy, s, c, &st, &en);
      out[(size_t)iy * W + ix + (size_t)W * H * view] =
          (O)march<R, float>(vol, side, side, nSlices, bx, by, bz, st, en, R(1) / R(2), ORC_BORDER, &n);
    }
    counts[item] = n;
  });
  int64_t tot = 0;
  for (auto v : counts) tot += v;
  return tot;
}

// General-3D ray-driven backprojection (CUDAmex_general3D_BP.cu:61-188).
// Each thread owns a z-slab of the output and walks every ray in the
// canonical order (view, iy, ix, sample), so every voxel receives its
// contributions in that fixed order whatever the thread count.
template <typename R, typename O>
void g3d_bp(const float* sino, int W, int H, int nProj, const float* angles, const float* pts,
            int ppv, int M, int N, O* out) {
  std::memset(out, 0, sizeof(O) * (size_t)M * M * N);
  const R bx = -(R)M / R(2), by = bx, bz = -(R)N / R(2);
  const unsigned imSizeX = (unsigned)M, imSizeY = (unsigned)M, imSizeZ = (unsigned)N;
  int nt = std::max(1, std::min(nthreads(), N));
  parallel_for(nt, [&](int64_t tid) {
    int k_lo = (int)((int64_t)N * tid / nt), k_hi = (int)((int64_t)N * (tid + 1) / nt);
    std::vector<R> acc((size_t)M * M * (k_hi - k_lo), R(0));
    for (int view = 0; view < nProj; ++view) {
      R sinB, cosB;
      sincos_of<R>(angles[view], &sinB, &cosB);
      const float* vp = pts + (ppv ? (size_t)view * 6 * W * H : 0);
      const float* P = sino + (size_t)W * H * view;
      for (int iy = 0; iy < H; ++iy)
        for (int ix = 0; ix < W; ++ix) {
          V3<R> st, en;
          g3d_ray<R>(vp, W, H, ix, iy, sinB, cosB, &st, &en);
          R aX1, aX2, aY1, aY2, aZ1, aZ2;
          if ((double)std::fabs(en.x - st.x) >= 1e-5) {
            aX1 = (bx - st.x) / (en.x - st.x);
            aX2 = (-bx - st.x) / (en.x - st.x);
          } else {
            aX1 = aX2 = nan_of<R>();
          }
          if ((double)std::fabs(en.y - st.y) >= 1e-5) {
            aY1 = (by - st.y) / (en.y - st.y);
            aY2 = (-by - st.y) / (en.y - st.y);
          } else {
            aY1 = aY2 = nan_of<R>();
          }
          if ((double)std::fabs(en.z - st.z) >= 1e-5) {
            aZ1 = (bz - st.z) / (en.z - st.z);
            aZ2 = (-bz - st.z) / (en.z - st.z);
          } else {
            aZ1 = aZ2 = nan_of<R>();
          }
          R aXmin = std::fmin(aX1, aX2), aXmax = std::fmax(aX1, aX2);
          R aYmin = std::fmin(aY1, aY2), aYmax = std::fmax(aY1, aY2);
          R aZmin = std::fmin(aZ1, aZ2), aZmax = std::fmax(aZ1, aZ2);
          R tmp = std::fmax(aXmin, std::fmax(aYmin, aZmin));
          R aMin = std::fmax(R(0), tmp);
          tmp = std::fmin(aXmax, std::fmin(aYmax, aZmax));
          R aMax = std::fmin(R(1), tmp);
          V3<R> cur, e2;
          cur.x = st.x + aMin * (en.x - st.x);
          cur.y = st.y + aMin * (en.y - st.y);
          cur.z = st.z + aMin * (en.z - st.z);
          e2.x = st.x + aMax * (en.x - st.x);
          e2.y = st.y + aMax * (en.y - st.y);
          e2.z = st.z + aMax * (en.z - st.z);
          R distance = std::sqrt((e2.x - cur.x) * (e2.x - cur.x) + (e2.y - cur.y) * (e2.y - cur.y) +
                                 (e2.z - cur.z) * (e2.z - cur.z));
          R travelled = R(0);
          const R stepsize = R(1) / R(2);
          R sX = (e2.x - cur.x) / distance * stepsize;
          R sY = (e2.y - cur.y) / distance * stepsize;
          R sZ = (e2.z - cur.z) / distance * stepsize;
          const R pval = (R)P[(size_t)iy * W + ix];
          while (travelled < distance) {
            int iL = to_index(std::floor(cur.x - bx));
            int iU = to_index(std::ceil(cur.x - bx));
            int jL = to_index(std::floor(cur.y - by));
            int jU = to_index(std::ceil(cur.y - by));
            int kL = to_index(std::floor(cur.z - bz));
            int kU = to_index(std::ceil(cur.z - bz));
            bool add = !(iL < 0 || jL < 0 || kL < 0 || (unsigned)iU >= imSizeX ||
                         (unsigned)jU >= imSizeY || (unsigned)kU >= imSizeZ);
            if (add) {
              R iXX = cur.x - bx - (R)iL + R(0.5);
              R iYY = cur.y - by - (R)jL + R(0.5);
              R iZ = cur.z - bz - (R)kL + R(0.5);
              R iZZ = R(1) - iZ;
              if (kL >= k_lo && kL < k_hi)
                acc[(size_t)iL + (size_t)M * jL + (size_t)M * M * (kL - k_lo)] +=
                    pval * stepsize * iXX * iYY * iZZ;
              if (kU >= k_lo && kU < k_hi)
                acc[(size_t)iL + (size_t)M * jL + (size_t)M * M * (kU - k_lo)] +=
                    pval * stepsize * iXX * iYY * iZ;
            }
            cur.x += sX;
            cur.y += sY;
            cur.z += sZ;
            travelled += stepsize;
          }
        }
    }
    for (size_t i = 0; i < acc.size(); ++i) out[(size_t)M * M * k_lo + i] = (O)acc[i];
  });
}

// ---------------------------------------------------------------- OSC
// One CUDAmex_oscIter call (CUDAmex_oscIter.cu:362-584): for each subset,
// FP of every view (K10, :140-271) rewriting pre/post in place, dual BP
// (K11, :272-358) into num/den, then the multiplicative update (K12,
// :113-136).  num/den start at zero (K9, :94-112) and are re-zeroed by K12.
template <typename R>
inline R exp_of(R v);
template <>
inline float exp_of<float>(float v) { return expf(v); }
template <>
inline double exp_of<double>(double v) { return std::exp(v); }

template <typename R>
int osc_iter(R* mu, int W, int H, R* pre, R* post, int nProj, int numSubsets,
             const int32_t* picks, const float* angles, R SR, int geom) {
  if (geom < 0 || geom > 2) return -1;
  if (numSubsets < 1) return -3;
  const R u0 = R(W - 1) / 2, v0 = R(H - 1) / 2;
  const R bx = -(R)W / R(2), by = -(R)W / R(2), bz = -(R)H / R(2);
  const int nb = (W + 31) / 32;
  const int rows = std::min(H, 32 * nb);
  const uint32_t hw = (uint32_t)W / 2u - 1u;
  const size_t vsz = (size_t)W * W * H, psz = (size_t)W * H;
  std::vector<R> vol(vsz);  // texture snapshot of mu for this subset
  std::vector<R> tpre(psz), tpost(psz);
  std::vector<R> num(vsz, R(0)), den(vsz, R(0));
  for (int sub = 0; sub < numSubsets; ++sub) {
    int sel = picks[sub];
    for (size_t i = 0; i < vsz; ++i) vol[i] = mu[i];
    std::vector<int> views;
    // size_t projInd = subsetSel: a negative pick selects no view
    // (CUDAmex_oscIter.cu:508-510); the update still runs.
    if (sel >= 0)
      for (int v = sel; v < nProj; v += numSubsets) views.push_back(v);
    // FP + ratio terms, K10
    for (int view : views) {
      R sinB, cosB;
      sincos_of<R>(angles[view], &sinB, &cosB);
      R* pr = pre + psz * view;
      R* po = post + psz * view;
      parallel_for(rows, [&](int64_t iyl) {
        int iy = (int)iyl;
        for (int ix = 0; ix < W; ++ix) {
          V3<R> st, en;
          R step;
          std_ray<R>(geom, ix, iy, u0, v0, sinB, cosB, SR, SR, &st, &en, &step, R(1) / R(2));
          step = R(1) / R(2);
          R dval = march<R, R>(vol.data(), W, W, H, bx, by, bz, st, en, step, ORC_BORDER, nullptr);
          size_t p = (size_t)ix + (size_t)W * iy;
          R preval = pr[p], postval = po[p];
          pr[p] = (preval * exp_of<R>(-dval) - postval);
          po[p] = preval * exp_of<R>(-dval) * dval;
        }
      });
    }
    // dual BP, K11
    for (int view : views) {
      R sinB, cosB;
      sincos_of<R>(angles[view], &sinB, &cosB);
      const R* pr = pre + psz * view;
      const R* po = post + psz * view;
      for (size_t i = 0; i < psz; ++i) {
        tpre[i] = pr[i];
        tpost[i] = po[i];
      }
      parallel_for((int64_t)W * W, [&](int64_t item) {
        int iy = (int)(item / W), ix = (int)(item % W);
        uint32_t dx = (uint32_t)ix - hw, dy = (uint32_t)iy - hw;
        if (!(dx * dx + dy * dy < hw * hw)) return;
        R x = (R)ix - ((R)hw + R(0.5));
        R y = (R)iy - ((R)hw + R(0.5));
        R s = x * cosB + y * sinB;
        R t = -x * sinB + y * cosB;
        R U, UU, py, pz;
        if (geom == 0) {
          U = R(1);
          UU = R(1);
          py = t * UU + (R)hw + R(1);
          pz = -v0 * UU + v0 + R(0.5);
        } else if (geom == 1) {
          U = R(1);
          UU = SR / (SR - s);
          py = t * UU + (R)hw + R(1);
          pz = R(0.5);
        } else {
          U = SR / (SR - s);
          UU = SR / (SR - s);
          py = t * U + (R)hw + R(1);
          pz = -v0 * UU + v0 + R(0.5);
        }
        for (int iz = 0; iz < H; ++iz) {
          size_t p = (size_t)ix + (size_t)W * iy + (size_t)W * W * iz;
          R fval = UU * UU * tex2<R, R>(tpre.data(), W, H, py, pz, ORC_BORDER);
          num[p] = num[p] + (fval);
          fval = UU * UU * tex2<R, R>(tpost.data(), W, H, py, pz, ORC_BORDER);
          den[p] = den[p] + (fval);
          pz += U;
        }
      }, 16);
    }
    // update, K12
    parallel_for((int64_t)W * W, [&](int64_t item) {
      int iy = (int)(item / W), ix = (int)(item % W);
      uint32_t dx = (uint32_t)ix - hw, dy = (uint32_t)iy - hw;
      if (!(dx * dx + dy * dy < hw * hw)) return;
      for (int iz = 0; iz < H; ++iz) {
        size_t p = (size_t)ix + (size_t)W * iy + (size_t)W * W * iz;
        R cval = num[p], cval2 = den[p], fval = mu[p];
        mu[p] = std::fmax(fval + (fval * cval / cval2), R(0));
        num[p] = R(0);
        den[p] = R(0);
      }
    }, 16);
  }
  return 0;
}

// ---------------------------------------------------------------- TV
// tvGradientKernel (CUDAmex_TVmin_3D.cu:101-158): interior voxels only.
template <typename R>
void tv_gradient(const R* in, int Nx, int Ny, int Nz, R* out) {
  const size_t sx = 1, sy = (size_t)Nx, sz = (size_t)Nx * Ny;
  parallel_for(Nz, [&](int64_t izl) {
    int iz = (int)izl;
    if (!(iz >= 1 && iz < Nz - 1)) return;
    for (int iy = 1; iy < Ny - 1; ++iy)
      for (int ix = 1; ix < Nx - 1; ++ix) {
        const R eps = (R)(float)1.0e-6;
        size_t c = ix * sx + iy * sy + iz * sz;
        R xyz = in[c];
        R xmyz = in[c - sx], xpyz = in[c + sx];
        R xymz = in[c - sy], xypz = in[c + sy];
        R xyzm = in[c - sz], xyzp = in[c + sz];
        R xmypz = in[c - sx + sy], xpymz = in[c + sx - sy];
        R xymzp = in[c - sy + sz], xypzm = in[c + sy - sz];
        R xmyzp = in[c - sx + sz], xpyzm = in[c + sx - sz];
        R tval;
        tval = ((xyz - xmyz) + (xyz - xymz) + (xyz - xyzm)) /
               (eps + std::sqrt((xyz - xmyz) * (xyz - xmyz) + (xyz - xymz) * (xyz - xymz) +
                                (xyz - xyzm) * (xyz - xyzm)));
        tval -= ((xpyz - xyz)) / (eps + std::sqrt((xpyz - xyz) * (xpyz - xyz) +
                                                  (xpyz - xpymz) * (xpyz - xpymz) +
                                                  (xpyz - xpyzm) * (xpyz - xpyzm)));
        tval -= ((xypz - xyz)) / (eps + std::sqrt((xypz - xyz) * (xypz - xyz) +
                                                  (xypz - xmypz) * (xypz - xmypz) +
                                                  (xypz - xypzm) * (xypz - xypzm)));
        tval -= ((xyzp - xyz)) / (eps + std::sqrt((xyzp - xyz) * (xyzp - xyz) +
                                                  (xyzp - xmyzp) * (xyzp - xmyzp) +
                                                  (xyzp - xymzp) * (xyzp - xymzp)));
        out[c] = tval;
      }
  });
}

// Driver (CUDAmex_TVmin_3D.cu:221-234): g; |g| (summed in double here);
// x <- (-c/|g|) * g + x over every voxel, border included.
template <typename R>
void tvmin(R* x, int Nx, int Ny, int Nz, R c, int steps) {
  const size_t n = (size_t)Nx * Ny * Nz;
  std::vector<R> g(n, R(0));
  for (int it = 0; it < steps; ++it) {
    tv_gradient<R>(x, Nx, Ny, Nz, g.data());
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += (double)g[i] * (double)g[i];
    R magnitude = (R)std::sqrt(sum);
    R a = -c / magnitude;
    for (size_t i = 0; i < n; ++i) x[i] = a * g[i] + x[i];
  }
}

// ---------------------------------------------------------------- FDK filter
void fft_inplace(std::vector<double>& re, std::vector<double>& im, bool inverse) {
  const size_t n = re.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      std::swap(re[i], re[j]);
      std::swap(im[i], im[j]);
    }
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    double ang = 2.0 * M_PI / (double)len * (inverse ? 1.0 : -1.0);
    for (size_t i = 0; i < n; i += len) {
      for (size_t k = 0; k < len / 2; ++k) {
        double wr = std::cos(ang * (double)k), wi = std::sin(ang * (double)k);
        size_t a = i + k, b = i + k + len / 2;
        double xr = re[b] * wr - im[b] * wi;
        double xi = re[b] * wi + im[b] * wr;
        re[b] = re[a] - xr;
        im[b] = im[a] - xi;
        re[a] += xr;
        im[a] += xi;
      }
    }
  }
  if (inverse)
    for (size_t i = 0; i < n; ++i) {
      re[i] /= (double)n;
      im[i] /= (double)n;
    }
}

int filter_order(int len) {
  // order = max(64, 2^nextpow2(2*len))  (preFilterSinogram.m:107)
  int o = 1;
  while (o < 2 * len) o <<= 1;
  return std::max(64, o);
}

// S_function (parker_weight.m:82-89) on a single argument: 0 at or below
// -1/2, 1 at or above 1/2, 0.5*(1+sin(pi*b)) between; NaN gives 0 (no mask
// of the reference selects it).
float parker_S(float b) {
  if (b <= -0.5f) return 0.0f;
  if (std::fabs(b) < 0.5f) return 0.5f * (1.0f + std::sin((float)M_PI * b));
  if (b >= 0.5f) return 1.0f;
  return 0.0f;
}

// parker_weight.m:29-76: the Parker / Wesarg window of every view, win[i*W+u].
// The angles and all that is derived from them are single, alpha and delta
// double; an operation with one single operand takes the other rounded to
// single.  alpha runs along the detector columns u (the fan direction) with
// W entries; the reference takes size(sino,2) entries and expands with
// repmat (:32, :73), the same window when W == H (a size error otherwise).
void parker_window(int W, int nProj, int geom, double sad, const float* angles, double q,
                   float* win) {
  std::vector<double> alpha((size_t)W, 0.0);
  if (geom != 0)  // par3d: alpha = 0 (:29-30)
    for (int u = 0; u < W; ++u) alpha[u] = std::atan((-(double)W / 2 + 0.5 + u) / sad);
  bool nondecreasing = true;  // all(diff(angles) >= 0) flips alpha (:36-40)
  for (int i = 1; i < nProj; ++i)
    if (!(angles[i] - angles[i - 1] >= 0.0f)) nondecreasing = false;
  if (nondecreasing)
    for (double& a : alpha) a = -a;
  const double delta = std::fabs(alpha[W - 1] - alpha[0]) / 2;            // :42
  const float extent = std::fabs(angles[nProj - 1] - angles[0]);          // :43
  const float eps = std::max(extent - (float)(M_PI + 2 * delta), 0.0f);   // :59
  const float pi_f = (float)M_PI, d2 = (float)(2 * delta), qf = (float)q;
  for (int i = 0; i < nProj; ++i) {
    const float beta = std::fabs(angles[i] - angles[0]);  // :66
    for (int u = 0; u < W; ++u) {
      const float a2 = (float)(2 * alpha[u]);
      // b(+-alpha) = q * (2 delta - 2 (+-alpha) + epsilon)  (:92-97)
      const float bp = qf * ((float)(2 * delta - 2 * alpha[u]) + eps);
      const float bm = qf * ((float)(2 * delta - 2 * -alpha[u]) + eps);
      const float s1 = parker_S(beta / bp - 0.5f);
      const float s2 = parker_S((((beta - d2) + a2) - eps) / bp + 0.5f);
      const float s3 = parker_S(((beta - pi_f) + a2) / bm - 0.5f);
      const float s4 = parker_S((((beta - pi_f) - d2) - eps) / bm + 0.5f);
      win[(size_t)i * W + u] = 0.5f * (((s1 + s2) - s3) - s4);  // :68-71
    }
  }
}

bool parker_on(double q) { return q >= 0.0; }  // a negative (or NaN) q: no Parker step

template <typename O>
int prefilter(const float* sino, int W, int H, int nProj, int geom, double sad, const float* angles,
              int window, double d, double parker_q, O* out) {
  if (nProj < 2) return -1;
  int order = filter_order(W);
  std::vector<double> filt(order);
  orc_design_filter(window, W, d, filt.data());
  // Parker / Wesarg short-scan window (preFilterSinogram.m:36,48,60)
  std::vector<float> win;
  if (parker_on(parker_q)) {
    win.resize((size_t)W * nProj);
    parker_window(W, nProj, geom, sad, angles, parker_q, win.data());
  }
  // cosine weights (preFilterSinogram.m:38-58): Distance in double, then
  // single .* double -> single (operand rounded to single first).
  std::vector<float> wgt((size_t)W * H, 1.0f);
  if (geom == 1 || geom == 2) {
    for (int v = 0; v < H; ++v)
      for (int u = 0; u < W; ++u) {
        double U = (double)(u + 1) - (double)(W - 1) / 2.0;
        double V = (double)(v + 1) - (double)(H - 1) / 2.0;
        double D = geom == 2 ? sad / std::sqrt(sad * sad + U * U + V * V)
                             : sad / std::sqrt(sad * sad + U * U);
        wgt[(size_t)u + (size_t)W * v] = (float)D;
      }
  }
  parallel_for((int64_t)nProj * H, [&](int64_t item) {
    int view = (int)(item / H), v = (int)(item % H);
    std::vector<double> re(order, 0.0), im(order, 0.0);
    const float* P = sino + (size_t)W * H * view + (size_t)W * v;
    for (int u = 0; u < W; ++u) {
      float w = (geom == 1 || geom == 2) ? P[u] * wgt[(size_t)u + (size_t)W * v] : P[u];
      if (!win.empty()) w = w * win[(size_t)view * W + u];
      re[u] = (double)w;
    }
    fft_inplace(re, im, false);
    for (int j = 0; j < order; ++j) {
      re[j] *= filt[j];
      im[j] *= filt[j];
    }
    fft_inplace(re, im, true);
    // tmp .* (angles(i)-angles(i-1))/2 in single (preFilterSinogram.m:80-84)
    float dA = view > 0 ? angles[view] - angles[view - 1] : angles[1] - angles[0];
    O* o = out + (size_t)W * H * view + (size_t)W * v;
    for (int u = 0; u < W; ++u) {
      if (sizeof(O) == sizeof(float)) {
        float y = (float)re[u];
        o[u] = (O)((y * dA) / 2.0f);
      } else {
        o[u] = (O)(re[u] * (double)dA / 2.0);
      }
    }
  }, 8);
  return 0;
}

}  // namespace

// OSC_TV_recon.m:1-102 in R arithmetic (float: the reference's; double: the
// FP64 twin the storage forms are measured against).
template <typename R>
int osc_tv_recon(const float* pre_in, const float* post_in, int W, int H, int nProj,
                 const float* angles, float sad, int geom, int nIter, float tv, int sInit,
                 int sFinal, double p, int tv_steps, const int32_t* first_picks, R* mu_out) {
  if (nIter < 2) return -1;  // calcNumSubsets divides by (nIter-1)^p
  const size_t vsz = (size_t)W * W * H, psz = (size_t)W * H * nProj;
  std::vector<R> mu(vsz, (R)(1.0 * 1e-10));
  std::vector<R> pre(pre_in, pre_in + psz), post(post_in, post_in + psz);
  const R epsf = (R)1e-6;
  for (auto& v : pre) if ((double)v < 1e-6) v = epsf;
  for (auto& v : post) if ((double)v < 1e-6) v = epsf;
  std::vector<R> pre2 = pre, post2 = post;
  pre2[0] = (R)0;
  post2[0] = (R)0;
  for (int k = 1; k <= nIter; ++k) {
    std::vector<R> mu_old = mu;
    mu_old[0] = (R)0;
    int n = orc_calc_num_subsets(sInit, sFinal, p, nIter, k);
    if (n < 1) return -2;
    std::vector<int32_t> picks(n);
    orc_calc_subset_order(n, first_picks[k - 1], picks.data());
    int rc = osc_iter<R>(mu.data(), W, H, pre.data(), post.data(), nProj, n, picks.data(),
                         angles, sad, geom);
    if (rc) return rc;
    pre = pre2;
    post = post2;
    if (k > 1) {
      double sum = 0.0;
      for (size_t i = 0; i < vsz; ++i) {
        R dlt = mu_old[i] - mu[i];
        if (std::isnan(dlt)) dlt = (R)0;
        sum += (double)dlt * (double)dlt;
      }
      R d_a = (R)std::sqrt(sum);
      tvmin<R>(mu.data(), W, W, H, (R)tv * d_a, tv_steps);
      for (auto& v : mu)
        if (std::isnan(v) || std::isinf(v)) v = (R)0;
    }
  }
  std::memcpy(mu_out, mu.data(), sizeof(R) * vsz);
  return 0;
}


extern "C" {

void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void) { return nthreads(); }
void orc_set_texture_weight_bits(int bits) { g_weight_bits = bits > 0 ? std::min(bits, 20) : 0; }
int orc_get_texture_weight_bits(void) { return g_weight_bits; }

float orc_tex3(const float* T, int nx, int ny, int nz, float x, float y, float z, int mode) {
  return tex3<float, float>(T, nx, ny, nz, x, y, z, mode);
}
float orc_tex2(const float* T, int nx, int ny, float x, float y, int mode) {
  return tex2<float, float>(T, nx, ny, x, y, mode);
}

int64_t orc_fp(const float* vol, int W, int H, float sad, const float* angles, int nProj, int geom,
               float* out) {
  if (geom < 0 || geom > 2) return -1;
  return fp_std<float, float>(vol, W, H, sad, angles, nProj, geom, out);
}
int64_t orc_fp_f64(const float* vol, int W, int H, double sad, const float* angles, int nProj,
                   int geom, double* out) {
  if (geom < 0 || geom > 2) return -1;
  return fp_std<double, double>(vol, W, H, sad, angles, nProj, geom, out);
}
void orc_fp_f64_rays(const float* vol, int W, int H, double sad, const float* angles, int geom,
                     int nrays, const int* ix, const int* iy, const int* view, double* out) {
  const double u0 = double(W - 1) / 2, v0 = double(H - 1) / 2;
  const double bx = -(double)W / 2, bz = -(double)H / 2;
  parallel_for(nrays, [&](int64_t r) {
    double s, c;
    sincos_of<double>(angles[view[r]], &s, &c);
    V3<double> st, en;
    double step;
    std_ray<double>(geom, ix[r], iy[r], u0, v0, s, c, sad, sad, &st, &en, &step, 1.0);
    out[r] = march<double, float>(vol, W, W, H, bx, bx, bz, st, en, step, ORC_CLAMP, nullptr);
  }, 64);
}

void orc_fp_rays(const float* vol, int W, int H, float sad, const float* angles, int geom,
                 int nrays, const int* ix, const int* iy, const int* view, float* out) {
  const float u0 = float(W - 1) / 2, v0 = float(H - 1) / 2;
  const float bx = -(float)W / 2.0f, bz = -(float)H / 2.0f;
  parallel_for(nrays, [&](int64_t r) {
    float s, c;
    sincos_of<float>(angles[view[r]], &s, &c);
    V3<float> st, en;
    float step;
    std_ray<float>(geom, ix[r], iy[r], u0, v0, s, c, sad, sad, &st, &en, &step, 1.0f);
    out[r] = march<float, float>(vol, W, W, H, bx, bx, bz, st, en, step, ORC_CLAMP, nullptr);
  }, 64);
}
void orc_bp_cols(const float* sino, int W, int H, int nProj, float sad, const float* angles,
                 int geom, int ncol, const int* ix, const int* iy, float* out) {
  std::memset(out, 0, sizeof(float) * (size_t)ncol * H);
  parallel_for(ncol, [&](int64_t c) {
    bp_column<float, float>(sino, W, H, nProj, sad, angles, geom, ix[c], iy[c],
                            out + (size_t)c * H, 1);
  });
}

void orc_bp(const float* sino, int W, int H, int nProj, float sad, const float* angles, int geom,
            float* out) {
  bp_std<float, float>(sino, W, H, nProj, sad, angles, geom, out);
}
void orc_bp_f64(const float* sino, int W, int H, int nProj, double sad, const float* angles,
                int geom, double* out) {
  bp_std<double, double>(sino, W, H, nProj, sad, angles, geom, out);
}
void orc_bp_f64_cols(const float* sino, int W, int H, int nProj, double sad, const float* angles,
                     int geom, int ncol, const int* ix, const int* iy, double* out) {
  std::memset(out, 0, sizeof(double) * (size_t)ncol * H);
  parallel_for(ncol, [&](int64_t c) {
    bp_column<double, double>(sino, W, H, nProj, sad, angles, geom, ix[c], iy[c],
                              out + (size_t)c * H, 1);
  });
}

int64_t orc_g3d_fp(const float* vol, int side, int nSlices, const float* angles, const float* pts,
                   int W, int H, int nProj, int ppv, float* out) {
  return g3d_fp<float, float>(vol, side, nSlices, angles, pts, W, H, nProj, ppv, out);
}
void orc_g3d_fp_f64(const float* vol, int side, int nSlices, const float* angles, const float* pts,
                    int W, int H, int nProj, int ppv, double* out) {
  g3d_fp<double, double>(vol, side, nSlices, angles, pts, W, H, nProj, ppv, out);
}
void orc_g3d_bp(const float* sino, int W, int H, int nProj, const float* angles, const float* pts,
                int ppv, int M, int N, float* out) {
  g3d_bp<float, float>(sino, W, H, nProj, angles, pts, ppv, M, N, out);
}
void orc_g3d_bp_f64(const float* sino, int W, int H, int nProj, const float* angles,
                    const float* pts, int ppv, int M, int N, double* out) {
  g3d_bp<double, double>(sino, W, H, nProj, angles, pts, ppv, M, N, out);
}

int orc_osc_iter(float* mu, int W, int H, float* pre, float* post, int nProj, int numSubsets,
                 const int32_t* picks, const float* angles, float sad, int geom) {
  return osc_iter<float>(mu, W, H, pre, post, nProj, numSubsets, picks, angles, sad, geom);
}
int orc_osc_iter_f64(double* mu, int W, int H, double* pre, double* post, int nProj,
                     int numSubsets, const int32_t* picks, const float* angles, double sad,
                     int geom) {
  return osc_iter<double>(mu, W, H, pre, post, nProj, numSubsets, picks, angles, sad, geom);
}

void orc_tvmin(float* x, int Nx, int Ny, int Nz, float c, int steps) {
  tvmin<float>(x, Nx, Ny, Nz, c, steps);
}
void orc_tvmin_f64(double* x, int Nx, int Ny, int Nz, double c, int steps) {
  tvmin<double>(x, Nx, Ny, Nz, c, steps);
}
void orc_tv_gradient(const float* x, int Nx, int Ny, int Nz, float* g) {
  std::memset(g, 0, sizeof(float) * (size_t)Nx * Ny * Nz);
  tv_gradient<float>(x, Nx, Ny, Nz, g);
}

int orc_design_filter(int window, int len, double d, double* out) {
  const int order = filter_order(len);
  if (!out) return order;
  if (window == 5) {
    for (int i = 0; i < order; ++i) out[i] = 1.0;
    return order;
  }
  // band-limited ramp impulse response (Kak & Slaney eq. 61), :117-121
  const int half = order / 2;
  std::vector<double> h(half + 1, 0.0), imp(order, 0.0);
  h[0] = 0.25;
  for (int n = 1; n <= half; n += 2) h[n] = -1.0 / ((M_PI * n) * (M_PI * n));
  for (int m = 0; m <= half; ++m) imp[m] = h[m];
  for (int m = half + 1; m < order; ++m) imp[m] = h[order - m];
  // filt = 2*real(fft(imp)), first half+1 bins, :122-123 (direct DFT)
  std::vector<double> filt(half + 1);
  for (int j = 0; j <= half; ++j) {
    double s = 0.0;
    for (int m = 0; m < order; ++m) {
      long long jm = ((long long)j * m) % order;
      s += imp[m] * std::cos(2.0 * M_PI * (double)jm / (double)order);
    }
    filt[j] = 2.0 * s;
  }
  for (int j = 1; j <= half; ++j) {
    double w = 2.0 * M_PI * j / order;
    switch (window) {
      case 0: break;
      case 1: filt[j] *= std::sin(w / (2 * d)) / (w / (2 * d)); break;
      case 2: filt[j] *= std::cos(w / (2 * d)); break;
      case 3: filt[j] *= (.54 + .46 * std::cos(w / d)); break;
      case 4: filt[j] *= (1 + std::cos(w / d)) / 2; break;
      default: return -1;
    }
  }
  for (int j = 0; j <= half; ++j) {
    double w = 2.0 * M_PI * j / order;
    if (w > M_PI * d) filt[j] = 0.0;
  }
  for (int j = 0; j <= half; ++j) out[j] = filt[j];
  for (int j = half + 1; j < order; ++j) out[j] = filt[order - j];
  return order;
}

int orc_prefilter(const float* sino, int W, int H, int nProj, int geom, double sad,
                  const float* angles, int window, double d, double parker_q, float* out) {
  return prefilter<float>(sino, W, H, nProj, geom, sad, angles, window, d, parker_q, out);
}
int orc_prefilter_f64(const float* sino, int W, int H, int nProj, int geom, double sad,
                      const float* angles, int window, double d, double parker_q, double* out) {
  return prefilter<double>(sino, W, H, nProj, geom, sad, angles, window, d, parker_q, out);
}

int orc_parker_window(int W, int nProj, int geom, double sad, const float* angles, double q,
                      float* win) {
  if (W < 1 || nProj < 1) return -1;
  parker_window(W, nProj, geom, sad, angles, q, win);
  return 0;
}

void orc_log_transform(const float* pre, const float* post, int64_t n, float* out) {
  // sino = -log(post./pre); NaN and +-Inf -> 0  (FBP_recon.m:26-28).  A
  // negative ratio, complex in MATLAB, is NaN here and so 0.
  for (int64_t i = 0; i < n; ++i) {
    const float y = -std::log(post[i] / pre[i]);
    out[i] = std::isfinite(y) ? y : 0.0f;
  }
}

int orc_calc_num_subsets(int sInit, int sFinal, double p, int K, int k) {
  double v = (double)(sInit - sFinal) / std::pow((double)(K - 1), p) * std::pow((double)(K - k), p) +
             (double)sFinal;
  return (int)std::round(v);  // MATLAB round: half away from zero
}

void orc_calc_subset_order(int n, int first_pick, int32_t* picks) {
  // calcSubsetOrder (OSC_TV_recon.m:327-349)
  std::vector<int> nums;
  for (int i = 1; i <= n; ++i) nums.push_back(i);
  int cur = first_pick;
  for (int s = 1; s <= n; ++s) {
    std::vector<int> lower, higher;
    for (int x : nums) {
      if (x < cur) lower.push_back(x);
      if (x > cur) higher.push_back(x);
    }
    if (s > 1) {
      if (lower.size() > higher.size()) {
        cur = lower[(lower.size() + 1) / 2 - 1];
        lower.erase(std::remove(lower.begin(), lower.end(), cur), lower.end());
      } else {
        cur = higher[(higher.size() + 1) / 2 - 1];
        higher.erase(std::remove(higher.begin(), higher.end(), cur), higher.end());
      }
      nums = lower;
      nums.insert(nums.end(), higher.begin(), higher.end());
    }
    picks[s - 1] = cur - 1;
  }
}

int orc_osc_tv_recon(const float* pre_in, const float* post_in, int W, int H, int nProj,
                     const float* angles, float sad, int geom, int nIter, float tv, int sInit,
                     int sFinal, double p, int tv_steps, const int32_t* first_picks,
                     float* mu_out) {
  return osc_tv_recon<float>(pre_in, post_in, W, H, nProj, angles, sad, geom, nIter, tv, sInit,
                             sFinal, p, tv_steps, first_picks, mu_out);
}

int orc_osc_tv_recon_f64(const float* pre_in, const float* post_in, int W, int H, int nProj,
                         const float* angles, float sad, int geom, int nIter, float tv,
                         int sInit, int sFinal, double p, int tv_steps,
                         const int32_t* first_picks, double* mu_out) {
  return osc_tv_recon<double>(pre_in, post_in, W, H, nProj, angles, sad, geom, nIter, tv, sInit,
                              sFinal, p, tv_steps, first_picks, mu_out);
}

}  // extern "C"
