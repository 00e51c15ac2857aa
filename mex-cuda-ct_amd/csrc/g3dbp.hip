// g3dbp.hip -- general-3D backprojection (CUDAmex_general3D_BP.cu:61-188)
// by voxel bricks, for gfx950.
//
// The reference scatters each 0.5-voxel sample of each ray into two voxels,
// (iL, jL, kL) and (iL, jL, kL + 1), with atomicAdd.  A voxel of the helical
// 320^3 x 360 workload receives about 290 samples, under one ray per view:
// the rays of one view barely overlap, so staging one view's rays in LDS
// merges nothing, and one float atomic per sample corner costs ~90% of the
// ray-driven kernel's time.  Here the volume is cut into bricks of
// 16 x 16 x 12 voxels; a brick is one workgroup that adds, in LDS, every
// sample whose cell (iL, jL, kL) lies in it, from every ray of every view,
// and writes each of its voxels once with a plain store.  Per chunk of views:
//
//   bin    one thread per ray marches the reference's running sum once,
//          keeps the position every kCkSamples samples (checkpoints) and its
//          set-up, and marks, for its 8x8 tile of detector pixels, each brick
//          its samples can reach (one bit per (brick, tile)), walking its
//          exact line 16 samples at a time;
//   brick  for each tile marked for it (each of the four waves takes a
//          quarter of the brick's row of tile bits), each ray of the tile
//          bounds the samples the brick can own from its exact line (widened
//          by the running sum's drift), resumes the running sum at the
//          checkpoint before them and adds the samples it owns, with the
//          reference's two-corner weights.  The corner kL + 1 of the brick's top plane
//          belongs to the brick above and goes to a halo plane;
//   halo   after the last chunk each halo plane is added to its plane.
//
// The LDS sums are 64-bit fixed-point integers (ds_add_u64): on gfx950 an
// LDS float atomic add costs about 15 times a 64-bit integer one (measured:
// 2.0e11 against 3.0e12 lane additions per second over the chip), and the
// integer sums are exact in any order, so every run gives the same bits.
// The brick's z extent of 12 keeps a workgroup's LDS at 31 KB (26.6 KB of
// sums, 4 KB of work queues), five workgroups per CU; when the brick loop
// was still latency-bound, 16 x 16 x 12 ran the helical workload in 63 ms
// against 71 ms for 16 x 16 x 16 and 93 ms for 16 x 16 x 32.  Rays of more
// samples than the volume's diagonal admits cannot cross it; they go
// through the ray-driven atomic kernel (launch_g3d_bp, with the same fixed
// point), which checks every sample as the reference does, and which exits
// at once when the bin passes counted none.
#include <stdlib.h>

#include <algorithm>
#include <cmath>

#include "common.h"
#include "internal.h"
#include "ray.h"

namespace mexct {

constexpr int BX = 16, BY = 16, BZ = 12;  // brick (voxels)
constexpr int BZH = BZ + 1;               // LDS z extent: the brick plus its halo plane
constexpr int kTileSide = 8;              // tiles of 8 x 8 detector pixels: one wave

struct G3DBrickArgs {
  const float* sino;  // [W,H,nProj]
  const float* pts;
  int64_t pts_view_stride;
  const float2* sc;
  int W, H, M, N;
  int v0;  // first view of the chunk
  float bx, by, bz;
  int n_cap;  // rays of more samples are left to the atomic kernel
  int S;      // checkpoint slots per ray
  uint32_t tiles_x, tiles_per_view, ntiles;
  uint32_t nbx, nby, nbz, nbricks, tw;
  uint32_t* mask;    // [nbricks][tw] bits over the chunk's tiles
  float* ck;         // checkpoints: x, y, z planes of ck_plane floats
  size_t ck_plane;
  float* rays;       // per ray of the chunk: inc x, y, z and the sample count (bits)
  size_t ray_plane;  // planes of ray_plane entries, one per (tile, lane)
  float* out;        // [M,M,N] floats, written at the end
  int64_t* out64;    // [M,M,N] fixed-point sums
  int64_t* halo64;   // [nbz][M][M]: corners kL + 1 of each brick's top plane
  const uint32_t* pmax;  // bits of max |P| (the fixed point's scale)
  int accumulate;        // a later chunk: add to out/halo
  uint32_t* long_rays;   // count of the rays of more than n_cap samples
};

// Ray of chunk tile t, lane `lane`: detector pixel and view, and whether it
// exists.
struct TileRay {
  int ix, iy, view;
  bool ok;
};
__device__ __forceinline__ TileRay tile_ray(const G3DBrickArgs& a, uint32_t t, uint32_t lane) {
  TileRay r;
  const uint32_t vl = t / a.tiles_per_view, rem = t - vl * a.tiles_per_view;
  const uint32_t ty = rem / a.tiles_x, tx = rem - ty * a.tiles_x;
  r.ix = (int)(tx * kTileSide + (lane & 7u));
  r.iy = (int)(ty * kTileSide + (lane >> 3));
  r.view = a.v0 + (int)vl;
  r.ok = r.ix < a.W && r.iy < a.H;
  return r;
}

__device__ __forceinline__ int ray_setup(const G3DBrickArgs& a, const TileRay& r, R3& cur,
                                         R3& inc) {
  const float2 scv = a.sc[r.view];
  R3 st, en;
  g3d_ray(a.pts + (size_t)r.view * a.pts_view_stride + (size_t)r.iy * a.W + r.ix,
          (size_t)a.W * a.H, scv.x, scv.y, st, en);
  // slab clip and sample count, as CUDAmex_general3D_BP.cu:84-143
  return clip_ray(st, en, a.bx, a.by, a.bz, 0.5f, cur, inc);
}

// Bound on how far the running sum cur + k inc (float additions, k < n)
// strays from the exact line: at most half an ulp of the largest coordinate
// per step (doubled here), plus room for the rounding of the coordinates'
// offsets from the box corner and of the line's own evaluation.
__device__ __forceinline__ float ray_drift(const R3& cur, const R3& inc, int n) {
  const float fn = (float)n;
  const float big = fmaxf(fmaxf(fabsf(cur.x) + fabsf(inc.x) * fn, fabsf(cur.y) + fabsf(inc.y) * fn),
                          fabsf(cur.z) + fabsf(inc.z) * fn) + 1.f;
  return fn * big * 1.2e-7f + 1e-3f;
}

// ------------------------------------------------------------------ bin
// One wave per tile.  LDS: one bit per brick per wave.
constexpr int kBinRun = 16;  // samples per box of the brick walk
static_assert(kCkSamples % kBinRun == 0, "checkpoints fall on the first sample of a run");
__global__ __launch_bounds__(256) void g3d_bin_kernel(G3DBrickArgs a) {
  extern __shared__ uint32_t smask[];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t t = blockIdx.x * 4u + wave;
  const uint32_t mw = (a.nbricks + 31u) >> 5;
  uint32_t* m = smask + (size_t)wave * mw;
  const float fm = (float)(a.M - 1), fn1 = (float)(a.N - 1);
  for (uint32_t i = lane; i < mw; i += 64u) m[i] = 0u;
  __syncthreads();
  if (t < a.ntiles) {
    const TileRay r = tile_ray(a, t, lane);
    if (r.ok) {
      R3 cur, inc;
      const int n = ray_setup(a, r, cur, inc);
      // the brick pass reads the ray's set-up from here (its first
      // position is checkpoint 0)
      const size_t ri = (size_t)t * 64u + lane;
      a.rays[ri] = inc.x;
      a.rays[a.ray_plane + ri] = inc.y;
      a.rays[2 * a.ray_plane + ri] = inc.z;
      a.rays[3 * a.ray_plane + ri] = __int_as_float(n);
      if (n > a.n_cap) atomicAdd(a.long_rays, 1u);  // left to the atomic kernel
      if (n > 0 && n <= a.n_cap) {
        const size_t cbase = (size_t)t * a.S * 64u + lane;
        // The bricks come from the exact line X0 + k s, a run of kBinRun
        // samples at a time: every brick that the run's box (the line between
        // the run's first and last sample, widened by the running sum's
        // drift) covers within the volume.  The running sum's samples lie in
        // that box, so every brick owning one of them is marked; a brick
        // marked without one costs the brick pass a range test and nothing
        // more.  The reference's index test (CUDAmex_general3D_BP.cu:148-156),
        // floor(f) >= 0 and ceil(f) < size on each axis, is 0 <= f <= size - 1
        // on the exact coordinates, so the box is clipped to that range.
        const float X0 = cur.x - a.bx, Y0 = cur.y - a.by, Z0 = cur.z - a.bz;
        const float d = ray_drift(cur, inc, n);
        int px0 = -1, px1 = -1, py0 = -1, py1 = -1, pz0 = -1, pz1 = -1;  // the previous run's bricks
        for (int k0 = 0; k0 < n; k0 += kBinRun) {
          if ((k0 & (kCkSamples - 1)) == 0) {
            const size_t ci = cbase + (size_t)(k0 / kCkSamples) * 64u;
            a.ck[ci] = cur.x;
            a.ck[a.ck_plane + ci] = cur.y;
            a.ck[2 * a.ck_plane + ci] = cur.z;
          }
          const float ka = (float)k0, kb = (float)min(k0 + kBinRun - 1, n - 1);
          const float xa = __builtin_fmaf(ka, inc.x, X0), xb = __builtin_fmaf(kb, inc.x, X0);
          const float ya = __builtin_fmaf(ka, inc.y, Y0), yb = __builtin_fmaf(kb, inc.y, Y0);
          const float za = __builtin_fmaf(ka, inc.z, Z0), zb = __builtin_fmaf(kb, inc.z, Z0);
          const float lx = fminf(xa, xb) - d, hx = fmaxf(xa, xb) + d;
          const float ly = fminf(ya, yb) - d, hy = fmaxf(ya, yb) + d;
          const float lz = fminf(za, zb) - d, hz = fmaxf(za, zb) + d;
          // written so that a NaN skips the run
          if (hx >= 0.f && lx <= fm && hy >= 0.f && ly <= fm && hz >= 0.f && lz <= fn1) {
            const int bx0 = cvt_flr(fmaxf(lx, 0.f)) / BX, bx1 = cvt_flr(fminf(hx, fm)) / BX;
            const int by0 = cvt_flr(fmaxf(ly, 0.f)) / BY, by1 = cvt_flr(fminf(hy, fm)) / BY;
            const int bz0 = cvt_flr(fmaxf(lz, 0.f)) / BZ, bz1 = cvt_flr(fminf(hz, fn1)) / BZ;
            // consecutive runs mostly cover the same bricks
            if (bx0 != px0 || bx1 != px1 || by0 != py0 || by1 != py1 || bz0 != pz0 || bz1 != pz1) {
              px0 = bx0;
              px1 = bx1;
              py0 = by0;
              py1 = by1;
              pz0 = bz0;
              pz1 = bz1;
              for (int bz = bz0; bz <= bz1; ++bz)
                for (int by = by0; by <= by1; ++by)
                  for (int bx = bx0; bx <= bx1; ++bx) {
                    const uint32_t b = (uint32_t)bx + a.nbx * ((uint32_t)by + a.nby * (uint32_t)bz);
                    atomicOr(&m[b >> 5], 1u << (b & 31u));
                  }
            }
          }
#pragma unroll
          for (int u = 0; u < kBinRun; ++u) {  // the running sum over the run
            cur.x = cur.x + inc.x;
            cur.y = cur.y + inc.y;
            cur.z = cur.z + inc.z;
          }
        }
      }
    }
  }
  __syncthreads();
  if (t < a.ntiles) {
    const uint32_t bit = 1u << (t & 31u);
    const size_t col = t >> 5;
    for (uint32_t i = lane; i < mw; i += 64u) {
      uint32_t w = m[i];
      while (w) {
        const uint32_t b = i * 32u + (uint32_t)__builtin_ctz(w);
        w &= w - 1u;
        atomicOr(&a.mask[(size_t)b * a.tw + col], bit);
      }
    }
  }
}

// ------------------------------------------------------------------ brick
// Range [klo, khi] of the samples k of the line X0 + k s (one axis) that
// can fall in [lo, hi) once the running sum's drift d is allowed for (with
// -d: the samples that fall in it for sure).  rs is 1/s to within an ulp;
// the caller moves the range's ends by a sample to cover its rounding.
__device__ __forceinline__ void axis_range(float X0, float s, float rs, float lo, float hi,
                                           float d, float& klo, float& khi) {
  if (s > 0.f) {
    klo = fmaxf(klo, (lo - d - X0) * rs);
    khi = fminf(khi, (hi + d - X0) * rs);
  } else if (s < 0.f) {
    klo = fmaxf(klo, (hi + d - X0) * rs);
    khi = fminf(khi, (lo - d - X0) * rs);
  } else if (!(X0 >= lo - d && X0 < hi + d)) {
    khi = -1.f;  // the axis never moves and stays outside
  }
}

// Exclusive float bound of a brick axis [lo, lo + B) intersected with the
// reference's index test on that axis (ceil(f) < n, i.e. f <= n - 1): the
// float just above n - 1 when the brick reaches past it.
__device__ __forceinline__ float axis_hi(int lo, int B, int n) {
  return lo + B <= n - 1 ? (float)(lo + B) : __uint_as_float(__float_as_uint((float)(n - 1)) + 1u);
}

// min(v, 1) as one v_min_u32 (the compiler would make it a compare and a
// select)
__device__ __forceinline__ uint32_t min1_u32(uint32_t v) {
  uint32_t r;
  asm("v_min_u32 %0, %1, 1" : "=v"(r) : "v"(v));
  return r;
}

// (v << 3) + c as one v_lshl_add_u32 (the compiler would subtract first and
// shift after); c is wave-uniform and stays in its SGPR
__device__ __forceinline__ int lshl3_add(int v, int c) {
  int r;
  asm("v_lshl_add_u32 %0, %1, 3, %2" : "=v"(r) : "v"(v), "s"(c));
  return r;
}

// One sample of a work item: its two corners added to the brick's LDS sums
// when its cell lies in the brick (CHECK: tested here; otherwise known), and
// the running sum advanced.  Each sample goes to LDS as it is taken: integer
// LDS adds are cheap, and summing the samples of one cell in registers first
// cost more instructions than it saved (50.2 against 45.5 ms, helical 320^3).
template <bool CHECK>
__device__ __forceinline__ void brick_sample(const G3DBrickArgs& a, unsigned long long* lds, R3& cur,
                                             const R3& inc, float w0, int base8, float lox, float loy,
                                             float loz, float hx, float hy, float hz) {
  const float fx = cur.x - a.bx, fy = cur.y - a.by, fz = cur.z - a.bz;
  const bool in = !CHECK || (fx >= lox && fx < hx && fy >= loy && fy < hy && fz >= loz && fz < hz);
  // CUDAmex_general3D_BP.cu:158-170: f - floor(f) + 0.5 per axis (exact, as
  // v_fract); kL gets 1 - iZ, kU = ceil gets iZ (kL itself when fz is a
  // whole number).  w0 carries the fixed point's scale.
  const float rz = __builtin_amdgcn_fractf(fz);
  const float iXX = __builtin_amdgcn_fractf(fx) + 0.5f;
  const float iYY = __builtin_amdgcn_fractf(fy) + 0.5f;
  const float iZ = rz + 0.5f;
  const float iZZ = 1 - iZ;
  const float wxy = w0 * iXX * iYY;
  const float lo = wxy * iZZ;
  const float hi = wxy * iZ;
  // byte offset of cell (iL, jL, kL) in the brick's LDS, and 8 more for kU
  // unless rz is zero (the bits of rz >= 0 are zero only then)
  const uint32_t at = (uint32_t)mad_i24(cvt_flr(fx), BY * BZH * 8,
                                        mad_i24(cvt_flr(fy), BZH * 8, lshl3_add(cvt_flr(fz), -base8)));
  const uint32_t at_hi = at + (min1_u32(__float_as_uint(rz)) << 3);
  if (in) {
    atomicAdd((unsigned long long*)((char*)lds + at), g3d_fixed(lo));
    atomicAdd((unsigned long long*)((char*)lds + at_hi), g3d_fixed(hi));
  }
  cur.x = cur.x + inc.x;
  cur.y = cur.y + inc.y;
  cur.z = cur.z + inc.z;
}

constexpr uint32_t kQueue = 128;  // work items per wave (a power of two above 127)

// Work items [head, head + cnt) of a wave's queue, one per lane: the samples
// [k0, k1] of the item's ray, resumed from the checkpoint before k0, added to
// the brick's LDS sums when their cell lies in the brick.
__device__ __forceinline__ void brick_items(const G3DBrickArgs& a, unsigned long long* lds,
                                            const uint2* queue, uint32_t head, uint32_t cnt,
                                            uint32_t lane, float lox, float loy, float loz,
                                            float hx, float hy, float hz, uint32_t base,
                                            float scale) {
  // the queue's entries were written by other lanes of this wave
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < cnt) {
    const uint2 it = queue[(head + lane) & (kQueue - 1u)];
    const uint32_t t = it.x >> 6, l = it.x & 63u;
    const size_t ri = it.x;
    const int k0 = (int)(it.y & 0xffffu), k1 = (int)(it.y >> 16);
    R3 cur, inc;
    inc.x = a.rays[ri];
    inc.y = a.rays[a.ray_plane + ri];
    inc.z = a.rays[2 * a.ray_plane + ri];
    // resume the running sum at the checkpoint before k0
    const int c = k0 / kCkSamples;
    const size_t ci = ((size_t)t * a.S + (size_t)c) * 64u + l;
    cur.x = a.ck[ci];
    cur.y = a.ck[a.ck_plane + ci];
    cur.z = a.ck[2 * a.ck_plane + ci];
    int kr = c * kCkSamples;
    for (; kr + 4 <= k0; kr += 4) {  // four additions per trip: fewer loop instructions
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        cur.x = cur.x + inc.x;
        cur.y = cur.y + inc.y;
        cur.z = cur.z + inc.z;
      }
    }
    for (; kr < k0; ++kr) {
      cur.x = cur.x + inc.x;
      cur.y = cur.y + inc.y;
      cur.z = cur.z + inc.z;
    }
    const TileRay r = tile_ray(a, t, l);
    const float pval = a.sino[(size_t)r.view * a.W * a.H + (size_t)r.iy * a.W + r.ix];
    // P * stepsize, times the fixed point's scale: both powers of two, so
    // every weight below equals the reference's float weight times the scale
    // exactly (for P above the float's subnormal range)
    const float w0 = pval * 0.5f * scale;
    const int base8 = (int)base * 8;
    // The samples [m0, m1] lie inside the brick for sure: the line from the
    // running sum at k0, X + (k - k0) s, stays at least the drift bound
    // inside every face of the brick's index range there (one sample less on
    // each side for the rounding of this bound).  They skip the inside test.
    int m0 = k1 + 1, m1 = k1;
    {
      const float fl = (float)(k1 - k0);
      const float dr = ray_drift(cur, inc, k1 - k0 + 1);
      float ilo = 0.f, ihi = fl;
      axis_range(cur.x - a.bx, inc.x, __builtin_amdgcn_rcpf(inc.x), lox, hx, -dr, ilo, ihi);
      axis_range(cur.y - a.by, inc.y, __builtin_amdgcn_rcpf(inc.y), loy, hy, -dr, ilo, ihi);
      axis_range(cur.z - a.bz, inc.z, __builtin_amdgcn_rcpf(inc.z), loz, hz, -dr, ilo, ihi);
      if (ilo <= ihi) {
        m0 = max(k0, k0 + (int)ceilf(ilo) + 1);
        m1 = min(k1, k0 + (int)floorf(ihi) - 1);
        if (m1 < m0) {
          m0 = k1 + 1;
          m1 = k1;
        }
      }
    }
    int k = k0;
    for (; k < m0; ++k) brick_sample<true>(a, lds, cur, inc, w0, base8, lox, loy, loz, hx, hy, hz);
    for (; k <= m1; ++k) brick_sample<false>(a, lds, cur, inc, w0, base8, lox, loy, loz, hx, hy, hz);
    for (; k <= k1; ++k) brick_sample<true>(a, lds, cur, inc, w0, base8, lox, loy, loz, hx, hy, hz);
  }
  __builtin_amdgcn_wave_barrier();
}

// The marked tiles of one wave's run of a brick's tile-bit row [q0, qe), in
// order: 64 words per vector load, the nonzero words and their bits walked
// one by one.  All of it is wave-uniform but the lane's word mv.
struct TileScan {
  uint32_t q0, qe;  // q0: the word at lane 0 of the current load (starts 64 before)
  uint32_t mv;      // this lane's word of the current load
  uint32_t w, q;    // bits left of word q
  uint64_t nzw;     // lanes whose word has bits left to walk
  __device__ __forceinline__ bool next(const uint32_t* row, uint32_t lane, uint32_t& t) {
    for (;;) {
      if (w) {
        t = q * 32u + (uint32_t)__builtin_ctz(w);
        w &= w - 1u;
        return true;
      }
      if (nzw) {
        const uint32_t l = (uint32_t)__builtin_ctzll(nzw);
        nzw &= nzw - 1ull;
        w = (uint32_t)__builtin_amdgcn_readlane((int)mv, (int)l);
        q = q0 + l;
        continue;
      }
      q0 += 64u;
      if (q0 >= qe) return false;
      mv = q0 + lane < qe ? row[q0 + lane] : 0u;
      nzw = __builtin_amdgcn_ballot_w64(mv != 0u);
    }
  }
};

// A lane's ray of tile t as the bin pass set it up: increments, sample count
// and first position (checkpoint 0).  Every entry exists for every (tile,
// lane) of the chunk, so the loads need no guard; the values are used only
// where ok and 0 < n <= n_cap, the rays whose set-up and checkpoints the bin
// pass wrote.
struct TileSetup {
  R3 cur, inc;
  int n;
  bool ok;
};
__device__ __forceinline__ TileSetup tile_setup(const G3DBrickArgs& a, uint32_t t, uint32_t lane) {
  TileSetup d;
  d.ok = tile_ray(a, t, lane).ok;
  const size_t ri = (size_t)t * 64u + lane;
  d.n = __float_as_int(a.rays[3 * a.ray_plane + ri]);
  d.inc.x = a.rays[ri];
  d.inc.y = a.rays[a.ray_plane + ri];
  d.inc.z = a.rays[2 * a.ray_plane + ri];
  const size_t c0 = (size_t)t * a.S * 64u + lane;
  d.cur.x = a.ck[c0];
  d.cur.y = a.ck[a.ck_plane + c0];
  d.cur.z = a.ck[2 * a.ck_plane + c0];
  return d;
}

__global__ __launch_bounds__(256) void g3d_brick_kernel(G3DBrickArgs a) {
  typedef unsigned long long Acc;
  __shared__ Acc lds[BX * BY * BZH];
  __shared__ uint2 wq[4][kQueue];
  const uint32_t brick = xcd_remap(blockIdx.x, gridDim.x);
  if (brick >= a.nbricks) return;
  const uint32_t bxi = brick % a.nbx, byi = (brick / a.nbx) % a.nby, bzi = brick / (a.nbx * a.nby);
  for (uint32_t i = threadIdx.x; i < (uint32_t)(BX * BY * BZH); i += 256u) lds[i] = (Acc)0;
  __syncthreads();
  const float scale = g3d_fixed_scale(a.pmax);
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const float lox = (float)(bxi * BX), loy = (float)(byi * BY), loz = (float)(bzi * BZ);
  // a sample is added here when its cell (iL, jL, kL) passes the reference's
  // index test (floor >= 0 and ceil < size on each axis) and lies in the
  // brick: on the exact coordinates f these are lo <= f < hi per axis
  const float hx = axis_hi((int)(bxi * BX), BX, a.M), hy = axis_hi((int)(byi * BY), BY, a.M),
              hz = axis_hi((int)(bzi * BZ), BZ, a.N);
  // LDS index of cell (i, j, k): (i - lox) * BY * BZH + (j - loy) * BZH + (k - loz)
  const uint32_t base = (bxi * BX * BY + byi * BY) * BZH + bzi * BZ;
  const uint32_t* row = a.mask + (size_t)brick * a.tw;
  // The rays of a marked tile that cross the brick (about half of them) are
  // queued, per wave, as (ray, k0, k1) work items in LDS, and the sample
  // loop runs on full waves of 64 items: ray = tile * 64 + lane, k0 | k1 << 16.
  uint2* queue = wq[wave];
  uint32_t head = 0, tail = 0;  // wave-uniform; at most 127 items wait
  // wave w scans the w-th quarter of the brick's row of tile bits (a run of
  // views), 64 words per vector load; the words with bits set are walked
  // one by one.  The set-up of the next marked tile is loaded while the
  // current one is ranged and queued.
  const uint32_t qw = (a.tw + 3u) / 4u, qb = wave * qw, qe = min(a.tw, qb + qw);
  TileScan sc{qb - 64u, qe, 0u, 0u, 0u, 0ull};
  uint32_t t;
  bool have = sc.next(row, lane, t);
  TileSetup cs;
  if (have) cs = tile_setup(a, t, lane);
  while (have) {
    const TileSetup d = cs;
    const uint32_t tc = t;
    have = sc.next(row, lane, t);
    if (have) cs = tile_setup(a, t, lane);  // in flight while tile tc is ranged
    bool active = false;
    int k0 = 0, k1 = -1;
    if (d.ok && d.n > 0 && d.n <= a.n_cap) {
      // samples the brick can own: the exact line X0 + k s inside the brick
      // on every axis, widened by the running sum's drift
      const float X0 = d.cur.x - a.bx, Y0 = d.cur.y - a.by, Z0 = d.cur.z - a.bz;
      const float dr = ray_drift(d.cur, d.inc, d.n);
      float klo = 0.f, khi = (float)(d.n - 1);
      axis_range(X0, d.inc.x, __builtin_amdgcn_rcpf(d.inc.x), lox, lox + (float)BX, dr, klo, khi);
      axis_range(Y0, d.inc.y, __builtin_amdgcn_rcpf(d.inc.y), loy, loy + (float)BY, dr, klo, khi);
      axis_range(Z0, d.inc.z, __builtin_amdgcn_rcpf(d.inc.z), loz, loz + (float)BZ, dr, klo, khi);
      if (klo <= khi) {
        active = true;
        k0 = max(0, (int)floorf(klo) - 1);
        k1 = min(d.n - 1, (int)ceilf(khi) + 1);
      }
    }
    const uint64_t m = __builtin_amdgcn_ballot_w64(active);
    if (active) {
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
          (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      queue[(tail + rank) & (kQueue - 1u)] =
          make_uint2(tc * 64u + lane, (uint32_t)k0 | ((uint32_t)k1 << 16));
    }
    tail += (uint32_t)__builtin_popcountll(m);
    if (tail - head >= 64u) {
      brick_items(a, lds, queue, head, 64u, lane, lox, loy, loz, hx, hy, hz, base, scale);
      head += 64u;
    }
  }
  while (tail != head) {
    const uint32_t cnt = min(64u, tail - head);
    brick_items(a, lds, queue, head, cnt, lane, lox, loy, loz, hx, hy, hz, base, scale);
    head += cnt;
  }
  __syncthreads();
  // the brick's voxels, x fastest across the threads
  const size_t MM = (size_t)a.M * a.M;
  for (uint32_t e = threadIdx.x; e < (uint32_t)(BX * BY * BZ); e += 256u) {
    const int li = (int)(e % BX), lj = (int)((e / BX) % BY), lk = (int)(e / (BX * BY));
    const int i = (int)bxi * BX + li, j = (int)byi * BY + lj, k = (int)bzi * BZ + lk;
    if (i >= a.M || j >= a.M || k >= a.N) continue;
    const int64_t v = (int64_t)lds[(li * BY + lj) * BZH + lk];
    const size_t o = (size_t)i + (size_t)a.M * j + MM * k;
    a.out64[o] = a.accumulate ? a.out64[o] + v : v;
  }
  // the halo plane: voxel plane k = (bzi + 1) * BZ, owned by the brick above
  if ((int)(bzi + 1) * BZ < a.N) {
    for (uint32_t e = threadIdx.x; e < (uint32_t)(BX * BY); e += 256u) {
      const int li = (int)(e % BX), lj = (int)(e / BX);
      const int i = (int)bxi * BX + li, j = (int)byi * BY + lj;
      if (i >= a.M || j >= a.M) continue;
      const int64_t v = (int64_t)lds[(li * BY + lj) * BZH + BZ];
      const size_t o = (size_t)bzi * MM + (size_t)j * a.M + i;
      a.halo64[o] = a.accumulate ? a.halo64[o] + v : v;
    }
  }
}

// ------------------------------------------------------------------ halo
// Adds halo plane bz to voxel plane (bz + 1) * BZ.
__global__ __launch_bounds__(256) void g3d_halo_kernel(G3DBrickArgs a) {
  const size_t MM = (size_t)a.M * a.M;
  const size_t nh = (size_t)(a.nbz - 1) * MM;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < nh; e += (size_t)gridDim.x * 256) {
    const size_t bz = e / MM, ij = e - bz * MM;
    a.out64[ij + MM * (size_t)((bz + 1) * BZ)] += a.halo64[e];
  }
}

// The fixed-point sums as floats.
__global__ __launch_bounds__(256) void g3d_fixed_to_float_kernel(G3DBrickArgs a, size_t n) {
  const double inv = 1.0 / (double)g3d_fixed_scale(a.pmax);
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256)
    a.out[e] = (float)((double)a.out64[e] * inv);
}

// ------------------------------------------------------------------ host
static unsigned grid_of(size_t n) {
  return (unsigned)std::min<size_t>((n + 255) / 256, 65536);
}

int g3d_brick_n_cap(int M, int N) {
  // the longest chord of the [M,M,N] box, in 0.5 steps, plus rounding
  const double diag = sqrt(2.0 * (double)M * M + (double)N * N);
  return (int)ceil(diag / 0.5) + 4;
}

G3DBrickPlan g3d_brick_plan(int W, int H, int M, int N, int nviews) {
  G3DBrickPlan p;
  p.tiles_x = (uint32_t)((W + kTileSide - 1) / kTileSide);
  p.tiles_per_view = p.tiles_x * (uint32_t)((H + kTileSide - 1) / kTileSide);
  p.nbx = (uint32_t)((M + BX - 1) / BX);
  p.nby = (uint32_t)((M + BY - 1) / BY);
  p.nbz = (uint32_t)((N + BZ - 1) / BZ);
  p.nbricks = p.nbx * p.nby * p.nbz;
  p.n_cap = g3d_brick_n_cap(M, N);
  p.S = (p.n_cap + kCkSamples - 1) / kCkSamples;
  // views per chunk: checkpoints and ray set-ups within 4 GiB and the tile
  // bits within 256 MiB (fewer chunks mean fewer brick passes, each with
  // its own load tail and output read-modify-write)
  const size_t ck_per_view = (size_t)(3 * p.S + 4) * 4 * 64u * p.tiles_per_view;
  size_t vc = std::max<size_t>(1, ((size_t)4 << 30) / ck_per_view);
  const size_t mask_per_view_bits = (size_t)p.nbricks * p.tiles_per_view;
  vc = std::min(vc, std::max<size_t>(1, ((size_t)256 << 20) * 8 / std::max<size_t>(mask_per_view_bits, 1)));
  // MEXCT_G3D_CHUNK_VIEWS caps the views per chunk (tests run several
  // chunks on small workloads with it)
  if (const char* cap = getenv("MEXCT_G3D_CHUNK_VIEWS")) {
    const long v = strtol(cap, nullptr, 10);
    if (v > 0) vc = std::min(vc, (size_t)v);
  }
  vc = std::min(vc, (size_t)std::max(nviews, 1));
  p.views_per_chunk = (int)vc;
  p.tw = (uint32_t)((vc * p.tiles_per_view + 31) / 32);
  p.ck_plane = vc * p.tiles_per_view * (size_t)p.S * 64u;
  p.ray_plane = vc * p.tiles_per_view * 64u;
  p.ck_bytes = (3 * p.ck_plane + 4 * p.ray_plane) * sizeof(float);
  p.mask_bytes = (size_t)p.nbricks * p.tw * sizeof(uint32_t);
  p.halo_elems = (size_t)p.nbz * M * M;
  // the bin kernel's per-wave brick bits fit the LDS
  p.ok = (size_t)4 * ((p.nbricks + 31) / 32) * 4 <= (size_t)64 << 10 &&
         p.n_cap < 65536;  // the brick pass packs sample indices in 16 bits
  return p;
}

hipError_t launch_g3d_bp_bricks(const G3DBrickLaunch& L, hipStream_t st) {
  const G3DBrickPlan& p = L.plan;
  G3DBrickArgs a{};
  a.sino = L.sino;
  a.pts = L.pts;
  a.pts_view_stride = L.pts_view_stride;
  a.sc = L.sc;
  a.W = L.W;
  a.H = L.H;
  a.M = L.M;
  a.N = L.N;
  a.bx = (float)(-(double)(float)L.M / 2.0);
  a.by = a.bx;
  a.bz = (float)(-(double)(float)L.N / 2.0);
  a.n_cap = p.n_cap;
  a.S = p.S;
  a.tiles_x = p.tiles_x;
  a.tiles_per_view = p.tiles_per_view;
  a.nbx = p.nbx;
  a.nby = p.nby;
  a.nbz = p.nbz;
  a.nbricks = p.nbricks;
  a.tw = p.tw;
  a.mask = L.mask;
  a.ck = L.ck;
  a.ck_plane = p.ck_plane;
  a.rays = L.ck + 3 * p.ck_plane;
  a.ray_plane = p.ray_plane;
  a.out = L.out;
  a.out64 = L.out64;
  a.halo64 = L.halo64;
  a.pmax = L.pmax;
  a.long_rays = L.long_rays;
  const size_t MMN = (size_t)L.M * L.M * L.N;
  if (L.nviews > 0) {
    const hipError_t e = launch_fill((float*)L.long_rays, 0.f, 1, st);
    if (e != hipSuccess) return e;
  }
  const size_t lds_bin = (size_t)4 * ((p.nbricks + 31) / 32) * sizeof(uint32_t);
  for (int v0 = 0; v0 < L.nviews; v0 += p.views_per_chunk) {
    const int nv = std::min(p.views_per_chunk, L.nviews - v0);
    a.v0 = v0;
    a.ntiles = (uint32_t)nv * p.tiles_per_view;
    a.accumulate = v0 > 0;
    hipError_t e = launch_fill((float*)L.mask, 0.f, p.mask_bytes / 4, st);
    if (e != hipSuccess) return e;
    g3d_bin_kernel<<<(a.ntiles + 3) / 4, 256, lds_bin, st>>>(a);
    const unsigned gb = (p.nbricks + 7u) & ~7u;
    g3d_brick_kernel<<<gb, 256, 0, st>>>(a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (L.nviews == 0) {  // no chunk ran: the volume is zero
    const hipError_t e = launch_fill((float*)L.out64, 0.f, 2 * MMN, st);
    if (e != hipSuccess) return e;
  }
  if (p.nbz > 1 && L.nviews > 0) {
    const unsigned g = grid_of((size_t)(p.nbz - 1) * L.M * L.M);
    g3d_halo_kernel<<<g, 256, 0, st>>>(a);
  }
  if (L.nviews > 0) {  // rays too long to cross the volume, checked sample by sample
    G3DBPLaunch f;
    f.sino = L.sino;
    f.pts = L.pts;
    f.pts_view_stride = L.pts_view_stride;
    f.sc = L.sc;
    f.out = L.out;
    f.W = L.W;
    f.H = L.H;
    f.nviews = L.nviews;
    f.M = L.M;
    f.N = L.N;
    f.n_min = p.n_cap;
    f.xfast = true;
    f.out64 = L.out64;
    f.pmax = L.pmax;
    f.gate = L.long_rays;  // nothing to do unless the bin passes counted a long ray
    const hipError_t e = launch_g3d_bp(f, st);
    if (e != hipSuccess) return e;
  }
  g3d_fixed_to_float_kernel<<<grid_of(MMN), 256, 0, st>>>(a, MMN);
  return hipGetLastError();
}

}  // namespace mexct
