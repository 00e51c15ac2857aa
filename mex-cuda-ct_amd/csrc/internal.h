// internal.h -- host-side launch interface between the C ABI (api.hip) and
// the kernel translation units.  Every launcher enqueues on `st` and
// returns the hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct mexct_ctx;
struct mexct_opts;

namespace mexct {

// ---- storage conversion + halo padding (pad.hip)
// 3D: in [nx,ny,nz] (dtype in_dt) -> out [nx+4, ny+4, nz+4] of storage st_dt.
hipError_t launch_pad3(const void* in, int in_dt, void* out, int st_dt, int nx, int ny, int nz,
                       int mode, hipStream_t st);
// z-fastest haloed copy: element (x,y,z) of the x-fastest input at
// in[x + in_sy*y + in_sz*z] -> out[(x+2)*(ny+4)*(nz+4) + (y+2)*(nz+4) + z+2].
hipError_t launch_pad3z(const void* in, int in_dt, int64_t in_sy, int64_t in_sz, void* out,
                        int st_dt, int nx, int ny, int nz, int mode, hipStream_t st,
                        const uint32_t* scale_max = nullptr);
// z-fastest x-pair ("dual") copy: element (x,y,z), x in [-2,n], y, z in
// [-2, n+1] / [-2, nz+1], at ((x+2)*(ny+4) + y+2)*(nz+4) + z+2 holds
// {V(x,y,z), V(x+1,y,z)} of the haloed volume V.
hipError_t launch_pad3d(const void* in, int in_dt, int64_t in_sy, int64_t in_sz, void* out,
                        int st_dt, int nx, int ny, int nz, int mode, hipStream_t st,
                        const uint32_t* scale_max = nullptr);
// z-fastest corner-quad copy: element (x,y,z), x,y in [-2,n], z in [-2,nz+1], at
// ((x+2)*(ny+3) + y+2)*(nz+4) + z+2 holds {V(x,y,z), V(x,y+1,z), V(x+1,y,z),
// V(x+1,y+1,z)} of the haloed volume V (same input addressing as launch_pad3z);
// f16 holds the differences {V, Dy, Dx, Dxy} of those four instead (pad.hip).
hipError_t launch_pad3q(const void* in, int in_dt, int64_t in_sy, int64_t in_sz, void* out,
                        int st_dt, int nx, int ny, int nz, int mode, hipStream_t st,
                        const uint32_t* scale_max = nullptr);
// 2D stack, row-grouped: in [w,h,n] -> out [rows, w+4, h+2*zh-1, n] where
// element (px, py) holds the haloed values at rows py-zh .. py-zh+rows-1 of
// column px (rows = 2, or 4 for 2-byte storage; zh >= 2 rows of halo).
hipError_t launch_pad2(const void* in, int in_dt, void* out, int st_dt, int w, int h, int n,
                       int mode, hipStream_t st, const uint32_t* scale_max = nullptr,
                       int rows = 2, int zh = 2);
// scale_max (pad3z, pad3q, pad2): when non-null, the bits of max|in|; the
// copy holds in * 2^-pow2_exp(scale_max) (see common.h).
// 3D interior copy between a padded f32 volume and a plain one.
hipError_t launch_unpad3(const float* padded, float* out, int nx, int ny, int nz, hipStream_t st);
hipError_t launch_fill(float* p, float v, size_t n, hipStream_t st);

// ---- forward projection (fp.hip)
struct FPLaunch {
  const void* vol;  // padded storage volume [NX+4, NY+4, NZ+4]
  int st_dt;        // storage dtype
  float* out;       // [W, H, nProj]
  const float2* sc; // device (sin, cos) per view
  int geom;         // GEOM_PAR/FAN/CONE/G3D
  int mode;         // ADDR_CLAMP / ADDR_BORDER
  int W, H, nviews;
  int NX, NY, NZ;
  float sad;
  float step;
  const float* pts;          // g3d rays (device)
  int64_t pts_view_stride;   // elements between per-view pts slices (0 = shared)
};
hipError_t launch_fp(const FPLaunch& a, hipStream_t st);

// standard-geometry FP on a z-fastest haloed volume (launch_pad3z layout)
// z-fastest volume layouts of the standard-geometry FP (see fpz.hip)
// DUAL64 / QUAD64: the DUAL / QUAD copies past 4 GiB, read with 64-bit
// addresses instead of 32-bit buffer offsets (every index clamped)
enum { LAYOUT_PLAIN = 0, LAYOUT_QUAD = 1, LAYOUT_DUAL = 2, LAYOUT_DUAL64 = 3, LAYOUT_QUAD64 = 4 };
constexpr int base_layout(int l) {
  return l == LAYOUT_DUAL64 ? LAYOUT_DUAL : l == LAYOUT_QUAD64 ? LAYOUT_QUAD : l;
}
constexpr bool wide64(int l) { return l == LAYOUT_DUAL64 || l == LAYOUT_QUAD64; }

struct FPZLaunch {
  const void* vol;  // launch_pad3z (PLAIN), launch_pad3d (DUAL) or launch_pad3q (QUAD) output
  int st_dt;
  int layout;
  const uint32_t* vol_max;  // non-null: vol holds V * 2^-e (pow2_exp); sums are rescaled
  float* out;
  const float2* sc;
  int geom, W, H, nviews;  // W x H detector
  float sad;
  int NX = 0, NZ = 0;      // volume NX x NX x NZ (0: W x W x H, the standard geometries)
  const float* pts = nullptr;  // GEOM_G3D: ray endpoints [W,H,6] (+ view * pts_view_stride)
  int64_t pts_view_stride = 0;
  int long_ray = 0;  // 2-byte storage: compensated sums above this many samples (0: 1024)
  // batch: `items` volumes vol_stride bytes apart (haloed), outputs out_stride
  // floats apart, vol_max[item] when scaled; one launch for all of them
  int items = 1;
  uint64_t vol_stride = 0, out_stride = 0;
  float* const* outs = nullptr;  // device array: item i's output (else out + i * out_stride)
  int vchunk_log2 = -1;  // block order: views per chunk 2^vchunk_log2 (-1: default 16)
  // the loads-only form: the same blocks, addresses and loads, the taps
  // XOR-folded instead of interpolated (par/fan/cone, DUAL f32 or QUAD 2-byte)
  bool loads_only = false;
};
hipError_t launch_fpz(const FPZLaunch& a, hipStream_t st);

// OSC forward projection + ratio terms over the views first, first+stride, ...
struct OSCFPLaunch {
  const void* mu_z;     // z-fastest volume of storage st_dt with a zero (border) halo
  int st_dt;            // storage of mu_z; pre_pad / post_pad: the same, BF16 for F16
  int layout;           // mu_z layout (LAYOUT_*)
  const uint32_t* vol_max;  // non-null: mu_z holds mu * 2^-e (pow2_exp)
  float* pre;           // [W,H,nProj] updated in place
  float* post;          // [W,H,nProj] updated in place
  void* pre_pad;        // [2,W+4,H+3,count] row-paired copies for the BP (zero halo)
  void* post_pad;
  const float2* sc;
  int geom, W, H, nProj, first, stride, count;
  float sad;
  int long_ray = 0;  // as FPZLaunch::long_ray
};
hipError_t launch_osc_fp(const OSCFPLaunch& a, hipStream_t st);

// ---- backprojection (bp.hip)
struct BPLaunch {
  const void* sino;  // row-grouped haloed sinogram [rows, W+4, H+2*zh-1, nProj] (launch_pad2, edge halo)
  int st_dt;
  const uint32_t* sino_max;  // non-null: sino holds P * 2^-e (pow2_exp); sums are rescaled
  float* out;        // [W, W, H]
  const float2* sc;
  int geom, W, H, nviews;
  float sad;
  int zr;            // z-voxels per thread (32 or 64)
  int rows;          // rows per sinogram element: 2, or 4 (2-byte storage, rows step <= 2)
  float* chain;      // cone f32: scratch of chain_views x (nz-1) x padded columns floats, or null
  int chain_views;   // views per launch the chain scratch holds
  // batch: `items` sinograms sino_stride bytes apart (haloed), volumes
  // out_stride floats apart, sino_max[item] when scaled; one launch per chunk
  int items = 1;
  uint64_t sino_stride = 0, out_stride = 0;
  float* const* outs = nullptr;  // device array: item i's volume (else out + i * out_stride)
  int super_ = 0;    // block order: super-tile side in 8x8 tiles (0: default 4)
  // the loads-only form: the same blocks and loads, each load XOR-folded
  // instead of interpolated (cone only)
  bool loads_only = false;
  // rows of halo above and below in the sinogram copy; zclamp false: every
  // in-circle voxel's detector rows lie within them (the host has shown it),
  // so the kernel reads them without clamping
  int zh = 2;
  bool zclamp = true;
};
// floats of BP chain scratch per view (cone f32, z-range zr, super-tile side)
size_t bp_chain_floats_per_view(int W, int H, int zr, int super_);
hipError_t launch_bp(const BPLaunch& a, hipStream_t st);

struct OSCBPLaunch {
  const void* pre_pad;    // [2, W+4, H+3, count] of storage st_dt (F32 or BF16)
  const void* post_pad;
  int st_dt;
  float* mu_pad;          // padded f32 volume, interior updated in place
  const float2* sc;       // all views
  int geom, W, H, first, stride, count;
  float sad;
};
hipError_t launch_osc_bp_update(const OSCBPLaunch& a, hipStream_t st);
// *out_bits = max(*out_bits, bits of max |x[i]|) over n floats (NaN ignored);
// the caller zero-fills *out_bits first.
hipError_t launch_absmax(const float* x, size_t n, uint32_t* out_bits, hipStream_t st);

// general-3D ray-driven backprojection (atomics)
struct G3DBPLaunch {
  const float* sino;  // [W,H,nProj]
  const float* pts;
  int64_t pts_view_stride;
  const float2* sc;
  float* out;         // z-fastest accumulator acc[(i*M + j)*N + k], zeroed by the caller
  int W, H, nviews, M, N;
  // only rays of more than n_min samples (the bricks take the others)
  int n_min = 0;
  // the [M,M,N] x-fastest volume instead of the z-fastest accumulator;
  // with out64 set, 64-bit fixed-point sums there (scale from pmax) instead
  bool xfast = false;
  int64_t* out64 = nullptr;
  const uint32_t* pmax = nullptr;
  // non-null: the launch does nothing when *gate is 0 (the bricks' count of
  // rays longer than n_min)
  const uint32_t* gate = nullptr;
};
hipError_t launch_g3d_bp(const G3DBPLaunch& a, hipStream_t st);

// general-3D backprojection by voxel bricks (g3dbp.hip)
constexpr int kCkSamples = 64;  // samples between a ray's checkpoints
struct G3DBrickPlan {
  uint32_t tiles_x, tiles_per_view, nbx, nby, nbz, nbricks, tw;
  int n_cap, S, views_per_chunk;
  size_t ck_plane, ray_plane, ck_bytes, mask_bytes, halo_elems;
  bool ok;  // the bricks' bits fit the binning kernel's LDS
};
G3DBrickPlan g3d_brick_plan(int W, int H, int M, int N, int nviews);
struct G3DBrickLaunch {
  G3DBrickPlan plan;
  const float* sino;
  const float* pts;
  int64_t pts_view_stride;
  const float2* sc;
  int W, H, nviews, M, N;
  uint32_t* mask;   // plan.mask_bytes
  float* ck;        // plan.ck_bytes
  float* out;       // [M,M,N]
  int64_t* out64;   // [M,M,N] fixed-point sums
  int64_t* halo64;  // plan.halo_elems fixed-point sums
  const uint32_t* pmax;  // bits of max |sino|: the fixed point's scale
  uint32_t* long_rays;   // one word: the count of rays left to the atomic kernel
};
// Every chunk of views, the halo planes, the rays left to the atomic kernel
// and the conversion to float: out is complete after.
hipError_t launch_g3d_bp_bricks(const G3DBrickLaunch& a, hipStream_t st);
// z-fastest accumulator -> [M,M,N] x-fastest volume
hipError_t launch_g3d_unz(const float* acc, float* out, int M, int N, hipStream_t st);

// ---- TV (tv.hip)
// Strided view of a float volume: element (i,j,k) at p[i + sy*j + sz*k].
struct Vol3 {
  float* p;
  int nx, ny, nz;
  int64_t sy, sz;
};
// `steps` TV gradient steps on x in place.  c is read from c_dev when it is
// not null, else from `c`.  g ([nx,ny,nz]), partials (tv_partials_count
// doubles) and a_dev (one float) are workspace.
hipError_t launch_tvmin(Vol3 x, float* g, double* partials, float* a_dev, const float* c_dev,
                        float c, int steps, hipStream_t st);
size_t tv_partials_count(int nx, int ny, int nz);

// ---- OSC-TV driver helpers (tv.hip)
// *sum = sum over voxels of (old - new)^2 in FP64, a NaN difference counted
// as 0 and old(0,0,0) read as 0 (OSC_TV_recon.m:292,312-314).
hipError_t launch_diff_norm(Vol3 old_v, Vol3 new_v, double* partials, double* sum,
                            hipStream_t st);
// *c_dev = tv * (float)sqrt(*sum)  (OSC_TV_recon.m:314-315)
hipError_t launch_tv_const(const double* sum, float tv, float* c_dev, hipStream_t st);
// NaN and +-Inf -> 0 (OSC_TV_recon.m:316-317)
hipError_t launch_clean(Vol3 x, hipStream_t st);

// ---- FDK prefilter (fdk.hip)
// The staged sample: x = sino[e], or -log(post[e] / sino[e]) with NaN/Inf
// -> 0 when post is set; then x * wgt[v][u] if wgt is set, then
// x * pw[view][u] if pw is set, each rounded to float.
struct FilterIn {
  const float* sino;  // the sinogram, or pre when post is set
  const float* post;  // FBP log-transform input, or null
  const float* wgt;   // cosine weights [H][W], or null
  const float* pw;    // Parker window [nviews][W], or null
};
// out = ((float)(sum_u' x[u'] g[u-u'+W-1]) * dA[view]) / 2 per detector row,
// x promoted to double.
hipError_t launch_prefilter(const FilterIn& in, float* out, const double* g, const float* dA,
                            int W, int H, int nviews, hipStream_t st);
// Parker / Wesarg window: single-precision constants of the scan
struct ParkerConsts {
  float pi_f;  // (float)pi
  float d2;    // (float)(2 delta)
  float eps;   // epsilon = max(extent - (pi + 2 delta), 0), single
};
// win[i*W + u] from col[u] = {(float)(2 alpha_u), b(alpha_u), b(-alpha_u), 0}
// and beta[i] = |angle_i - angle_0|.
hipError_t launch_parker(const float4* col, const float* beta, const ParkerConsts& k, int W,
                         int nviews, float* win, hipStream_t st);
// out[u,v,view] = sino[u,v,view] * win[view*W + u]
hipError_t launch_apply_window(const float* sino, const float* win, float* out, int W, int H,
                               int nviews, hipStream_t st);

// ---- gather microbenchmark (probe.hip)
// pattern 0: the FP's quad-volume loads; 1: the BP's row-paired loads.
// buf holds at least the quad volume / row-paired sinogram of W, H, nviews.
hipError_t launch_gather_probe(const void* buf, int st_dt, int pattern, int W, int H, int nviews,
                               float* sink, int64_t* samples, hipStream_t st);

// ---- context internals shared with the request server (api.hip)
hipStream_t ctx_stream(const mexct_ctx* c);
// sets the calling thread's mexct_last_error() text and returns code
int set_error(int code, const char* msg);
// Upper bound of the device bytes a context's workspace holds after one
// call of operator op (MEXCT_OP_*: FP, BP, FDK = pre-filter + BP, OSC-TV)
// over `items` batched operands, each workspace slot rounded up as the
// context allocates it.  min_subsets: the fewest subsets of any OSC-TV
// iteration (sets the largest per-subset ratio-term copies).
// mexct_fp_batch_dev / mexct_bp_batch_dev with each item's input and output
// at its own device address (host arrays of `count` pointers).
int fp_batch_ptrs(mexct_ctx* c, const void* const* vols, int32_t vol_dtype, int32_t count,
                  int64_t W, int64_t H, float sad, const float* angles, int32_t nProj,
                  int32_t geom, float* const* outs, const mexct_opts* opts);
int bp_batch_ptrs(mexct_ctx* c, const void* const* sinos, int32_t sino_dtype, int32_t count,
                  int64_t W, int64_t H, int64_t nProj, float sad, const float* angles,
                  int32_t geom, float* const* outs, const mexct_opts* opts);
size_t op_workspace_bytes(int op, int geom, int st, int64_t W, int64_t H, int64_t nProj, float sad,
                          int items, int min_subsets);

}  // namespace mexct
