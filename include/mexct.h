/*
 * mexct.h -- C ABI of the MI355X (gfx950) CT projection and
 * iterative-reconstruction operators.
 *
 * Each entry point replaces one mexFunction of kurtishdekker/mex-cuda-CT
 * (cited below) with the same argument meaning, array layout and result.
 * Arrays are MATLAB `single`, column-major:
 *   volume    vol[x + W*y + W*W*z]           dims [W, W, H]
 *   sinogram  P[u + W*v + W*H*view]          dims [W, H, nProj]
 *   rays      pts[ix + W*iy + W*H*k], k<6    dims [W, H, 6] (x1,y1,z1,x2,y2,z2)
 * Angles are float radians; their sine/cosine is taken with host sincosf
 * as in the reference.  geom: 0 parallel, 1 fan, 2 cone.
 *
 * Two families:
 *   mexct_*      host pointers, synchronous (the MEX call contract: copy in,
 *                compute, copy out).  Usable from any thread; each thread
 *                uses its own implicit context on the current HIP device.
 *   mexct_*_dev  device pointers, asynchronous on the context's stream.
 *
 * Return value: MEXCT_OK (0) or a negative MEXCT_E* code; the message of
 * the last failure on the calling thread is mexct_last_error().
 * Nothing in this header uses torch or HIP types.
 */
#ifndef MEXCT_H
#define MEXCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MEXCT_OK 0
#define MEXCT_EINVAL (-1)  /* bad argument: null pointer, dims, counts      */
#define MEXCT_EGEOM (-2)   /* geomFlag not 0/1/2 ("invalid geometry")         */
#define MEXCT_EHIP (-3)    /* HIP runtime error                                */
#define MEXCT_ENOMEM (-4)  /* device allocation failed or over the HBM budget */
#define MEXCT_ENOTSUP (-5) /* unsupported option combination                  */

/* storage type of the gathered operand (volume for FP, sinogram for BP);
 * arithmetic (weights, ray sums, accumulators) is always FP32.  F32 is the
 * reference's arithmetic bit for bit.  F16/BF16 also step the sample
 * coordinates in texel space (a few ulps from the reference's rounding);
 * F16 copies of f32 data hold x * 2^-e, e the binary exponent of max|x|, so
 * that any data scale stays inside FP16's range.  In the OSC update the
 * projected volume copy takes the storage type and the ratio terms take it
 * too, except that F16 stores them in BF16 (their range exceeds FP16's). */
#define MEXCT_F32 0
#define MEXCT_F16 1
#define MEXCT_BF16 2

typedef struct mexct_opts {
  int32_t storage;  /* MEXCT_F32 (default, reference-exact), MEXCT_F16, MEXCT_BF16 */
  int32_t tv_steps; /* TVmin gradient steps; 0 -> 20 as CUDAmex_TVmin_3D.cu:221  */
  int32_t zrange;   /* BP z-voxels per thread: 32, 48 or 64; 0 -> chosen by H    */
  int32_t long_ray; /* F16/BF16 FP: rays of more samples than this are summed with  */
                    /* compensated (Kahan) summation; 0 -> 1024, < 0 -> never       */
  int32_t g3d_bp;   /* general-3D BP: MEXCT_G3D_BRICKS (0, default),                */
                    /* MEXCT_G3D_DETERMINISTIC or MEXCT_G3D_ATOMICS (see below)      */
} mexct_opts;

typedef struct mexct_ctx mexct_ctx;

/* Context = device + stream + workspace.  stream may be NULL (the context
 * creates one at default priority) or an existing hipStream_t. */
int mexct_ctx_create(mexct_ctx** ctx, int device, void* stream);
int mexct_ctx_destroy(mexct_ctx* ctx);
int mexct_ctx_sync(mexct_ctx* ctx);
/* bytes of device workspace currently held by the context */
int64_t mexct_ctx_workspace_bytes(const mexct_ctx* ctx);
const char* mexct_last_error(void);
int mexct_version(void);

/* ------------------------------------------------------------------ FP
 * fp = CUDAmex_FP(vol, SAD, angles, nProjections, geomFlag)
 * replaces projection_kernels/CUDAmex_FP.cu:350-477 (kernels :56-346).
 * vol [W,W,H] -> out [W,H,nProj].  Cone marches with step 1.0, par/fan 0.5,
 * clamp addressing.  Detector rows iy >= 32*ceil(W/32) stay 0 as in the
 * reference's grid (CUDAmex_FP.cu:432-440).  n_angles >= nProj is required
 * (the reference reads past the end otherwise).  Volumes up to
 * (W+4)^2 (H+4) < 2^32 haloed elements (the reference's unsigned 32-bit
 * index range, CUDAmex_FP.cu:394); larger ones are MEXCT_EINVAL. */
int mexct_fp(const float* vol, int64_t W, int64_t H, float sad, const float* angles,
             int64_t n_angles, int32_t nProj, int32_t geom, float* out, const mexct_opts* opts);
int mexct_fp_dev(mexct_ctx* ctx, const void* vol, int32_t vol_dtype, int64_t W, int64_t H,
                 float sad, const float* angles, int32_t nProj, int32_t geom, float* out,
                 const mexct_opts* opts);

/* ------------------------------------------------------------------ BP
 * bp = CUDAmex_BP(P, SAD, angles, geomFlag)
 * replaces projection_kernels/CUDAmex_BP.cu:223-324 (kernels :44-221).
 * sino [W,H,nProj] -> out [W,W,H]; voxels outside the circle
 * (ix-hw)^2+(iy-hw)^2 < hw^2, hw = W/2-1, are 0.  Views accumulate in
 * view order; no 1/nProj or pi factor (as the reference).  Volumes below
 * 2^32 elements (CUDAmex_BP.cu:217) and any number of views (the reference
 * walks them one at a time, :287-291). */
int mexct_bp(const float* sino, int64_t W, int64_t H, int64_t nProj, float sad,
             const float* angles, int64_t n_angles, int32_t geom, float* out,
             const mexct_opts* opts);
int mexct_bp_dev(mexct_ctx* ctx, const void* sino, int32_t sino_dtype, int64_t W, int64_t H,
                 int64_t nProj, float sad, const float* angles, int32_t geom, float* out,
                 const mexct_opts* opts);

/* ------------------------------------------------------------------ batched FP / BP
 * No counterpart in the reference (one MEX call per volume).  `count`
 * (1..256) volumes or sinograms of one geometry, stored one after another,
 * projected or backprojected by one kernel launch (per chunk of views for
 * the BP); outputs one after another.  Item i's result has the same bits as
 * the single call on item i. */
int mexct_fp_batch_dev(mexct_ctx* ctx, const void* vols, int32_t vol_dtype, int32_t count,
                       int64_t W, int64_t H, float sad, const float* angles, int32_t nProj,
                       int32_t geom, float* outs, const mexct_opts* opts);
int mexct_bp_batch_dev(mexct_ctx* ctx, const void* sinos, int32_t sino_dtype, int32_t count,
                       int64_t W, int64_t H, int64_t nProj, float sad, const float* angles,
                       int32_t geom, float* outs, const mexct_opts* opts);

/* ------------------------------------------------------------------ autotuner
 * No counterpart in the reference.  The FP and BP kernels' block order (FP:
 * views per chunk; BP: super-tile side, and the z-range for FP32 storage
 * unless opts fixes it) is tuned per problem shape (operator, geometry,
 * storage, W, H, nProj; any batch count): the tuning call of a shape times
 * every candidate once on its first 128 (FP) or 32 (BP) views, keeps the
 * fastest for the
 * process and runs the call with it.  Candidates write identical bits.
 * mode 1: tune at a shape's first call; 2: at its second (one-off shapes
 * pay nothing; the request server's contexts use this); 0 off (default:
 * built-in choices).  A context's own setting (-1: follow the global one)
 * overrides.  mexct_tuned_config fills cfg[4] = {found, FP views-per-chunk
 * log2 or BP super-tile side, BP z-range, the winner's time on the tuning
 * views in microseconds}. */
int mexct_set_autotune(int32_t mode);
int mexct_ctx_set_autotune(mexct_ctx* ctx, int32_t mode);
int mexct_tuned_config(int32_t op, int32_t geom, int32_t storage, int64_t W, int64_t H,
                       int64_t nProj, int32_t* cfg);

/* ------------------------------------------------------------------ general-3D FP
 * fp = CUDAmex_general3D_FP(vol, angles, pts, nProjections)
 * replaces projection_kernels/CUDAmex_general3D_FP.cu:154-252 (kernel :59-150).
 * vol [side,side,nSlices], pts [W,H,6] rotated about z by each angle,
 * step 0.5, border addressing -> out [W,H,nProj].
 * pts_per_view = 1 is an additive extension: pts is [W,H,6,nProj] and
 * view k uses its own slice (still rotated by angles[k]). */
int mexct_g3d_fp(const float* vol, int64_t side, int64_t nSlices, const float* angles,
                 int64_t n_angles, const float* pts, int64_t W, int64_t H, int32_t nProj,
                 int32_t pts_per_view, float* out, const mexct_opts* opts);
int mexct_g3d_fp_dev(mexct_ctx* ctx, const void* vol, int32_t vol_dtype, int64_t side,
                     int64_t nSlices, const float* angles, const float* pts_dev, int64_t W,
                     int64_t H, int32_t nProj, int32_t pts_per_view, float* out,
                     const mexct_opts* opts);

/* ------------------------------------------------------------------ general-3D BP
 * bp = CUDAmex_general3D_BP(P, angles, pts, M, N)
 * replaces projection_kernels/CUDAmex_general3D_BP.cu:194-282 (kernel :61-188).
 * Every 0.5-voxel sample of every ray adds the reference's two weighted
 * corners to [M,M,N].  opts->g3d_bp selects how the sums are formed:
 *   MEXCT_G3D_BRICKS: per 16x16x12 voxel brick in LDS, written once, with
 *     64-bit fixed-point sums that are exact in any order, so every run
 *     gives the same bits (each added corner is rounded to a step of
 *     max|P| * 2^-27 or finer: one sample's corner in the bricks, a run of
 *     at most 4 samples in one cell in the long-ray kernel; up to 2^32 of
 *     them per voxel);
 *   MEXCT_G3D_DETERMINISTIC: the same, and MEXCT_ENOTSUP where the bricks
 *     cannot run instead of the atomics;
 *   MEXCT_G3D_ATOMICS: the ray-driven scatter with float atomics (the
 *     reference's; its sum order is not fixed).
 * Volumes whose bricks exceed the binning pass's LDS (about 730^3) take
 * the atomics in MEXCT_G3D_BRICKS mode. */
#define MEXCT_G3D_BRICKS 0
#define MEXCT_G3D_DETERMINISTIC 1
#define MEXCT_G3D_ATOMICS 2
int mexct_g3d_bp(const float* sino, int64_t W, int64_t H, int64_t nProj, const float* angles,
                 int64_t n_angles, const float* pts, int32_t pts_per_view, int64_t M, int64_t N,
                 float* out, const mexct_opts* opts);
int mexct_g3d_bp_dev(mexct_ctx* ctx, const float* sino, int64_t W, int64_t H, int64_t nProj,
                     const float* angles, const float* pts_dev, int32_t pts_per_view, int64_t M,
                     int64_t N, float* out, const mexct_opts* opts);

/* ------------------------------------------------------------------ OSC
 * CUDAmex_oscIter(mu, pre, post, numSubsets, subsetPicks, angles, SAD, geomFlag)
 * replaces iterative_reconstruction/CUDAmex_oscIter.cu:362-584
 * (kernels :94-358).  mu [W,W,H], pre/post [W,H,nProj] are updated IN
 * PLACE exactly as the reference mutates its MATLAB inputs: for every
 * processed view pre <- pre*exp(-l) - post and post <- pre*exp(-l)*l.
 * picks are 0-based int32 subset offsets (n_picks >= numSubsets). */
int mexct_osc_iter(float* mu, int64_t W, int64_t H, float* pre, float* post, int64_t nProj,
                   int32_t numSubsets, const int32_t* picks, int64_t n_picks,
                   const float* angles, int64_t n_angles, float sad, int32_t geom,
                   const mexct_opts* opts);
int mexct_osc_iter_dev(mexct_ctx* ctx, float* mu, int64_t W, int64_t H, float* pre,
                       float* post, int64_t nProj, int32_t numSubsets, const int32_t* picks,
                       const float* angles, float sad, int32_t geom, const mexct_opts* opts);

/* ------------------------------------------------------------------ TV
 * out = CUDAmex_TVmin_3D(vol, c)
 * replaces iterative_reconstruction/CUDAmex_TVmin_3D.cu:162-246
 * (kernel :101-158).  opts->tv_steps steps (default 20) of
 * x <- x - (c/|g|) g; |g| is reduced in FP64 on the device.  The gradient
 * border, uninitialised in the reference, is 0. */
int mexct_tvmin(const float* vol, int64_t Nx, int64_t Ny, int64_t Nz, float c, float* out,
                const mexct_opts* opts);
int mexct_tvmin_dev(mexct_ctx* ctx, float* vol, int64_t Nx, int64_t Ny, int64_t Nz, float c,
                    const mexct_opts* opts);

/* ------------------------------------------------------------------ FDK filter
 * filtered = preFilterSinogram(sino, geom, angles, filter, cutoff[, parker_q])
 * (fbp_reconstruction/preFilterSinogram.m:29-152): cosine weight
 * (fan/cone), Parker window when parker_q >= 0, band-limited ramp x window
 * on a zero-padded FFT grid of order max(64, 2^nextpow2(2W)), then
 * x (angle step)/2.  parker_q = MEXCT_NO_PARKER is the 5-argument call.
 * window: 0 ramp, 1 shepp-logan, 2 cosine, 3 hamming, 4 hann, 5 none.
 * The convolution is evaluated in FP64 on the device.  The _dev form reads
 * nProj angles; the host form reads n_angles, of which the Parker window
 * takes the scan direction and extent (diff(angles), angles(end)). */
#define MEXCT_NO_PARKER (-1.0)
int mexct_prefilter(const float* sino, int64_t W, int64_t H, int64_t nProj, int32_t geom,
                    double sad, const float* angles, int64_t n_angles, int32_t window,
                    double cutoff, double parker_q, float* out);
int mexct_prefilter_dev(mexct_ctx* ctx, const float* sino, int64_t W, int64_t H, int64_t nProj,
                        int32_t geom, double sad, const float* angles, int32_t window,
                        double cutoff, double parker_q, float* out);

/* ------------------------------------------------------------------ allocation-free pre-filter
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mexct_prefilter_dev with the same arguments, kernels
 * and results, but its tables (filter taps, cosine weights, angle steps,
 * Parker window) live in a fixed 16 MiB device array that every GPU holds
 * from library load to process end, not in the context's workspace.  It
 * calls no hipMalloc, hipFree or hipHostMalloc.  Tables needing more than
 * 16 MiB (about 8*(2W-1) + 4*W*H (fan/cone) + 4*nProj + 4*W*nProj (Parker)
 * bytes) are rejected with MEXCT_EINVAL.  Every call on a GPU shares that
 * GPU's array and it keeps the last call's tables, so calls that overlap on
 * one GPU race: serialise them.  A call whose tables differ from the last
 * call's copies them in and synchronises the context's stream. */
int mexct_prefilter_noalloc_dev(mexct_ctx* ctx, const float* sino, int64_t W, int64_t H,
                                int64_t nProj, int32_t geom, double sad, const float* angles,
                                int32_t window, double cutoff, double parker_q, float* out);

/* ------------------------------------------------------------------ Parker weights
 * weightedSino = parker_weight(sino, angles, geom, q)
 * replaces fbp_reconstruction/parker_weight.m:1-79 (Wesarg et al. 2002;
 * q = 1 is Parker 1982).  q >= 0; SAD is ignored for par3d (alpha = 0).
 * The window runs along the detector columns u; the reference computes it
 * over size(sino,2) and needs W == H, where the two agree.  The
 * reference's warnings and disp lines are not reproduced. */
int mexct_parker_weight(const float* sino, int64_t W, int64_t H, int64_t nProj,
                        const float* angles, int64_t n_angles, int32_t geom, double sad, double q,
                        float* out);
int mexct_parker_weight_dev(mexct_ctx* ctx, const float* sino, int64_t W, int64_t H,
                            int64_t nProj, const float* angles, int32_t geom, double sad, double q,
                            float* out);

/* ------------------------------------------------------------------ FBP driver
 * recon = FBP_recon(pre, post, scanAngles, geom, filter, filterCutoff, parker_q)
 * replaces fbp_reconstruction/FBP_recon.m:1-40: -log(post./pre) with NaN
 * and Inf set to 0 (a negative ratio, complex in MATLAB, gives 0), the
 * pre-filter above, then CUDAmex_BP of the geometry into vol [W,W,H].
 * The log transform is fused into the filter's staging.  opts selects the
 * backprojector's storage and z-range. */
int mexct_fbp_recon(const float* pre, const float* post, int64_t W, int64_t H, int64_t nProj,
                    const float* angles, int64_t n_angles, int32_t geom, double sad,
                    int32_t window, double cutoff, double parker_q, float* vol,
                    const mexct_opts* opts);
int mexct_fbp_recon_dev(mexct_ctx* ctx, const float* pre, const float* post, int64_t W,
                        int64_t H, int64_t nProj, const float* angles, int32_t geom, double sad,
                        int32_t window, double cutoff, double parker_q, float* vol,
                        const mexct_opts* opts);

/* ------------------------------------------------------------------ OSC-TV driver
 * recon = OSC_TV_recon(pre, post, angles, geom, nIter, TV, sInit, sFinal, p)
 * (iterative_reconstruction/OSC_TV_recon.m:1-102) with all state resident
 * on the device.  first_picks[k] (1-based, k < nIter) replaces MATLAB's
 * randperm(n,1) draw of iteration k.  pre/post are not modified. */
int mexct_osc_tv_recon(const float* pre, const float* post, int64_t W, int64_t H,
                       int64_t nProj, const float* angles, int64_t n_angles, float sad,
                       int32_t geom, int32_t nIter, float tv, int32_t sInit, int32_t sFinal,
                       double p, const int32_t* first_picks, float* mu_out,
                       const mexct_opts* opts);
int mexct_osc_tv_recon_dev(mexct_ctx* ctx, const float* pre, const float* post, int64_t W,
                           int64_t H, int64_t nProj, const float* angles, float sad,
                           int32_t geom, int32_t nIter, float tv, int32_t sInit,
                           int32_t sFinal, double p, const int32_t* first_picks, float* mu_out,
                           const mexct_opts* opts);

/* Checkpoint / resume of the driver above (no counterpart in the
 * reference, whose only state between iterations is mu, OSC_TV_recon.m:44-71):
 * runs iterations k_begin..k_end (1 <= k_begin <= k_end <= nIter) of the
 * nIter-iteration schedule on mu [W,W,H] in place.  On entry mu holds the
 * volume after iteration k_begin - 1 (ignored for k_begin = 1, which starts
 * from 1e-10); on return the volume after k_end.  Splitting 1..nIter into
 * consecutive ranges gives the bits of one mexct_osc_tv_recon call. */
int mexct_osc_tv_iterations(const float* pre, const float* post, int64_t W, int64_t H,
                            int64_t nProj, const float* angles, int64_t n_angles, float sad,
                            int32_t geom, int32_t nIter, float tv, int32_t sInit, int32_t sFinal,
                            double p, const int32_t* first_picks, int32_t k_begin, int32_t k_end,
                            float* mu, const mexct_opts* opts);
int mexct_osc_tv_iterations_dev(mexct_ctx* ctx, const float* pre, const float* post, int64_t W,
                                int64_t H, int64_t nProj, const float* angles, float sad,
                                int32_t geom, int32_t nIter, float tv, int32_t sInit,
                                int32_t sFinal, double p, const int32_t* first_picks,
                                int32_t k_begin, int32_t k_end, float* mu,
                                const mexct_opts* opts);

/* ------------------------------------------------------------------ measurement
 * No reference counterpart: the gather-throughput microbenchmark behind the
 * projectors' "measured gather roofline".  Launches, on the context's stream,
 * a kernel that issues the load instructions of the forward projector
 * (MEXCT_PROBE_FP: per sample two 16-byte loads of the x-pair volume in f32,
 * one 16-byte load of the corner-quad volume in f16/bf16) or of the
 * backprojector (MEXCT_PROBE_BP: per
 * update one 16-byte / 8-byte load of the row-paired sinogram) with ideal
 * coherence, over a buffer of the projector's size for W, H, nviews held in
 * the context's workspace (its contents are whatever the last projection
 * left there).  *samples receives the number of samples (FP) or updates
 * (BP) the launch performs. */
/* The forward projector's own gather ceiling: mexct_fp_dev's halo pass and
 * projector launch with the same block order, per-lane addresses and loads,
 * but each ray's taps XOR-folded instead of interpolated and summed; out
 * receives the folds (not a projection).  The projector cannot outrun it.
 * par/fan/cone while the volume fits the z-fastest layouts
 * (MEXCT_ENOTSUP otherwise). */
int mexct_fp_loads_dev(mexct_ctx* ctx, const void* vol, int32_t vol_dtype, int64_t W, int64_t H,
                       float sad, const float* angles, int32_t nProj, int32_t geom, float* out,
                       const mexct_opts* opts);
/* The backprojector's own gather ceiling, as mexct_fp_loads_dev: mexct_bp_dev's
 * copy and launches (cone; MEXCT_ENOTSUP otherwise) with every load but
 * each loaded element XOR-folded instead of interpolated; out receives the
 * folds. */
int mexct_bp_loads_dev(mexct_ctx* ctx, const void* sino, int32_t sino_dtype, int64_t W, int64_t H,
                       int64_t nProj, float sad, const float* angles, int32_t geom, float* out,
                       const mexct_opts* opts);
#define MEXCT_PROBE_FP 0
#define MEXCT_PROBE_BP 1
/* pattern | MEXCT_PROBE_UNROLL(k): the probe loop unrolled 2^k times (k = 0..3,
 * up to 2^k samples' loads in flight per wave); the ceiling is the best one */
#define MEXCT_PROBE_UNROLL(k) ((k) << 4)
int mexct_probe_gather_dev(mexct_ctx* ctx, int32_t storage, int64_t W, int64_t H, int32_t nviews,
                           int32_t pattern, int64_t* samples);

/* ------------------------------------------------------------------ request server
 * No reference counterpart (each MEX call owns the GPU while it runs).  A
 * continuous-batching server for many callers sharing one GPU:
 * `slots` workers, each with its own stream and workspace, take queued
 * requests in arrival order as long as the sum of their device footprints
 * stays within hbm_budget bytes; a worker that takes a request also takes up
 * to max_group-1 queued requests with the same operator, geometry, sizes,
 * angles, SAD and storage (at most an even share of those among the idle
 * workers) and runs an FP, BP or FDK group as one batched launch, an OSC-TV
 * group back to back.  Requests run the
 * same operators as the entry points above, so a result does not depend on
 * what else is in flight.  Arrays are host memory in the MATLAB layout, or,
 * with resident = 1, device memory of the server's device (checked at
 * submit: each array must lie in a device allocation of that device that
 * holds the request's bytes from that address on);
 * the caller keeps input/output alive and unchanged until mexct_server_wait
 * returns for the ticket.
 *   FP:    input vol [W,W,H]          -> output sino [W,H,nProj]
 *   BP:    input sino [W,H,nProj]     -> output vol [W,W,H]
 *   FDK:   input sino, prefilter(window, cutoff) then BP -> output vol
 *   OSCTV: input pre, input2 post     -> output mu [W,W,H] (mexct_osc_tv_recon) */
#define MEXCT_OP_FP 0
#define MEXCT_OP_BP 1
#define MEXCT_OP_FDK 2
#define MEXCT_OP_OSCTV 3

typedef struct mexct_request {
  int32_t op, geom, storage, window; /* window as mexct_prefilter (FDK only) */
  int64_t W, H, nProj;               /* angles holds nProj entries          */
  float sad;
  float osc_tv;                      /* OSCTV: TV, nIter, sInit, sFinal, p  */
  double cutoff;
  double osc_p;
  int32_t osc_iters, osc_s_init, osc_s_final, tv_steps; /* tv_steps as mexct_opts */
  const float* angles;
  const float* input;
  const float* input2;
  float* output;
  const int32_t* first_picks; /* OSCTV: osc_iters entries */
  int32_t resident;           /* 1: input, input2 and output are device arrays */
} mexct_request;

typedef struct mexct_server mexct_server;

int mexct_server_create(mexct_server** sv, int device, int64_t hbm_budget, int32_t slots,
                        int32_t max_group);
/* Device bytes a request is reserved against the budget: its staging
 * arrays plus the workspace the entry point it runs takes (an upper bound
 * of what a worker holds while running it alone); MEXCT_EINVAL (negative)
 * for a malformed request. */
int64_t mexct_server_footprint(const mexct_request* req);
/* Queues a request; MEXCT_ENOMEM if its footprint alone exceeds the budget. */
int mexct_server_submit(mexct_server* sv, const mexct_request* req, uint64_t* ticket);
/* Blocks until the request has finished; *status is its return code (on
 * error mexct_last_error() holds its message), latency_s = submit -> output
 * in host memory, queue_s = submit -> start.  Any pointer may be NULL. */
int mexct_server_wait(mexct_server* sv, uint64_t ticket, int32_t* status, double* latency_s,
                      double* queue_s);
/* stats[0..n): done, groups, requests run in groups, queued, bytes reserved,
 * peak bytes reserved, workspace releases, requests run in batched
 * launches, batched groups. */
int mexct_server_stats(mexct_server* sv, int64_t* stats, int32_t n);
/* Finishes running requests, drops queued ones, frees everything. */
int mexct_server_destroy(mexct_server* sv);

/* OSC_TV_recon.m:323-349 helpers (host only). */
int mexct_calc_num_subsets(int32_t sInit, int32_t sFinal, double p, int32_t K, int32_t k);
int mexct_calc_subset_order(int32_t n, int32_t first_pick, int32_t* picks);

#ifdef __cplusplus
}
#endif
#endif /* MEXCT_H */
