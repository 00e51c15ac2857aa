/*
 * oracle.h -- CPU restatement of the mex-cuda-CT operators.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against; it is loaded only by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  The
 * product (libmexct) never links, loads or calls it.
 *
 * Every function restates one reference kernel or host routine literally:
 * same float operations in the same order, no FP contraction
 * (-ffp-contract=off), glibc sincosf/sqrtf/expf, exact FP32 texture
 * weights (lerp = fmaf(t, b - a, a), x first, then y, then z).
 * Each *_f64 entry point is the "FP64 twin": the identical algorithm run
 * in double (geometry, weights and sums), used as the accuracy reference
 * for the reduced-precision storage modes.
 *
 * Arrays are MATLAB column-major: vol[x + W*y + W*W*z] for a [W,W,H]
 * volume and P[u + W*v + W*H*view] for a [W,H,nProj] sinogram.
 */
#ifndef MEXCT_ORACLE_H
#define MEXCT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* address modes of the reference's textures */
#define ORC_CLAMP 0
#define ORC_BORDER 1

/* thread count used by every call below (0 = hardware concurrency) */
void orc_set_threads(int n);
int orc_get_threads(void);

/* Texture weight mode of every sampler call below: 0 (default) exact FP32
 * fractions; 8 rounds each fraction to a multiple of 2^-8, the 9-bit
 * fixed-point weight (8 fraction bits) of the CUDA Programming Guide's
 * texture-fetching appendix, to bound the distance between exact weights
 * and NVIDIA's texture unit.  Process-wide; not for concurrent callers. */
void orc_set_texture_weight_bits(int bits);
int orc_get_texture_weight_bits(void);

/* FP32 texture fetch restatement (for pin tests of the sampler). */
float orc_tex3(const float* T, int nx, int ny, int nz, float x, float y, float z, int mode);
float orc_tex2(const float* T, int nx, int ny, float x, float y, int mode);

/* CUDAmex_FP (CUDAmex_FP.cu:56-346 kernels, :350-477 driver).
 * vol [W,W,H] -> out [W,H,nProj]; returns the total ray-march sample count. */
int64_t orc_fp(const float* vol, int W, int H, float sad, const float* angles,
               int nProj, int geom, float* out);
int64_t orc_fp_f64(const float* vol, int W, int H, double sad, const float* angles,
                   int nProj, int geom, double* out);
/* FP64 twin on a subset of rays: ray r = (ix[r], iy[r], view[r]). */
void orc_fp_f64_rays(const float* vol, int W, int H, double sad, const float* angles,
                     int geom, int nrays, const int* ix, const int* iy, const int* view,
                     double* out);

/* FP32 restatement on a subset of rays (ix[r], iy[r], view[r]); the ray
 * sums equal orc_fp's entries for the same pixels. */
void orc_fp_rays(const float* vol, int W, int H, float sad, const float* angles, int geom,
                 int nrays, const int* ix, const int* iy, const int* view, float* out);
/* FP32 BP of a subset of voxel columns; out [ncol, H] (row c = column c). */
void orc_bp_cols(const float* sino, int W, int H, int nProj, float sad, const float* angles,
                 int geom, int ncol, const int* ix, const int* iy, float* out);

/* CUDAmex_BP (CUDAmex_BP.cu:44-221 kernels, :223-324 driver).
 * sino [W,H,nProj] -> out [W,W,H] (zero outside the circle). */
void orc_bp(const float* sino, int W, int H, int nProj, float sad, const float* angles,
            int geom, float* out);
void orc_bp_f64(const float* sino, int W, int H, int nProj, double sad, const float* angles,
                int geom, double* out);
/* FP64 twin on a subset of voxel columns (ix[c], iy[c]); out [ncol, H]. */
void orc_bp_f64_cols(const float* sino, int W, int H, int nProj, double sad,
                     const float* angles, int geom, int ncol, const int* ix, const int* iy,
                     double* out);

/* CUDAmex_general3D_FP (CUDAmex_general3D_FP.cu:59-150, :154-252).
 * vol [side,side,nSlices]; pts [W,H,6] (pts_per_view = 0) or
 * [W,H,6,nProj] (pts_per_view = 1, additive extension) -> out [W,H,nProj]. */
int64_t orc_g3d_fp(const float* vol, int side, int nSlices, const float* angles,
                   const float* pts, int W, int H, int nProj, int pts_per_view, float* out);
void orc_g3d_fp_f64(const float* vol, int side, int nSlices, const float* angles,
                    const float* pts, int W, int H, int nProj, int pts_per_view, double* out);

/* CUDAmex_general3D_BP (CUDAmex_general3D_BP.cu:61-188, :194-282).
 * Scatter order is canonical: view, then iy, then ix, then sample. */
void orc_g3d_bp(const float* sino, int W, int H, int nProj, const float* angles,
                const float* pts, int pts_per_view, int M, int N, float* out);
void orc_g3d_bp_f64(const float* sino, int W, int H, int nProj, const float* angles,
                    const float* pts, int pts_per_view, int M, int N, double* out);

/* CUDAmex_oscIter (CUDAmex_oscIter.cu:94-358 kernels, :362-584 driver).
 * mu [W,W,H], pre/post [W,H,nProj] are updated in place. */
int orc_osc_iter(float* mu, int W, int H, float* pre, float* post, int nProj,
                 int numSubsets, const int32_t* picks, const float* angles, float sad,
                 int geom);
int orc_osc_iter_f64(double* mu, int W, int H, double* pre, double* post, int nProj,
                     int numSubsets, const int32_t* picks, const float* angles, double sad,
                     int geom);

/* CUDAmex_TVmin_3D (CUDAmex_TVmin_3D.cu:101-158 kernel, :162-246 driver),
 * x [Nx,Ny,Nz] updated in place over `steps` steps (reference: 20).
 * The gradient norm is summed in double (the reference sums in float). */
void orc_tvmin(float* x, int Nx, int Ny, int Nz, float c, int steps);
void orc_tvmin_f64(double* x, int Nx, int Ny, int Nz, double c, int steps);
/* one gradient evaluation (interior only, border left 0) */
void orc_tv_gradient(const float* x, int Nx, int Ny, int Nz, float* g);

/* designFilter (preFilterSinogram.m:95-152): returns `order`, fills
 * filt[order] (double).  window: 0 ramp, 1 shepp-logan, 2 cosine,
 * 3 hamming, 4 hann, 5 none.  Pass filt = NULL to query order. */
int orc_design_filter(int window, int len, double d, double* filt);
/* preFilterSinogram.m:29-90.  geom: 0 par3d (no cosine weight), 1 fan3d,
 * 2 cone3d.  parker_q >= 0 applies parker_weight.m with q = parker_q after
 * the cosine weight; parker_q < 0 skips it (the 5-argument call).
 * sino [W,H,nProj] -> out. */
int orc_prefilter(const float* sino, int W, int H, int nProj, int geom, double sad,
                  const float* angles, int window, double d, double parker_q, float* out);
int orc_prefilter_f64(const float* sino, int W, int H, int nProj, int geom, double sad,
                      const float* angles, int window, double d, double parker_q, double* out);
/* parker_weight.m:29-76: win[i*W + u], the weight of view i, column u. */
int orc_parker_window(int W, int nProj, int geom, double sad, const float* angles, double q,
                      float* win);
/* FBP_recon.m:26-28: out = -log(post./pre) with NaN/Inf -> 0. */
void orc_log_transform(const float* pre, const float* post, int64_t n, float* out);

/* OSC_TV_recon.m:323-349 helpers. */
int orc_calc_num_subsets(int sInit, int sFinal, double p, int K, int k);
void orc_calc_subset_order(int n, int first_pick /* 1-based */, int32_t* picks);

/* OSC_TV_recon.m:1-102 driver (figure() dropped).  first_picks[k] is the
 * 1-based randperm draw of iteration k.  pre/post are not modified. */
int orc_osc_tv_recon(const float* pre, const float* post, int W, int H, int nProj,
                     const float* angles, float sad, int geom, int nIter, float tv,
                     int sInit, int sFinal, double p, int tv_steps,
                     const int32_t* first_picks, float* mu_out);
/* The same driver in FP64 throughout (inputs promoted from float). */
int orc_osc_tv_recon_f64(const float* pre, const float* post, int W, int H, int nProj,
                         const float* angles, float sad, int geom, int nIter, float tv,
                         int sInit, int sFinal, double p, int tv_steps,
                         const int32_t* first_picks, double* mu_out);

#ifdef __cplusplus
}
#endif
#endif
