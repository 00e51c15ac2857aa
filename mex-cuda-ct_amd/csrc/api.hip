// api.hip -- the C ABI (include/mexct.h): argument checks, workspace,
// per-view geometry tables and kernel launches.
//
// The reference's mexFunctions allocate, copy and free device memory on
// every call and synchronise after every view.  Here a context owns a
// stream and a grow-only workspace reused across calls; one call is a short
// fixed sequence of launches on that stream (halo/convert pass, projector
// kernel), and the host-pointer entry points add one copy in and one copy
// out around it.
#include "../../include/mexct.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <climits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "common.h"
#include "internal.h"

using namespace mexct;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// "HIP ERROR: <call> = <code> (<string>) at (<file>:<line>)", the format of
// the reference's CHECK_CUDA_ERROR (CUDAmex_FP.cu:43-50).
#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess)                                                                      \
      return fail(MEXCT_EHIP, "HIP ERROR: %s = %d (%s) at (%s:%d)", #x, (int)e_,              \
                  hipGetErrorString(e_), __FILE__, __LINE__);                                  \
  } while (0)

#define RC(x)              \
  do {                     \
    int r_ = (x);          \
    if (r_) return r_;     \
  } while (0)

enum Slot {
  WS_HOST_A,  // host-entry staging of the first array argument
  WS_HOST_B,
  WS_HOST_C,
  WS_HOST_OUT,
  WS_VOL,      // haloed volume (storage dtype)
  WS_SINO,     // haloed sinogram (storage dtype)
  WS_PTS,
  WS_PREPAD,
  WS_POSTPAD,
  WS_MUPAD,
  WS_MUOLD,
  WS_TVG,
  WS_PARTIALS,
  WS_SCALARS,
  WS_FILT,
  WS_WGT,
  WS_DA,
  WS_PRE0,
  WS_POST0,
  WS_PRE2,
  WS_POST2,
  WS_PREW,
  WS_POSTW,
  WS_PLAIN,
  WS_CHAIN,    // cone f32 BP: row coordinates at the z-range starts
  WS_MUZ,      // z-fastest copy of the OSC volume
  WS_MAXBITS,  // max|x| bits for the FP16 copies: [1] volume, [2] sinogram
  WS_PARKER,   // Parker window: column terms, view betas, then [nProj][W] weights
  WS_FILTERED, // FBP: the filtered sinogram handed to the backprojector
  WS_G3D_CK,   // general-3D BP bricks: ray checkpoints
  WS_G3D_MASK, // (brick, tile) bits
  WS_G3D_HALO, // halo planes (64-bit fixed point)
  WS_G3D_ACC,  // 64-bit fixed-point volume
  WS_OUTPTRS,  // batches with an output pointer per item
  WS_COUNT
};

struct DevGuard {
  int prev = -1;
  explicit DevGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DevGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

int store_bytes(int dt) { return dt == MEXCT_F32 ? 4 : 2; }

}  // namespace

// Fixed device scratch of mexct_prefilter_noalloc_dev: the pre-filter's
// tables (filter taps, cosine weights, angle steps, Parker window) of the
// last call on each GPU.  Every GPU the library runs on holds its own 16 MiB
// copy from module load to process end.  In host code `scratch` only names
// the symbol; kernels and copies get its device address from
// scratch_address().
constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

// What `scratch` holds on one device: the tables of the call described by
// the key, when valid.
struct ScratchState {
  unsigned char* ptr = nullptr;  // device address of `scratch` on this device
  bool valid = false;
  int64_t W = -1, H = -1, nProj = -1;
  int geom = -1, window = -1;
  double sad = 0.0, cutoff = 0.0, q = 0.0;
  std::vector<float> angles;
};
std::mutex g_scratch_mu;  // guards g_scratch (host side only)
ScratchState g_scratch[kMaxDevices];

// Looks up the device address of `scratch` once per device, with that
// device current.  The caller holds g_scratch_mu.
int scratch_address_locked(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(MEXCT_EINVAL, "device %d beyond the %d the scratch table covers", device, kMaxDevices);
  ScratchState& s = g_scratch[device];
  if (!s.ptr) {
    DevGuard g(device);
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(MEXCT_EHIP, "hipGetSymbolAddress(scratch) on device %d failed (%s)", device,
                  hipGetErrorString(e));
    }
    s.ptr = (unsigned char*)p;
  }
  *out = s.ptr;
  return MEXCT_OK;
}

int scratch_address(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return scratch_address_locked(device, out);
}

}  // namespace

struct mexct_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  struct Buf {
    void* p = nullptr;
    size_t cap = 0;
  } ws[WS_COUNT];
  // per-view (sin, cos) table cache, keyed by its host contents
  std::vector<float2> sc_key;
  float2* sc_dev = nullptr;
  size_t sc_cap = 0;
  void* pinned = nullptr;
  size_t pinned_cap = 0;
  // FDK filter cache
  int f_W = -1, f_window = -1;
  double f_cut = 0.0;
  int w_W = -1, w_H = -1, w_geom = -1;
  double w_sad = 0.0;
  std::vector<float> dA_key;  // angle-step table last uploaded to WS_DA
  // Parker window cache (WS_PARKER), keyed by everything it depends on
  std::vector<float> pk_angles;
  int64_t pk_W = -1, pk_n = -1;
  int pk_geom = -1;
  double pk_sad = 0.0, pk_q = -1.0;
  int autotune = -1;  // -1: follow mexct_set_autotune; 0 off; 1 tune unseen shapes

  int get(int slot, size_t bytes, void** out) {
    Buf& b = ws[slot];
    if (bytes == 0) bytes = 256;
    if (b.cap < bytes) {
      if (b.p) {
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
      }
      size_t cap = (bytes + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
      hipError_t e = hipMalloc(&b.p, cap);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return fail(MEXCT_ENOMEM, "device allocation of %zu bytes failed (%s)", cap,
                    hipGetErrorString(e));
      }
      b.cap = cap;
    }
    *out = b.p;
    return MEXCT_OK;
  }

  // Stage `bytes` of host data through pinned memory into device memory.
  int upload(void* dst, const void* src, size_t bytes) {
    if (pinned_cap < bytes) {
      HIPCHK(hipStreamSynchronize(stream));
      if (pinned) HIPCHK(hipHostFree(pinned));
      pinned = nullptr;
      pinned_cap = 0;
      size_t cap = std::max<size_t>(bytes, 1 << 16);
      HIPCHK(hipHostMalloc(&pinned, cap, hipHostMallocDefault));
      pinned_cap = cap;
    } else {
      HIPCHK(hipStreamSynchronize(stream));  // the previous upload may still read it
    }
    memcpy(pinned, src, bytes);
    HIPCHK(hipMemcpyAsync(dst, pinned, bytes, hipMemcpyHostToDevice, stream));
    return MEXCT_OK;
  }

  // Device table of (sinf, cosf) of each angle, from host sincosf on the
  // float angle as the reference's host loop (CUDAmex_FP.cu:447-449).
  int sincos_table(const float* angles, int n, const float2** out) {
    std::vector<float2> h((size_t)std::max(n, 0));
    for (int i = 0; i < n; ++i) sincosf(angles[i], &h[i].x, &h[i].y);
    if (sc_dev && h.size() == sc_key.size() &&
        (h.empty() || memcmp(h.data(), sc_key.data(), h.size() * sizeof(float2)) == 0)) {
      *out = sc_dev;
      return MEXCT_OK;
    }
    if (sc_cap < h.size() || !sc_dev) {
      if (sc_dev) {
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipFree(sc_dev));
      }
      sc_dev = nullptr;
      size_t cap = std::max<size_t>(h.size(), 1024);
      HIPCHK(hipMalloc((void**)&sc_dev, cap * sizeof(float2)));
      sc_cap = cap;
    }
    if (!h.empty()) RC(upload(sc_dev, h.data(), h.size() * sizeof(float2)));
    sc_key.swap(h);
    *out = sc_dev;
    return MEXCT_OK;
  }

  ~mexct_ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& b : ws)
      if (b.p) (void)hipFree(b.p);
    if (sc_dev) (void)hipFree(sc_dev);
    if (pinned) (void)hipHostFree(pinned);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int ctx_init(mexct_ctx* c, int device, void* stream) {
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fail(MEXCT_EINVAL, "device %d out of range (%d devices)", device, ndev);
  c->device = device;
  DevGuard g(device);
  unsigned char* sp;  // set-up of the fixed scratch, once per device
  RC(scratch_address(device, &sp));
  if (stream) {
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
  } else {
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
  }
  return MEXCT_OK;
}

// Implicit per-thread context of the host-pointer entry points.
thread_local std::vector<std::unique_ptr<mexct_ctx>> t_ctx;

int default_ctx(mexct_ctx** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if ((int)t_ctx.size() <= dev) t_ctx.resize(dev + 1);
  if (!t_ctx[dev]) {
    std::unique_ptr<mexct_ctx> c(new mexct_ctx());
    RC(ctx_init(c.get(), dev, nullptr));
    t_ctx[dev] = std::move(c);
  }
  *out = t_ctx[dev].get();
  return MEXCT_OK;
}

int check_geom(int geom) {
  if (geom < 0 || geom > 2) return fail(MEXCT_EGEOM, "invalid geometry");
  return MEXCT_OK;
}

int check_dtype(int dt, const char* what) {
  if (dt != MEXCT_F32 && dt != MEXCT_F16 && dt != MEXCT_BF16)
    return fail(MEXCT_EINVAL, "%s: unknown dtype %d", what, dt);
  return MEXCT_OK;
}

// Padded operands are indexed with 32-bit element offsets: signed (below
// 2^31 haloed elements) by default, unsigned where `limit` says so.
int check_padded(int64_t a, int64_t b, int64_t c, const char* what,
                 double limit = 2147483648.0) {
  double n = (double)(a + 2 * HALO) * (double)(b + 2 * HALO) * (double)(c + 2 * HALO);
  if (n >= limit) return fail(MEXCT_EINVAL, "%s too large (%g padded elements)", what, n);
  return MEXCT_OK;
}

int storage_of(const mexct_opts* o) { return o ? o->storage : MEXCT_F32; }
// FPZLaunch::long_ray from the options: 0 keeps the kernel's default
int long_ray_of(const mexct_opts* o) {
  if (!o || o->long_ray == 0) return 0;
  return o->long_ray < 0 ? INT_MAX : o->long_ray;
}

// z-fastest FP volume layout for storage st (see fpz.hip): DUAL for f32,
// QUAD for 2-byte storage while the copy stays below 4 GiB (32-bit buffer
// offsets); past that the same copies read through 64-bit addresses
// (DUAL64 / QUAD64, below 2^32 elements); else -1 (the x-fastest
// projector).  Measured at 384^3 x 512 cone: f32 PLAIN 47.6 ms, DUAL
// 32.3 ms; f16 PLAIN 50.8 ms, DUAL 24.1 ms, QUAD 20.1 ms.
// MEXCT_FP_LAYOUT=wide64 takes the 64-bit forms below 4 GiB as well, and
// MEXCT_FP_LAYOUT=plain the one-value-per-element copy (below 2 GiB), so
// that every address path can be compared with the default on one volume.
int fp_layout(int64_t W, int64_t H, int st, size_t* bytes) {
  const size_t es = (size_t)store_bytes(st);
  // the z-fastest projector's 24-bit index multiplies need (W+4)(H+4) < 2^23
  if ((W + 4) * (H + 4) >= ((int64_t)1 << 23)) return -1;
  const char* force = getenv("MEXCT_FP_LAYOUT");
  if (force && !strcmp(force, "plain")) {
    *bytes = (size_t)(W + 4) * (W + 4) * (H + 4) * es;
    return *bytes < ((size_t)1 << 31) ? LAYOUT_PLAIN : -1;
  }
  const bool force64 = force && !strcmp(force, "wide64");
  const int want = st == MEXCT_F32 ? LAYOUT_DUAL : LAYOUT_QUAD;
  const size_t elems = want == LAYOUT_DUAL ? (size_t)(W + 3) * (W + 4) * (H + 4)
                                           : (size_t)(W + 3) * (W + 3) * (H + 4);
  *bytes = elems * (want == LAYOUT_DUAL ? 2 : 4) * es;
  if (*bytes < ((size_t)1 << 32) && !force64) return want;
  if (elems < ((size_t)1 << 32)) return want == LAYOUT_DUAL ? LAYOUT_DUAL64 : LAYOUT_QUAD64;
  return -1;
}

hipError_t build_fp_volume(int layout, const void* in, int in_dt, int64_t sy, int64_t sz,
                           void* out, int st, int64_t W, int64_t H, int mode, hipStream_t s,
                           const uint32_t* scale) {
  layout = base_layout(layout);  // the 64-bit layouts hold the same copies
  if (layout == LAYOUT_DUAL)
    return launch_pad3d(in, in_dt, sy, sz, out, st, (int)W, (int)W, (int)H, mode, s, scale);
  if (layout == LAYOUT_QUAD)
    return launch_pad3q(in, in_dt, sy, sz, out, st, (int)W, (int)W, (int)H, mode, s, scale);
  return launch_pad3z(in, in_dt, sy, sz, out, st, (int)W, (int)W, (int)H, mode, s, scale);
}

// Device words holding max|x| of each of `items` arrays of n floats, `stride`
// floats apart (as float bits), from slot k of WS_MAXBITS on, for the
// power-of-two scaling of FP16 copies.  Slots: 1 the OSC volume, then
// kMaxItems each for FP volumes and BP sinograms.
constexpr int kMaxItems = 256;  // batch items per launch
constexpr int kSlotVol = 64, kSlotSino = 64 + kMaxItems;
int max_bits(mexct_ctx* c, const float* x, size_t n, int k, uint32_t** out, int items = 1,
             size_t stride = 0, const void* const* ptrs = nullptr) {
  void* m;
  RC(c->get(WS_MAXBITS, 4 * (64 + 2 * kMaxItems), &m));
  uint32_t* p = (uint32_t*)m + k;
  HIPCHK(launch_fill((float*)p, 0.f, (size_t)items, c->stream));
  for (int g = 0; g < items; ++g)
    HIPCHK(launch_absmax(ptrs ? (const float*)ptrs[g] : x + g * stride, n, p + g, c->stream));
  *out = p;
  return MEXCT_OK;
}

// A batch's per-item output pointers (host array of device pointers) as a
// device array in the workspace; null when the outputs are one array.
int upload_ptrs(mexct_ctx* c, float* const* ptrs, int items, float* const** dev) {
  *dev = nullptr;
  if (!ptrs) return MEXCT_OK;
  void* d;
  RC(c->get(WS_OUTPTRS, sizeof(float*) * (size_t)kMaxItems, &d));
  HIPCHK(hipMemcpyAsync(d, ptrs, sizeof(float*) * (size_t)items, hipMemcpyHostToDevice, c->stream));
  *dev = (float* const*)d;
  return MEXCT_OK;
}

int tv_steps_of(const mexct_opts* o) { return (o && o->tv_steps > 0) ? o->tv_steps : 20; }

int pick_zrange(const mexct_opts* o, int geom, int64_t H) {
  if (o && (o->zrange == 32 || o->zrange == 48 || o->zrange == 64)) return o->zrange;
  (void)geom;
  // 48 at 384 rows: 8 ranges, 18432 blocks at 384^3 fill the SIMDs in whole
  // rounds (64: 13824 blocks, 4.5 rounds); 1-3% faster than 64 there
  return H >= 256 ? 48 : 32;
}

// ---- autotuner of the kernels' block order
//
// Candidates differ only in which workgroup computes which rays or voxels
// (FP: views per chunk of the block order; BP: super-tile side, and the
// z-range for FP32 storage, whose row starts replay the same running sum
// whatever the range), so every candidate writes the same bits.  The tuning
// call of a problem shape times each candidate once on the first
// kTuneViews (BP) or kTuneViewsFP (FP) views of its own data, keeps the
// fastest for later calls of the shape, and then runs the whole call with
// it.  The timed launches cost about candidates x those views / nProj of
// the call.
// One choice per shape serves every batch size, so that batches of varying
// size do not each pay for tuning.
struct TuneKey {
  int op, geom, st;
  int64_t W, H, n;
  bool operator<(const TuneKey& o) const {
    return std::tie(op, geom, st, W, H, n) < std::tie(o.op, o.geom, o.st, o.W, o.H, o.n);
  }
};
struct TuneVal {
  int a = -1, b = -1;  // FP: vchunk_log2; BP: super-tile side, z-range
  float ms = 0.f;      // the winner's time
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;
std::map<TuneKey, int> g_seen;  // calls per untuned shape
std::atomic<int> g_autotune{0};

bool tune_find(const TuneKey& k, TuneVal* v) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(k);
  if (it == g_tune.end()) return false;
  *v = it->second;
  return true;
}

void tune_store(const TuneKey& k, const TuneVal& v) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[k] = v;
}

// Whether this call of an untuned shape should tune it: mode 1 tunes at the
// first call, mode 2 at the second (so that one-off shapes pay nothing).
bool want_tune(const mexct_ctx* c, const TuneKey& k) {
  const int mode = c->autotune >= 0 ? c->autotune : g_autotune.load();
  if (mode <= 0) return false;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return ++g_seen[k] >= mode;
}

constexpr int kTuneViews = 32;     // BP candidates
constexpr int kTuneViewsFP = 128;  // FP: the view chunk's L2 reuse needs several chunks

// Times run(i) for each candidate i on the context's stream and returns the
// fastest index; candidate `dflt` (the built-in choice) is kept unless
// another is at least 3% faster, so that timing noise does not move it.
template <typename F>
int tune_pick(mexct_ctx* c, int ncand, int dflt, F run, float* best_ms) {
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(MEXCT_EHIP, "hipEventCreate failed");
  }
  int best = 0;
  float bt = 1e30f, dt = 1e30f;
  int rc = MEXCT_OK;
  for (int i = 0; i < ncand && rc == MEXCT_OK; ++i) {
    for (int rep = 0; rep < 1 && rc == MEXCT_OK; ++rep) {
      if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fail(MEXCT_EHIP, "event record failed"); break; }
      rc = run(i);
      if (rc) break;
      float ms = 0.f;
      if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(MEXCT_EHIP, "event timing failed");
        break;
      }
      if (i == dflt) dt = ms;
      if (ms < bt) {
        bt = ms;
        best = i;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (dflt >= 0 && dflt < ncand && !(bt < 0.97f * dt)) {
    best = dflt;
    bt = dt;
  }
  *best_ms = bt;
  return rc ? rc : best;  // error codes are negative
}

// ---- FDK filter design (preFilterSinogram.m:95-152), host FP64
int filter_order(int len) {
  int o = 1;
  while (o < 2 * len) o <<= 1;
  return std::max(64, o);
}

int design_filter(int window, int len, double d, std::vector<double>& full) {
  const int order = filter_order(len);
  full.assign(order, 1.0);
  if (window == 5) return order;
  if (window < 0 || window > 4) return -1;
  const int half = order / 2;
  // band-limited ramp impulse response, Kak & Slaney eq. 61 (:117-121)
  std::vector<double> imp(order, 0.0);
  imp[0] = 0.25;
  for (int n = 1; n <= half; n += 2) {
    double v = -1.0 / ((M_PI * n) * (M_PI * n));
    imp[n] = v;
    if (n < half) imp[order - n] = v;
  }
  // 2*real(fft(imp)) on bins 0..half (:122-123)
  std::vector<double> filt(half + 1);
  for (int j = 0; j <= half; ++j) {
    double s = 0.0;
    for (int m = 0; m < order; ++m) {
      if (imp[m] == 0.0) continue;
      long long jm = ((long long)j * m) % order;
      s += imp[m] * cos(2.0 * M_PI * (double)jm / (double)order);
    }
    filt[j] = 2.0 * s;
  }
  for (int j = 1; j <= half; ++j) {
    const double w = 2.0 * M_PI * j / order;
    switch (window) {
      case 1: filt[j] *= sin(w / (2 * d)) / (w / (2 * d)); break;
      case 2: filt[j] *= cos(w / (2 * d)); break;
      case 3: filt[j] *= (.54 + .46 * cos(w / d)); break;
      case 4: filt[j] *= (1 + cos(w / d)) / 2; break;
      default: break;
    }
  }
  for (int j = 0; j <= half; ++j)
    if (2.0 * M_PI * j / order > M_PI * d) filt[j] = 0.0;
  for (int j = 0; j <= half; ++j) full[j] = filt[j];
  for (int j = half + 1; j < order; ++j) full[j] = filt[order - j];
  return order;
}

// Circular impulse response of the (real, even) filter and its taps
// g[d + W - 1] = h[d mod order], d in [-(W-1), W-1].
void conv_taps(const std::vector<double>& full, int W, std::vector<double>& g) {
  const int order = (int)full.size();
  std::vector<double> h(order);
  const int half = order / 2;
  for (int m = 0; m < order; ++m) {
    double s = full[0] + ((m & 1) ? -full[half] : full[half]);
    for (int j = 1; j < half; ++j) {
      long long jm = ((long long)j * m) % order;
      s += 2.0 * full[j] * cos(2.0 * M_PI * (double)jm / (double)order);
    }
    h[m] = s / (double)order;
  }
  g.resize(2 * W - 1);
  for (int d = -(W - 1); d <= W - 1; ++d) g[d + W - 1] = h[((d % order) + order) % order];
}

}  // namespace

namespace mexct {
hipStream_t ctx_stream(const mexct_ctx* c) { return c->stream; }
int set_error(int code, const char* msg) { return fail(code, "%s", msg); }

// Mirrors the slots fp_run, bp_run, prefilter_run and mexct_osc_tv_recon_dev
// take from the context (mexct_ctx::get rounds each up to 2 MiB).
// Rows of halo above and below the BP's row-grouped sinogram copy, enough
// that no in-circle voxel's detector row needs the kernel's clamp (the
// halo's edge copies are the clamped taps): par and fan rows stay in
// [0, H-1]; a cone voxel at z sits on row v0 + (z - v0) U, U = SAD / (SAD - s),
// |s| < hw + 1 for the in-circle columns, so within v0 (Umax - 1) of
// [0, H-1], plus the running sum's drift and a margin.  *zclamp: a source
// within 5% of the circle, or a copy that would pass 2 GiB per view, keeps
// the kernel's clamp and the 2-row halo; so does f32 storage, whose
// backprojector ran no faster without the clamp (384^3 x 512 cone: 15.8
// against 15.7 ms at SAD 1000, 16.9 against 16.3 at SAD 500, the taller copy
// costing what the clamp saved), and so does a halo adding more than a
// third of the copy's rows (f16 at SAD 1000, a quarter more rows: 9.5
// against 9.8 ms; at SAD 500, 62% more: 10.2 against 9.9 ms).
int bp_row_halo(int64_t W, int64_t H, int geom, float sad, int rows, int st, bool* zclamp) {
  int zh = HALO;
  *zclamp = st == MEXCT_F32;
  if (*zclamp) return zh;
  if (geom == GEOM_CONE) {
    const double r = (double)(W / 2 - 1) + 1.0, v0 = (double)(H - 1) / 2.0;
    const double ext = (double)sad > 1.05 * r ? v0 * ((double)sad / ((double)sad - r) - 1.0) + 4.0
                                              : 1e30;
    if (ext < (double)H + 8.0)
      zh = std::max(HALO, (int)std::ceil(ext));
    else
      *zclamp = true;
  }
  if (2 * (zh - HALO) > (H + 3) / 3) {
    zh = HALO;
    *zclamp = true;
  }
  if ((double)(W + 4) * (double)(H + 2 * zh - 1) * rows * store_bytes(st) >= 2147483648.0) {
    zh = HALO;
    *zclamp = true;
  }
  return zh;
}

size_t op_workspace_bytes(int op, int geom, int st, int64_t W, int64_t H, int64_t nProj, float sad,
                          int items, int min_subsets) {
  auto slot = [](size_t b) {
    if (b == 0) b = 256;
    return (b + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
  };
  const size_t es = (size_t)store_bytes(st);
  const size_t n = (size_t)std::max<int64_t>(nProj, 1);
  const size_t nitems = (size_t)std::max(items, 1);
  const size_t common = slot(4 * (64 + 2 * kMaxItems)) + std::max<size_t>(n, 1024) * sizeof(float2);
  size_t b = common;
  if (op == MEXCT_OP_FP) {
    size_t lbytes = 0;
    const int layout = fp_layout(W, H, st, &lbytes);
    b += layout >= 0 ? slot(lbytes * nitems) : slot((size_t)(W + 4) * (W + 4) * (H + 4) * es);
    return b;
  }
  if (op == MEXCT_OP_BP || op == MEXCT_OP_FDK) {
    const double hw = (double)(W / 2 - 1);
    const int rows = st != MEXCT_F32 && (geom != 2 || (double)sad > 2.02 * hw) ? 4 : 2;
    bool zc;
    const int zh = bp_row_halo(W, H, geom, sad, rows, st, &zc);
    b += slot((size_t)(W + 4) * (H + 2 * zh - 1) * rows * es * n * nitems);
    if (geom == 2 && st == MEXCT_F32) {
      size_t cfv = 0;  // the largest chain of any tuning candidate
      for (int sup : {2, 4, 8})
        for (int zr : {32, 48, 64})
          cfv = std::max(cfv, bp_chain_floats_per_view((int)W, (int)H, zr, sup));
      if (cfv > 0) {
        const size_t cv = std::min<size_t>(n, std::max<size_t>(1, ((size_t)1 << 28) / cfv));
        b += slot(cv * cfv * 4);
      }
    }
    if (op == MEXCT_OP_FDK)
      b += slot(sizeof(double) * (size_t)(2 * W - 1)) +
           (geom != GEOM_PAR ? slot(sizeof(float) * (size_t)W * H) : 0) + slot(sizeof(float) * n);
    return b;
  }
  // OSC-TV: raised and zeroed pre/post, their working copies, the padded
  // state and its previous value, the plain volume, the TV gradient, the
  // reduction partials and scalars, the ratio-term copies of the largest
  // subset and the z-fastest projected volume
  const size_t psz = (size_t)W * H * n * 4, vsz = (size_t)W * W * H * 4;
  const size_t padsz = (size_t)(W + 4) * (W + 4) * (H + 4) * 4;
  const size_t cmax = (n + (size_t)std::max(min_subsets, 1) - 1) / (size_t)std::max(min_subsets, 1);
  const size_t pv = (size_t)2 * (W + 4) * (H + 3);
  const size_t pbytes = (pv * cmax * es + 3) / 4 * 4;
  size_t lbytes = 0;
  (void)fp_layout(W, H, st, &lbytes);
  b += 6 * slot(psz) + 2 * slot(padsz) + 2 * slot(vsz) +
       slot(tv_partials_count((int)W, (int)W, (int)H) * 8) + slot(256) + 2 * slot(pbytes) +
       slot(lbytes);
  return b;
}
}  // namespace mexct

// =============================================================== C ABI
extern "C" {

int mexct_version(void) { return 10000; }

const char* mexct_last_error(void) { return g_err.c_str(); }

int mexct_ctx_create(mexct_ctx** ctx, int device, void* stream) {
  if (!ctx) return fail(MEXCT_EINVAL, "ctx is null");
  std::unique_ptr<mexct_ctx> c(new mexct_ctx());
  RC(ctx_init(c.get(), device, stream));
  *ctx = c.release();
  return MEXCT_OK;
}

int mexct_ctx_destroy(mexct_ctx* ctx) {
  if (!ctx) return MEXCT_OK;
  DevGuard g(ctx->device);
  delete ctx;
  return MEXCT_OK;
}

int mexct_ctx_sync(mexct_ctx* ctx) {
  if (!ctx) return fail(MEXCT_EINVAL, "ctx is null");
  DevGuard g(ctx->device);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MEXCT_OK;
}

int64_t mexct_ctx_workspace_bytes(const mexct_ctx* ctx) {
  if (!ctx) return 0;
  int64_t s = 0;
  for (auto& b : ctx->ws) s += (int64_t)b.cap;
  return s + (int64_t)ctx->sc_cap * 8;
}

// ------------------------------------------------------------------ FP
namespace {
// FP of `items` volumes of one geometry stored one after another (vol_dtype
// elements), outputs one after another: one launch for all of them.
// in_ptrs / out_ptrs (host arrays of `items` device pointers, or null):
// item i's volume and output, in place of vol and out stored one after
// another.
int fp_run(mexct_ctx* c, const void* vol, int32_t vol_dtype, int items, int64_t W, int64_t H,
           float sad, const float* angles, int32_t nProj, int32_t geom, float* out,
           const mexct_opts* opts, bool loads_only = false, const void* const* in_ptrs = nullptr,
           float* const* out_ptrs = nullptr) {
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  RC(check_geom(geom));
  if (W < 1 || H < 1 || nProj < 0) return fail(MEXCT_EINVAL, "bad dims W=%lld H=%lld nProj=%d", (long long)W, (long long)H, nProj);
  if (items < 1 || items > kMaxItems) return fail(MEXCT_EINVAL, "batch of %d volumes (1..%d)", items, kMaxItems);
  if (nProj > 0 && ((!vol && !in_ptrs) || !angles || (!out && !out_ptrs)))
    return fail(MEXCT_EINVAL, "null array argument");
  RC(check_dtype(vol_dtype, "vol"));
  const int st = storage_of(opts);
  RC(check_dtype(st, "storage"));
  // below 2^32 haloed elements, the reference's own unsigned 32-bit index
  // range (CUDAmex_FP.cu:394): the z-fastest copies past 4 GiB read through
  // 64-bit addresses from 32-bit element indices, and the x-fastest
  // projector's element offsets are unsigned 32-bit
  RC(check_padded(W, W, H, "volume", 4294967296.0));
  if (nProj == 0) return MEXCT_OK;
  DevGuard g(c->device);
  void* pad;
  const size_t pbytes = (size_t)(W + 4) * (W + 4) * (H + 4) * store_bytes(st);
  size_t lbytes = 0;
  const int layout = fp_layout(W, H, st, &lbytes);
  const float2* sc;
  const size_t vel = (size_t)W * W * H;  // elements of one input volume
  if (loads_only && (layout < 0 || wide64(layout)))
    return fail(MEXCT_ENOTSUP, "no loads-only form for a volume copy past 4 GiB");
  if (layout >= 0) {
    // z-fastest haloed volumes, one wave per 64 detector rows of a column
    RC(c->get(WS_VOL, lbytes * items, &pad));
    uint32_t* vmax = nullptr;
    if (st == MEXCT_F16 && vol_dtype == MEXCT_F32)
      RC(max_bits(c, (const float*)vol, vel, kSlotVol, &vmax, items, vel, in_ptrs));
    const size_t in_es = (size_t)store_bytes(vol_dtype);
    for (int i = 0; i < items; ++i)
      HIPCHK(build_fp_volume(layout, in_ptrs ? in_ptrs[i] : (const char*)vol + i * vel * in_es,
                             vol_dtype, W, W * W,
                             (char*)pad + i * lbytes, st, W, H, ADDR_CLAMP, c->stream,
                             vmax ? vmax + i : nullptr));
    RC(c->sincos_table(angles, nProj, &sc));
    FPZLaunch z;
    z.vol = pad;
    z.st_dt = st;
    z.layout = layout;
    z.vol_max = vmax;
    z.out = out;
    z.sc = sc;
    z.geom = geom;
    z.W = (int)W;
    z.H = (int)H;
    z.nviews = nProj;
    z.sad = sad;
    z.long_ray = long_ray_of(opts);
    z.items = items;
    z.vol_stride = lbytes;
    z.out_stride = (uint64_t)W * H * nProj;
    RC(upload_ptrs(c, out_ptrs, items, &z.outs));
    z.loads_only = loads_only;
    const TuneKey key{MEXCT_OP_FP, geom, st, W, H, nProj};
    TuneVal tv;
    if (tune_find(key, &tv)) {
      z.vchunk_log2 = tv.a;
    } else if (!loads_only && want_tune(c, key)) {
      static const int cand[] = {2, 3, 4, 5};  // 4 .. 32 views per chunk
      const int best = tune_pick(c, 4, 2, [&](int i) {
        FPZLaunch t = z;
        t.vchunk_log2 = cand[i];
        t.nviews = std::min<int>(nProj, kTuneViewsFP);
        HIPCHK(launch_fpz(t, c->stream));
        return MEXCT_OK;
      }, &tv.ms);
      if (best < 0) return best;
      tv.a = cand[best];
      tune_store(key, tv);
      z.vchunk_log2 = tv.a;
    }
    HIPCHK(launch_fpz(z, c->stream));
    return MEXCT_OK;
  }
  if (items != 1) return fail(MEXCT_ENOTSUP, "volume too large to batch");
  if (loads_only) return fail(MEXCT_ENOTSUP, "no loads-only form for the x-fastest projector");
  RC(c->get(WS_VOL, pbytes, &pad));
  HIPCHK(launch_pad3(in_ptrs ? in_ptrs[0] : vol, vol_dtype, pad, st, (int)W, (int)W, (int)H,
                     ADDR_CLAMP, c->stream));
  RC(c->sincos_table(angles, nProj, &sc));
  FPLaunch a;
  a.vol = pad;
  a.st_dt = st;
  a.out = out_ptrs ? out_ptrs[0] : out;
  a.sc = sc;
  a.geom = geom;
  a.mode = ADDR_CLAMP;
  a.W = (int)W;
  a.H = (int)H;
  a.nviews = nProj;
  a.NX = (int)W;
  a.NY = (int)W;
  a.NZ = (int)H;
  a.sad = sad;
  a.step = geom == GEOM_CONE ? 1.0f : 0.5f;  // CUDAmex_FP.cu:133,231,328
  a.pts = nullptr;
  a.pts_view_stride = 0;
  HIPCHK(launch_fp(a, c->stream));
  return MEXCT_OK;
}
}  // namespace

int mexct_fp_dev(mexct_ctx* c, const void* vol, int32_t vol_dtype, int64_t W, int64_t H,
                 float sad, const float* angles, int32_t nProj, int32_t geom, float* out,
                 const mexct_opts* opts) {
  return fp_run(c, vol, vol_dtype, 1, W, H, sad, angles, nProj, geom, out, opts);
}

int mexct_fp_loads_dev(mexct_ctx* c, const void* vol, int32_t vol_dtype, int64_t W, int64_t H,
                       float sad, const float* angles, int32_t nProj, int32_t geom, float* out,
                       const mexct_opts* opts) {
  return fp_run(c, vol, vol_dtype, 1, W, H, sad, angles, nProj, geom, out, opts, true);
}

int mexct_fp_batch_dev(mexct_ctx* c, const void* vols, int32_t vol_dtype, int32_t count,
                       int64_t W, int64_t H, float sad, const float* angles, int32_t nProj,
                       int32_t geom, float* outs, const mexct_opts* opts) {
  return fp_run(c, vols, vol_dtype, count, W, H, sad, angles, nProj, geom, outs, opts);
}

int mexct_fp(const float* vol, int64_t W, int64_t H, float sad, const float* angles,
             int64_t n_angles, int32_t nProj, int32_t geom, float* out, const mexct_opts* opts) {
  RC(check_geom(geom));
  if (nProj < 0 || n_angles < nProj)
    return fail(MEXCT_EINVAL, "nProj (%d) exceeds the number of angles (%lld)", nProj, (long long)n_angles);
  if (W < 1 || H < 1) return fail(MEXCT_EINVAL, "bad dims");
  if (!vol || (nProj > 0 && (!angles || !out))) return fail(MEXCT_EINVAL, "null array argument");
  mexct_ctx* c;
  RC(default_ctx(&c));
  DevGuard g(c->device);
  const size_t vb = (size_t)W * W * H * 4, ob = (size_t)W * H * nProj * 4;
  void *dv, *dout;
  RC(c->get(WS_HOST_A, vb, &dv));
  RC(c->get(WS_HOST_OUT, ob, &dout));
  HIPCHK(hipMemcpyAsync(dv, vol, vb, hipMemcpyHostToDevice, c->stream));
  RC(mexct_fp_dev(c, dv, MEXCT_F32, W, H, sad, angles, nProj, geom, (float*)dout, opts));
  if (ob) HIPCHK(hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MEXCT_OK;
}

// ------------------------------------------------------------------ BP
namespace {
// BP of `items` sinograms of one geometry stored one after another
// (sino_dtype elements), volumes one after another: one launch per chunk of
// views for all of them.
// in_ptrs / out_ptrs as for fp_run: item i's sinogram and volume.
int bp_run(mexct_ctx* c, const void* sino, int32_t sino_dtype, int items, int64_t W, int64_t H,
           int64_t nProj, float sad, const float* angles, int32_t geom, float* out,
           const mexct_opts* opts, bool loads_only = false, const void* const* in_ptrs = nullptr,
           float* const* out_ptrs = nullptr) {
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  RC(check_geom(geom));
  if (W < 1 || H < 1 || nProj < 0) return fail(MEXCT_EINVAL, "bad dims");
  if (items < 1 || items > kMaxItems) return fail(MEXCT_EINVAL, "batch of %d sinograms (1..%d)", items, kMaxItems);
  if ((!out && !out_ptrs) || (nProj > 0 && ((!sino && !in_ptrs) || !angles)))
    return fail(MEXCT_EINVAL, "null array argument");
  RC(check_dtype(sino_dtype, "sino"));
  const int st = storage_of(opts);
  RC(check_dtype(st, "storage"));
  // any number of views: the row-grouped copy is indexed per view and the
  // kernel takes the views in chunks below 2 GiB (the reference walks them
  // one at a time, CUDAmex_BP.cu:287-291)
  if (nProj > INT32_MAX) return fail(MEXCT_EINVAL, "%lld views", (long long)nProj);
  if (loads_only && geom != GEOM_CONE) return fail(MEXCT_ENOTSUP, "the loads-only BP is cone only");
  // row-grouped haloed copy: (W+4) x (H+3) elements of two values per view,
  // four for 2-byte storage when the rows of consecutive z-voxels step by
  // less than 2 (par, fan; cone with SAD > 2.02 hw, hw = W/2-1, so that
  // U = SAD/(SAD - s) < 1.99 for every voxel in the circle |s| < hw); the
  // kernel takes views in chunks below 2 GiB, so one view must fit
  const double hw = (double)(W / 2 - 1);
  const int rows = st != MEXCT_F32 && (geom != 2 || (double)sad > 2.02 * hw) ? 4 : 2;
  bool zclamp;
  const int zh = bp_row_halo(W, H, geom, sad, rows, st, &zclamp);
  const size_t rp_bytes = (size_t)(W + 4) * (H + 2 * zh - 1) * rows * store_bytes(st);
  if ((double)rp_bytes >= 2147483648.0)
    return fail(MEXCT_EINVAL, "detector too large (%lld x %lld)", (long long)W, (long long)H);
  if ((double)W * W * H >= 4294967296.0) return fail(MEXCT_EINVAL, "volume too large");
  DevGuard g(c->device);
  void* pad;
  const size_t sino_stride = rp_bytes * (size_t)std::max<int64_t>(nProj, 1);
  RC(c->get(WS_SINO, sino_stride * items, &pad));
  uint32_t* smax = nullptr;
  const size_t sel = (size_t)W * H * nProj;  // elements of one input sinogram
  if (st == MEXCT_F16 && sino_dtype == MEXCT_F32 && nProj > 0)
    RC(max_bits(c, (const float*)sino, sel, kSlotSino, &smax, items, sel, in_ptrs));
  const size_t in_es = (size_t)store_bytes(sino_dtype);
  if (nProj > 0)
    for (int i = 0; i < items; ++i)
      HIPCHK(launch_pad2(in_ptrs ? in_ptrs[i] : (const char*)sino + i * sel * in_es, sino_dtype,
                         (char*)pad + i * sino_stride,
                         st, (int)W, (int)H, (int)nProj, ADDR_CLAMP, c->stream,
                         smax ? smax + i : nullptr, rows, zh));
  const float2* sc = nullptr;
  RC(c->sincos_table(angles, (int)nProj, &sc));
  BPLaunch a;
  a.sino = pad;
  a.st_dt = st;
  a.sino_max = smax;
  a.out = out;
  a.sc = sc;
  a.geom = geom;
  a.W = (int)W;
  a.H = (int)H;
  a.nviews = (int)nProj;
  a.sad = sad;
  a.zr = pick_zrange(opts, geom, H);
  a.rows = rows;
  a.zh = zh;
  a.zclamp = zclamp;
  a.chain = nullptr;
  a.chain_views = 0;
  a.items = items;
  a.sino_stride = sino_stride;
  a.out_stride = (uint64_t)W * W * H;
  RC(upload_ptrs(c, out_ptrs, items, &a.outs));
  a.loads_only = loads_only;
  const TuneKey key{MEXCT_OP_BP, geom, st, W, H, nProj};
  TuneVal tv;
  const bool found = tune_find(key, &tv);
  const bool tune = !found && !loads_only && nProj > 0 && want_tune(c, key);
  // candidates: super-tile side 2, 4, 8; the z-range too for FP32 storage
  // unless the caller fixed it
  const bool zr_free =
      st == MEXCT_F32 && !(opts && (opts->zrange == 32 || opts->zrange == 48 || opts->zrange == 64));
  std::vector<std::pair<int, int>> cand;
  for (int sup : {2, 4, 8}) {
    if (zr_free)
      for (int zr : {32, 48, 64}) cand.push_back({sup, zr});
    else
      cand.push_back({sup, a.zr});
  }
  if (found) {
    // a z-range the caller fixed, or one in texel space (2-byte storage,
    // whose row start depends on it), is never replaced by a tuned one
    a.super_ = tv.a;
    if (zr_free) a.zr = tv.b;
  }
  // chain scratch for the largest candidate (super-tile 8, z-range 32)
  const bool chain = geom == 2 && st == MEXCT_F32;
  size_t cfv = 0;
  if (chain) {
    if (tune)
      for (auto& cd : cand) cfv = std::max(cfv, bp_chain_floats_per_view((int)W, (int)H, cd.second, cd.first));
    else
      cfv = bp_chain_floats_per_view((int)W, (int)H, a.zr, a.super_);
  }
  if (cfv > 0 && nProj > 0) {
    // at most 1 GiB of chain scratch; more views run as further launches
    const int64_t cv = std::min<int64_t>(nProj, std::max<int64_t>(1, ((int64_t)1 << 28) / (int64_t)cfv));
    void* ch;
    RC(c->get(WS_CHAIN, (size_t)cv * cfv * 4, &ch));
    a.chain = (float*)ch;
    a.chain_views = (int)cv;
  }
  if (tune) {
    int dflt = 0;  // the built-in choice: super-tile 4 at the default z-range
    for (size_t i = 0; i < cand.size(); ++i)
      if (cand[i].first == 4 && cand[i].second == pick_zrange(opts, geom, H)) dflt = (int)i;
    const int best = tune_pick(c, (int)cand.size(), dflt, [&](int i) {
      BPLaunch t = a;
      t.super_ = cand[i].first;
      t.zr = cand[i].second;
      t.nviews = (int)std::min<int64_t>(nProj, kTuneViews);
      HIPCHK(launch_bp(t, c->stream));
      return MEXCT_OK;
    }, &tv.ms);
    if (best < 0) return best;
    tv.a = cand[best].first;
    tv.b = cand[best].second;
    tune_store(key, tv);
    a.super_ = tv.a;
    if (zr_free) a.zr = tv.b;
  }
  HIPCHK(launch_bp(a, c->stream));
  return MEXCT_OK;
}
}  // namespace

int mexct_bp_dev(mexct_ctx* c, const void* sino, int32_t sino_dtype, int64_t W, int64_t H,
                 int64_t nProj, float sad, const float* angles, int32_t geom, float* out,
                 const mexct_opts* opts) {
  return bp_run(c, sino, sino_dtype, 1, W, H, nProj, sad, angles, geom, out, opts);
}

int mexct_bp_loads_dev(mexct_ctx* c, const void* sino, int32_t sino_dtype, int64_t W, int64_t H,
                       int64_t nProj, float sad, const float* angles, int32_t geom, float* out,
                       const mexct_opts* opts) {
  return bp_run(c, sino, sino_dtype, 1, W, H, nProj, sad, angles, geom, out, opts, true);
}

int mexct_bp_batch_dev(mexct_ctx* c, const void* sinos, int32_t sino_dtype, int32_t count,
                       int64_t W, int64_t H, int64_t nProj, float sad, const float* angles,
                       int32_t geom, float* outs, const mexct_opts* opts) {
  return bp_run(c, sinos, sino_dtype, count, W, H, nProj, sad, angles, geom, outs, opts);
}

extern "C++" {
namespace mexct {
int fp_batch_ptrs(mexct_ctx* c, const void* const* vols, int32_t vol_dtype, int32_t count,
                  int64_t W, int64_t H, float sad, const float* angles, int32_t nProj,
                  int32_t geom, float* const* outs, const mexct_opts* opts) {
  return fp_run(c, nullptr, vol_dtype, count, W, H, sad, angles, nProj, geom, nullptr, opts, false,
                vols, outs);
}
int bp_batch_ptrs(mexct_ctx* c, const void* const* sinos, int32_t sino_dtype, int32_t count,
                  int64_t W, int64_t H, int64_t nProj, float sad, const float* angles,
                  int32_t geom, float* const* outs, const mexct_opts* opts) {
  return bp_run(c, nullptr, sino_dtype, count, W, H, nProj, sad, angles, geom, nullptr, opts, false,
                sinos, outs);
}
}  // namespace mexct
}  // extern "C++"

// ------------------------------------------------------------------ autotuner
int mexct_set_autotune(int32_t mode) {
  if (mode < 0 || mode > 2) return fail(MEXCT_EINVAL, "autotune mode must be 0, 1 or 2");
  g_autotune.store(mode);
  return MEXCT_OK;
}

int mexct_ctx_set_autotune(mexct_ctx* c, int32_t mode) {
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  if (mode < -1 || mode > 2) return fail(MEXCT_EINVAL, "autotune mode must be -1, 0, 1 or 2");
  c->autotune = mode;
  return MEXCT_OK;
}

int mexct_tuned_config(int32_t op, int32_t geom, int32_t storage, int64_t W, int64_t H,
                       int64_t nProj, int32_t* cfg) {
  if (!cfg) return fail(MEXCT_EINVAL, "cfg is null");
  TuneVal v;
  const bool f = tune_find(TuneKey{op, geom, storage, W, H, nProj}, &v);
  cfg[0] = f ? 1 : 0;
  cfg[1] = f ? v.a : -1;
  cfg[2] = f ? v.b : -1;
  cfg[3] = f ? (int32_t)(v.ms * 1000.f) : 0;
  return MEXCT_OK;
}

int mexct_bp(const float* sino, int64_t W, int64_t H, int64_t nProj, float sad,
             const float* angles, int64_t n_angles, int32_t geom, float* out,
             const mexct_opts* opts) {
  RC(check_geom(geom));
  if (W < 1 || H < 1 || nProj < 0) return fail(MEXCT_EINVAL, "bad dims");
  if (n_angles < nProj)
    return fail(MEXCT_EINVAL, "nProj (%lld) exceeds the number of angles (%lld)", (long long)nProj, (long long)n_angles);
  if (!out || (nProj > 0 && (!sino || !angles))) return fail(MEXCT_EINVAL, "null array argument");
  mexct_ctx* c;
  RC(default_ctx(&c));
  DevGuard g(c->device);
  const size_t sb = (size_t)W * H * nProj * 4, ob = (size_t)W * W * H * 4;
  void *ds, *dout;
  RC(c->get(WS_HOST_A, sb, &ds));
  RC(c->get(WS_HOST_OUT, ob, &dout));
  if (sb) HIPCHK(hipMemcpyAsync(ds, sino, sb, hipMemcpyHostToDevice, c->stream));
  RC(mexct_bp_dev(c, ds, MEXCT_F32, W, H, nProj, sad, angles, geom, (float*)dout, opts));
  HIPCHK(hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MEXCT_OK;
}

// ------------------------------------------------------------------ general-3D FP
int mexct_g3d_fp_dev(mexct_ctx* c, const void* vol, int32_t vol_dtype, int64_t side,
                     int64_t nSlices, const float* angles, const float* pts_dev, int64_t W,
                     int64_t H, int32_t nProj, int32_t ppv, float* out, const mexct_opts* opts) {
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  if (side < 1 || nSlices < 1 || W < 1 || H < 1 || nProj < 0) return fail(MEXCT_EINVAL, "bad dims");
  if (nProj > 0 && (!vol || !angles || !pts_dev || !out)) return fail(MEXCT_EINVAL, "null array argument");
  RC(check_dtype(vol_dtype, "vol"));
  const int st = storage_of(opts);
  RC(check_dtype(st, "storage"));
  RC(check_padded(side, side, nSlices, "volume"));
  if (nProj == 0) return MEXCT_OK;
  DevGuard g(c->device);
  void* pad;
  const float2* sc;
  size_t lbytes = 0;
  const int layout = fp_layout(side, nSlices, st, &lbytes);
  if (layout >= 0) {
    // z-fastest volume as for the standard geometries; a wave marches 64
    // consecutive detector rows of one column, whose rays are near each
    // other for any detector whose columns are lines in space
    RC(c->get(WS_VOL, lbytes, &pad));
    uint32_t* vmax = nullptr;
    if (st == MEXCT_F16 && vol_dtype == MEXCT_F32)
      RC(max_bits(c, (const float*)vol, (size_t)side * side * nSlices, 1, &vmax));
    HIPCHK(build_fp_volume(layout, vol, vol_dtype, side, side * side, pad, st, side, nSlices,
                           ADDR_BORDER, c->stream, vmax));
    RC(c->sincos_table(angles, nProj, &sc));
    FPZLaunch z;
    z.vol = pad;
    z.st_dt = st;
    z.layout = layout;
    z.vol_max = vmax;
    z.out = out;
    z.sc = sc;
    z.geom = GEOM_G3D;
    z.W = (int)W;
    z.H = (int)H;
    z.nviews = nProj;
    z.sad = 0.f;
    z.NX = (int)side;
    z.NZ = (int)nSlices;
    z.pts = pts_dev;
    z.pts_view_stride = ppv ? (int64_t)6 * W * H : 0;
    z.long_ray = long_ray_of(opts);
    HIPCHK(launch_fpz(z, c->stream));
    return MEXCT_OK;
  }
  RC(c->get(WS_VOL, (size_t)(side + 4) * (side + 4) * (nSlices + 4) * store_bytes(st), &pad));
  HIPCHK(launch_pad3(vol, vol_dtype, pad, st, (int)side, (int)side, (int)nSlices, ADDR_BORDER, c->stream));
  RC(c->sincos_table(angles, nProj, &sc));
  FPLaunch a;
  a.vol = pad;
  a.st_dt = st;
  a.out = out;
  a.sc = sc;
  a.geom = GEOM_G3D;
  a.mode = ADDR_BORDER;
  a.W = (int)W;
  a.H = (int)H;
  a.nviews = nProj;
  a.NX = (int)side;
  a.NY = (int)side;
  a.NZ = (int)nSlices;
  a.sad = 0.f;
  a.step = 0.5f;  // CUDAmex_general3D_FP.cu:132
  a.pts = pts_dev;
  a.pts_view_stride = ppv ? (int64_t)6 * W * H : 0;
  HIPCHK(launch_fp(a, c->stream));
  return MEXCT_OK;
}

int mexct_g3d_fp(const float* vol, int64_t side, int64_t nSlices, const float* angles,
                 int64_t n_angles, const float* pts, int64_t W, int64_t H, int32_t nProj,
                 int32_t ppv, float* out, const mexct_opts* opts) {
  if (nProj < 0 || n_angles < nProj) return fail(MEXCT_EINVAL, "nProj exceeds the number of angles");
  if (side < 1 || nSlices < 1 || W < 1 || H < 1) return fail(MEXCT_EINVAL, "bad dims");
  if (!vol || !pts || (nProj > 0 && (!angles || !out))) return fail(MEXCT_EINVAL, "null array argument");
  mexct_ctx* c;
  RC(default_ctx(&c));
  DevGuard g(c->device);
  const size_t vb = (size_t)side * side * nSlices * 4;
  const size_t pb = (size_t)W * H * 6 * 4 * (ppv ? std::max(nProj, 1) : 1);
  const size_t ob = (size_t)W * H * nProj * 4;
  void *dv, *dp, *dout;
  RC(c->get(WS_HOST_A, vb, &dv));
  RC(c->get(WS_HOST_B, pb, &dp));
  RC(c->get(WS_HOST_OUT, ob, &dout));
  HIPCHK(hipMemcpyAsync(dv, vol, vb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dp, pts, pb, hipMemcpyHostToDevice, c->stream));
  RC(mexct_g3d_fp_dev(c, dv, MEXCT_F32, side, nSlices, angles, (const float*)dp, W, H, nProj, ppv,
                      (float*)dout, opts));
  if (ob) HIPCHK(hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MEXCT_OK;
}

// ------------------------------------------------------------------ general-3D BP
int mexct_g3d_bp_dev(mexct_ctx* c, const float* sino, int64_t W, int64_t H, int64_t nProj,
                     const float* angles, const float* pts_dev, int32_t ppv, int64_t M, int64_t N,
                     float* out, const mexct_opts* opts) {
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  if (W < 1 || H < 1 || nProj < 0 || M < 1 || N < 1) return fail(MEXCT_EINVAL, "bad dims");
  if (!out || (nProj > 0 && (!sino || !angles || !pts_dev))) return fail(MEXCT_EINVAL, "null array argument");
  const int mode = opts ? opts->g3d_bp : MEXCT_G3D_BRICKS;
  if (mode < MEXCT_G3D_BRICKS || mode > MEXCT_G3D_ATOMICS)
    return fail(MEXCT_EINVAL, "unknown general-3D BP mode %d", mode);
  if ((double)M * M * N >= 2147483648.0 || (double)W * H * nProj >= 4294967296.0)
    return fail(MEXCT_EINVAL, "general-3D BP too large (M=%lld N=%lld)", (long long)M, (long long)N);
  DevGuard g(c->device);
  if (nProj == 0) {
    HIPCHK(launch_fill(out, 0.f, (size_t)M * M * N, c->stream));
    return MEXCT_OK;
  }
  const G3DBrickPlan plan = g3d_brick_plan((int)W, (int)H, (int)M, (int)N, (int)nProj);
  if (mode == MEXCT_G3D_DETERMINISTIC && !plan.ok)
    return fail(MEXCT_ENOTSUP, "deterministic general-3D BP: %u bricks exceed the binning LDS",
                plan.nbricks);
  if (mode != MEXCT_G3D_ATOMICS && plan.ok) {
    G3DBrickLaunch L;
    L.plan = plan;
    L.sino = sino;
    L.pts = pts_dev;
    L.pts_view_stride = ppv ? (int64_t)6 * W * H : 0;
    RC(c->sincos_table(angles, (int)nProj, &L.sc));
    L.W = (int)W;
    L.H = (int)H;
    L.nviews = (int)nProj;
    L.M = (int)M;
    L.N = (int)N;
    void *ck, *mk, *hl, *acc;
    RC(c->get(WS_G3D_CK, plan.ck_bytes, &ck));
    RC(c->get(WS_G3D_MASK, plan.mask_bytes + 256, &mk));  // + the long-ray count
    RC(c->get(WS_G3D_HALO, plan.halo_elems * 8, &hl));
    RC(c->get(WS_G3D_ACC, (size_t)M * M * N * 8, &acc));
    L.mask = (uint32_t*)mk;
    L.long_rays = (uint32_t*)((char*)mk + plan.mask_bytes);
    L.ck = (float*)ck;
    L.out = out;
    L.halo64 = (int64_t*)hl;
    L.out64 = (int64_t*)acc;
    uint32_t* pm;
    RC(max_bits(c, sino, (size_t)W * H * nProj, 2, &pm));
    L.pmax = pm;
    HIPCHK(launch_g3d_bp_bricks(L, c->stream));
    return MEXCT_OK;
  }
  void* acc;
  RC(c->get(WS_TVG, (size_t)M * M * N * 4, &acc));
  HIPCHK(launch_fill((float*)acc, 0.f, (size_t)M * M * N, c->stream));
  const float2* sc;
  RC(c->sincos_table(angles, (int)nProj, &sc));
  G3DBPLaunch a;
  a.sino = sino;
  a.pts = pts_dev;
  a.pts_view_stride = ppv ? (int64_t)6 * W * H : 0;
  a.sc = sc;
  a.out = out;
  a.W = (int)W;
  a.H = (int)H;
  a.nviews = (int)nProj;
  a.M = (int)M;
  a.N = (int)N;
  a.out = (float*)acc;
  HIPCHK(launch_g3d_bp(a, c->stream));
  HIPCHK(launch_g3d_unz((const float*)acc, out, (int)M, (int)N, c->stream));
  return MEXCT_OK;
}

int mexct_g3d_bp(const float* sino, int64_t W, int64_t H, int64_t nProj, const float* angles,
                 int64_t n_angles, const float* pts, int32_t ppv, int64_t M, int64_t N,
                 float* out, const mexct_opts* opts) {
  if (n_angles < nProj) return fail(MEXCT_EINVAL, "nProj exceeds the number of angles");
  if (W < 1 || H < 1 || nProj < 0 || M < 1 || N < 1) return fail(MEXCT_EINVAL, "bad dims");
  if (!out || !pts || (nProj > 0 && (!sino || !angles))) return fail(MEXCT_EINVAL, "null array argument");
  mexct_ctx* c;
  RC(default_ctx(&c));
  DevGuard g(c->device);
  const size_t sb = (size_t)W * H * nProj * 4;
  const size_t pb = (size_t)W * H * 6 * 4 * (ppv ? std::max<int64_t>(nProj, 1) : 1);
  const size_t ob = (size_t)M * M * N * 4;
  void *ds, *dp, *dout;
  RC(c->get(WS_HOST_A, sb, &ds));
  RC(c->get(WS_HOST_B, pb, &dp));
  RC(c->get(WS_HOST_OUT, ob, &dout));
  if (sb) HIPCHK(hipMemcpyAsync(ds, sino, sb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dp, pts, pb, hipMemcpyHostToDevice, c->stream));
  RC(mexct_g3d_bp_dev(c, (const float*)ds, W, H, nProj, angles, (const float*)dp, ppv, M, N,
                      (float*)dout, opts));
  HIPCHK(hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MEXCT_OK;
}

// ------------------------------------------------------------------ OSC
namespace {

// One CUDAmex_oscIter pass over the subsets on the device state: mu_pad is
// the zero-haloed volume, pre/post [W,H,nProj] are rewritten in place.
// One CUDAmex_oscIter call on the padded f32 state mu_pad.  The subset
// projections read a z-fastest copy of mu_pad in storage st (FP16: scaled by
// a power of two from max|mu|), and the dual backprojection reads the ratio
// terms in storage st (BF16 when st is FP16: their range exceeds FP16's).
int osc_subsets(mexct_ctx* c, float* mu_pad, int W, int H, float* pre, float* post, int nProj,
                int numSubsets, const int32_t* picks, const float2* sc, float sad, int geom,
                int st) {
  const int cmax = (nProj + numSubsets - 1) / numSubsets;
  const size_t pv = (size_t)2 * (W + 4) * (H + 3);  // row-paired copies
  const size_t es = (size_t)store_bytes(st);
  const int pst = st == MEXCT_F16 ? MEXCT_BF16 : st;  // ratio-term storage
  // whole 32-bit words so the zero fill covers every 2-byte element
  const size_t pwords = (pv * std::max(cmax, 1) * es + 3) / 4;
  void *pp, *qp;
  RC(c->get(WS_PREPAD, pwords * 4, &pp));
  RC(c->get(WS_POSTPAD, pwords * 4, &qp));
  // halos stay zero: the kernels only ever write the interior
  HIPCHK(launch_fill((float*)pp, 0.f, pwords, c->stream));
  HIPCHK(launch_fill((float*)qp, 0.f, pwords, c->stream));
  void* mz;
  const int64_t MPX = W + 4, MPXY = (int64_t)(W + 4) * (W + 4);
  size_t lbytes = 0;
  const int layout = fp_layout(W, H, st, &lbytes);
  if (layout < 0)
    return fail(MEXCT_EINVAL,
                "OSC: the projected volume copy needs (W+4)*(H+4) = %lld below 2^23 and fewer "
                "than 2^32 elements",
                (long long)(W + 4) * (H + 4));
  RC(c->get(WS_MUZ, lbytes, &mz));
  const float* interior = mu_pad + 2 + 2 * MPX + 2 * MPXY;
  for (int s = 0; s < numSubsets; ++s) {
    const int sel = picks[s];
    const int count = (sel >= 0 && sel < nProj) ? (nProj - 1 - sel) / numSubsets + 1 : 0;
    if (count > 0) {
      // the subset projects the volume as left by the previous update
      uint32_t* vmax = nullptr;
      if (st == MEXCT_F16) RC(max_bits(c, mu_pad, (size_t)MPXY * (H + 4), 1, &vmax));
      HIPCHK(build_fp_volume(layout, interior, MEXCT_F32, MPX, MPXY, mz, st, W, H, ADDR_BORDER,
                             c->stream, vmax));
      OSCFPLaunch f;
      f.mu_z = mz;
      f.st_dt = st;
      f.layout = layout;
      f.vol_max = vmax;
      f.pre = pre;
      f.post = post;
      f.pre_pad = pp;
      f.post_pad = qp;
      f.sc = sc;
      f.geom = geom;
      f.W = W;
      f.H = H;
      f.nProj = nProj;
      f.first = sel;
      f.stride = numSubsets;
      f.count = count;
      f.sad = sad;
      HIPCHK(launch_osc_fp(f, c->stream));
    }
    OSCBPLaunch b;
    b.pre_pad = pp;
    b.post_pad = qp;
    b.st_dt = pst;
    b.mu_pad = mu_pad;
    b.sc = sc;
    b.geom = geom;
    b.W = W;
    b.H = H;
    b.first = sel;
    b.stride = numSubsets;
    b.count = count;
    b.sad = sad;
    HIPCHK(launch_osc_bp_update(b, c->stream));
  }
  return MEXCT_OK;
}

int osc_check(int64_t W, int64_t H, int64_t nProj, int32_t numSubsets, int32_t geom) {
  RC(check_geom(geom));
  if (W < 2 || H < 1 || nProj < 0) return fail(MEXCT_EINVAL, "bad dims");
  if (numSubsets < 1) return fail(MEXCT_EINVAL, "numSubsets must be >= 1 (got %d)", numSubsets);
  RC(check_padded(W, W, H, "volume"));
  if ((double)(W + 4) * (W + 4) * (H + 4) * 4.0 >= 2147483648.0)
    return fail(MEXCT_EINVAL, "volume too large for the OSC state (%lld x %lld x %lld)",
                (long long)W, (long long)W, (long long)H);
  return MEXCT_OK;
}

}  // namespace

int mexct_osc_iter_dev(mexct_ctx* c, float* mu, int64_t W, int64_t H, float* pre, float* post,
                       int64_t nProj, int32_t numSubsets, const int32_t* picks,
                       const float* angles, float sad, int32_t geom, const mexct_opts* opts) {
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  RC(osc_check(W, H, nProj, numSubsets, geom));
  const int st = storage_of(opts);
  RC(check_dtype(st, "storage"));
  if (!mu || !picks || (nProj > 0 && (!pre || !post || !angles)))
    return fail(MEXCT_EINVAL, "null array argument");
  DevGuard g(c->device);
  void* mp;
  RC(c->get(WS_MUPAD, (size_t)(W + 4) * (W + 4) * (H + 4) * 4, &mp));
  HIPCHK(launch_pad3(mu, MEXCT_F32, mp, MEXCT_F32, (int)W, (int)W, (int)H, ADDR_BORDER, c->stream));
  const float2* sc = nullptr;
  RC(c->sincos_table(angles, (int)nProj, &sc));
  RC(osc_subsets(c, (float*)mp, (int)W, (int)H, pre, post, (int)nProj, numSubsets, picks, sc, sad,
                 geom, st));
  HIPCHK(launch_unpad3((const float*)mp, mu, (int)W, (int)W, (int)H, c->stream));
  return MEXCT_OK;
}

int mexct_osc_iter(float* mu, int64_t W, int64_t H, float* pre, float* post, int64_t nProj,
                   int32_t numSubsets, const int32_t* picks, int64_t n_picks, const float* angles,
                   int64_t n_angles, float sad, int32_t geom, const mexct_opts* opts) {
  RC(osc_check(W, H, nProj, numSubsets, geom));
  if (n_picks < numSubsets) return fail(MEXCT_EINVAL, "fewer subset picks than subsets");
  if (n_angles < nProj) return fail(MEXCT_EINVAL, "nProj exceeds the number of angles");
  if (!mu || !picks || (nProj > 0 && (!pre || !post || !angles)))
    return fail(MEXCT_EINVAL, "null array argument");
  mexct_ctx* c;
  RC(default_ctx(&c));
  DevGuard g(c->device);
  const size_t vb = (size_t)W * W * H * 4, pb = (size_t)W * H * nProj * 4;
  void *dm, *dpre, *dpost;
  RC(c->get(WS_HOST_A, vb, &dm));
  RC(c->get(WS_HOST_B, pb, &dpre));
  RC(c->get(WS_HOST_C, pb, &dpost));
  HIPCHK(hipMemcpyAsync(dm, mu, vb, hipMemcpyHostToDevice, c->stream));
  if (pb) {
    HIPCHK(hipMemcpyAsync(dpre, pre, pb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dpost, post, pb, hipMemcpyHostToDevice, c->stream));
  }
  RC(mexct_osc_iter_dev(c, (float*)dm, W, H, (float*)dpre, (float*)dpost, nProj, numSubsets, picks,
                        angles, sad, geom, opts));
  HIPCHK(hipMemcpyAsync(mu, dm, vb, hipMemcpyDeviceToHost, c->stream));
  if (pb) {
    HIPCHK(hipMemcpyAsync(pre, dpre, pb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(post, dpost, pb, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return MEXCT_OK;
}

// ------------------------------------------------------------------ TV
namespace {
int tv_run(mexct_ctx* c, Vol3 x, const float* c_dev, float cc, int steps) {
  void *gp, *pp, *sp;
  RC(c->get(WS_TVG, (size_t)x.nx * x.ny * x.nz * 4, &gp));
  RC(c->get(WS_PARTIALS, tv_partials_count(x.nx, x.ny, x.nz) * 8, &pp));
  RC(c->get(WS_SCALARS, 256, &sp));
  HIPCHK(launch_tvmin(x, (float*)gp, (double*)pp, (float*)sp, c_dev, cc, steps, c->stream));
  return MEXCT_OK;
}
}  // namespace

int mexct_tvmin_dev(mexct_ctx* c, float* vol, int64_t Nx, int64_t Ny, int64_t Nz, float cc,
                    const mexct_opts* opts) {
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  if (Nx < 1 || Ny < 1 || Nz < 1) return fail(MEXCT_EINVAL, "bad dims");
  if (!vol) return fail(MEXCT_EINVAL, "null array argument");
  if (Nz > 65535) return fail(MEXCT_EINVAL, "Nz too large");
  DevGuard g(c->device);
  Vol3 x{vol, (int)Nx, (int)Ny, (int)Nz, (int64_t)Nx, (int64_t)Nx * Ny};
  return tv_run(c, x, nullptr, cc, tv_steps_of(opts));
}

int mexct_tvmin(const float* vol, int64_t Nx, int64_t Ny, int64_t Nz, float cc, float* out,
                const mexct_opts* opts) {
  if (Nx < 1 || Ny < 1 || Nz < 1) return fail(MEXCT_EINVAL, "bad dims");
  if (!vol || !out) return fail(MEXCT_EINVAL, "null array argument");
  mexct_ctx* c;
  RC(default_ctx(&c));
  DevGuard g(c->device);
  const size_t vb = (size_t)Nx * Ny * Nz * 4;
  void* dv;
  RC(c->get(WS_HOST_A, vb, &dv));
  HIPCHK(hipMemcpyAsync(dv, vol, vb, hipMemcpyHostToDevice, c->stream));
  RC(mexct_tvmin_dev(c, (float*)dv, Nx, Ny, Nz, cc, opts));
  HIPCHK(hipMemcpyAsync(out, dv, vb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MEXCT_OK;
}

// ------------------------------------------------------------------ FDK filter
namespace {

bool parker_on(double q) { return q >= 0.0; }  // negative or NaN: the 5-argument call

// The Parker / Wesarg window of the first nProj views (parker_weight.m:29-76)
// in WS_PARKER, as [nProj][W] floats at *win.  n_all >= nProj angles set the
// scan direction and extent (diff(angles), angles(end)).  The geometry terms
// are formed here in MATLAB's single/double mix -- alpha and delta double,
// anything combined with the single angles single -- and the window itself
// on the device.  alpha runs along the detector columns (the fan direction);
// the reference takes size(sino,2) entries and expands with repmat (:32,:73),
// the same window when W == H.
size_t parker_col_bytes(int64_t W) { return (size_t)W * sizeof(float4); }
size_t parker_beta_bytes(int64_t nProj) { return ((size_t)nProj * sizeof(float) + 255) & ~(size_t)255; }

// Host part of the window: the per-column terms and per-view betas, laid
// out as parker_col_bytes(W) of float4 then parker_beta_bytes(nProj) of
// float in h, and the scan constants k.
void parker_tables(int64_t W, int64_t nProj, int64_t n_all, int geom, double sad,
                   const float* angles, double q, std::vector<char>& h, ParkerConsts& k) {
  const size_t col_b = parker_col_bytes(W), beta_b = parker_beta_bytes(nProj);
  std::vector<double> alpha((size_t)W, 0.0);
  if (geom != GEOM_PAR)
    for (int64_t u = 0; u < W; ++u) alpha[u] = atan((-(double)W / 2 + 0.5 + (double)u) / sad);
  bool increasing = true;
  for (int64_t i = 1; i < n_all; ++i)
    if (!(angles[i] - angles[i - 1] >= 0.0f)) increasing = false;
  if (increasing)
    for (double& a : alpha) a = -a;
  const double delta = fabs(alpha[W - 1] - alpha[0]) / 2;
  const float extent = fabsf(angles[n_all - 1] - angles[0]);
  k.pi_f = (float)M_PI;
  k.d2 = (float)(2 * delta);
  k.eps = std::max(extent - (float)(M_PI + 2 * delta), 0.0f);
  const float qf = (float)q;
  h.assign(col_b + beta_b, 0);
  float4* col = (float4*)h.data();
  for (int64_t u = 0; u < W; ++u)
    col[u] = make_float4((float)(2 * alpha[u]), qf * ((float)(2 * delta - 2 * alpha[u]) + k.eps),
                         qf * ((float)(2 * delta + 2 * alpha[u]) + k.eps), 0.f);
  float* beta = (float*)(h.data() + col_b);
  for (int64_t i = 0; i < nProj; ++i) beta[i] = fabsf(angles[i] - angles[0]);
}

int parker_window(mexct_ctx* c, int64_t W, int64_t nProj, int64_t n_all, int geom, double sad,
                  const float* angles, double q, const float** win) {
  const size_t col_b = parker_col_bytes(W), beta_b = parker_beta_bytes(nProj);
  void* base;
  RC(c->get(WS_PARKER, col_b + beta_b + (size_t)W * nProj * sizeof(float), &base));
  float* w = (float*)((char*)base + col_b + beta_b);
  *win = w;
  const bool same = c->pk_W == W && c->pk_n == nProj && c->pk_geom == geom && c->pk_sad == sad &&
                    c->pk_q == q && c->pk_angles.size() == (size_t)n_all &&
                    memcmp(c->pk_angles.data(), angles, (size_t)n_all * sizeof(float)) == 0;
  if (same) return MEXCT_OK;
  std::vector<char> h;
  ParkerConsts k;
  parker_tables(W, nProj, n_all, geom, sad, angles, q, h, k);
  c->pk_W = -1;  // invalid until the kernel is queued
  RC(c->upload(base, h.data(), h.size()));
  HIPCHK(launch_parker((const float4*)base, (const float*)((char*)base + col_b), k, (int)W,
                       (int)nProj, w, c->stream));
  c->pk_angles.assign(angles, angles + n_all);
  c->pk_W = W;
  c->pk_n = nProj;
  c->pk_geom = geom;
  c->pk_sad = sad;
  c->pk_q = q;
  return MEXCT_OK;
}

// Cosine weights Distance = SAD ./ sqrt(SAD^2 + U.^2 [+ V.^2]) with
// U = (1..W) - (W-1)/2 (preFilterSinogram.m:40-42, :54-57), rounded to
// single, as w[u + W*v] (fan and cone).
void cosine_weights(int64_t W, int64_t H, int geom, double sad, std::vector<float>& w) {
  w.resize((size_t)W * H);
  for (int64_t v = 0; v < H; ++v)
    for (int64_t u = 0; u < W; ++u) {
      const double U = (double)(u + 1) - (double)(W - 1) / 2.0;
      const double V = (double)(v + 1) - (double)(H - 1) / 2.0;
      const double D = geom == GEOM_CONE ? sad / sqrt(sad * sad + U * U + V * V)
                                         : sad / sqrt(sad * sad + U * U);
      w[(size_t)u + (size_t)W * v] = (float)D;
    }
}

// (angles(i) - angles(i-1)) in single; view 1 takes the next step (:80-84).
void angle_steps(int64_t nProj, const float* angles, std::vector<float>& dA) {
  dA.resize((size_t)nProj);
  for (int64_t i = 0; i < nProj; ++i) dA[i] = i > 0 ? angles[i] - angles[i - 1] : angles[1] - angles[0];
}

int check_prefilter_args(int32_t geom, int64_t W, int64_t H, int64_t nProj, int32_t window) {
  RC(check_geom(geom));
  if (W < 1 || H < 1) return fail(MEXCT_EINVAL, "bad dims");
  if (nProj < 2) return fail(MEXCT_EINVAL, "the angle step needs at least 2 views");
  if (window < 0 || window > 5) return fail(MEXCT_EINVAL, "unknown filter window %d", window);
  if (3 * W - 1 > 16384) return fail(MEXCT_EINVAL, "detector too wide for the LDS filter (W=%lld)", (long long)W);
  if ((double)H * nProj >= 2147483648.0) return fail(MEXCT_EINVAL, "too many detector rows (H*nProj)");
  return MEXCT_OK;
}

// preFilterSinogram.m:29-90 on in.sino (or on -log(in.post ./ in.sino)),
// the first nProj of n_all angles.
int prefilter_run(mexct_ctx* c, FilterIn in, int64_t W, int64_t H, int64_t nProj, int64_t n_all,
                  int32_t geom, double sad, const float* angles, int32_t window, double cutoff,
                  double parker_q, float* out) {
  void *gp, *dp;
  RC(c->get(WS_FILT, sizeof(double) * (2 * W - 1), &gp));
  if (c->f_W != W || c->f_window != window || c->f_cut != cutoff) {
    std::vector<double> full, taps;
    design_filter(window, (int)W, cutoff, full);
    conv_taps(full, (int)W, taps);
    RC(c->upload(gp, taps.data(), taps.size() * sizeof(double)));
    c->f_W = (int)W;
    c->f_window = window;
    c->f_cut = cutoff;
  }
  in.wgt = nullptr;
  if (geom != GEOM_PAR) {
    void* wp;
    RC(c->get(WS_WGT, sizeof(float) * W * H, &wp));
    if (c->w_W != W || c->w_H != H || c->w_geom != geom || c->w_sad != sad) {
      std::vector<float> w;
      cosine_weights(W, H, geom, sad, w);
      RC(c->upload(wp, w.data(), w.size() * sizeof(float)));
      c->w_W = (int)W;
      c->w_H = (int)H;
      c->w_geom = geom;
      c->w_sad = sad;
    }
    in.wgt = (const float*)wp;
  }
  in.pw = nullptr;
  if (parker_on(parker_q)) RC(parker_window(c, W, nProj, n_all, geom, sad, angles, parker_q, &in.pw));
  std::vector<float> dA;
  angle_steps(nProj, angles, dA);
  const bool have = c->ws[WS_DA].p && c->dA_key.size() == dA.size() &&
                    memcmp(c->dA_key.data(), dA.data(), dA.size() * sizeof(float)) == 0;
  RC(c->get(WS_DA, sizeof(float) * nProj, &dp));
  if (!have) {
    RC(c->upload(dp, dA.data(), dA.size() * sizeof(float)));
    c->dA_key.swap(dA);
  }
  HIPCHK(launch_prefilter(in, out, (const double*)gp, (const float*)dp, (int)W, (int)H,
                          (int)nProj, c->stream));
  return MEXCT_OK;
}

}  // namespace

int mexct_prefilter_dev(mexct_ctx* c, const float* sino, int64_t W, int64_t H, int64_t nProj,
                        int32_t geom, double sad, const float* angles, int32_t window,
                        double cutoff, double parker_q, float* out) {
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  RC(check_prefilter_args(geom, W, H, nProj, window));
  if (!sino || !out || !angles) return fail(MEXCT_EINVAL, "null array argument");
  DevGuard g(c->device);
  return prefilter_run(c, FilterIn{sino, nullptr, nullptr, nullptr}, W, H, nProj, nProj, geom,
                       sad, angles, window, cutoff, parker_q, out);
}

int mexct_prefilter(const float* sino, int64_t W, int64_t H, int64_t nProj, int32_t geom,
                    double sad, const float* angles, int64_t n_angles, int32_t window,
                    double cutoff, double parker_q, float* out) {
  if (n_angles < nProj) return fail(MEXCT_EINVAL, "nProj exceeds the number of angles");
  RC(check_prefilter_args(geom, W, H, nProj, window));
  if (!sino || !out || !angles) return fail(MEXCT_EINVAL, "null array argument");
  mexct_ctx* c;
  RC(default_ctx(&c));
  DevGuard g(c->device);
  const size_t sb = (size_t)W * H * nProj * 4;
  void *ds, *dout;
  RC(c->get(WS_HOST_A, sb, &ds));
  RC(c->get(WS_HOST_OUT, sb, &dout));
  HIPCHK(hipMemcpyAsync(ds, sino, sb, hipMemcpyHostToDevice, c->stream));
  RC(prefilter_run(c, FilterIn{(const float*)ds, nullptr, nullptr, nullptr}, W, H, nProj,
                   n_angles, geom, sad, angles, window, cutoff, parker_q, (float*)dout));
  HIPCHK(hipMemcpyAsync(out, dout, sb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MEXCT_OK;
}

// ------------------------------------------------------------------ allocation-free pre-filter
namespace {
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte offsets of the pre-filter tables inside `scratch`.
struct ScratchLayout {
  size_t taps, wgt, dA, pk, win, total;
};
ScratchLayout scratch_layout(int64_t W, int64_t H, int64_t nProj, int geom, bool parker) {
  ScratchLayout l;
  size_t off = 0;
  l.taps = off;
  off += align256(sizeof(double) * (size_t)(2 * W - 1));
  l.wgt = off;
  if (geom != GEOM_PAR) off += align256(sizeof(float) * (size_t)W * H);
  l.dA = off;
  off += align256(sizeof(float) * (size_t)nProj);
  l.pk = off;
  l.win = off;
  if (parker) {
    l.win = off + parker_col_bytes(W) + parker_beta_bytes(nProj);
    off = l.win + align256(sizeof(float) * (size_t)W * nProj);
  }
  l.total = off;
  return l;
}
}  // namespace

int mexct_prefilter_noalloc_dev(mexct_ctx* c, const float* sino, int64_t W, int64_t H,
                                int64_t nProj, int32_t geom, double sad, const float* angles,
                                int32_t window, double cutoff, double parker_q, float* out) {
  RC(check_prefilter_args(geom, W, H, nProj, window));
  const bool pk = parker_on(parker_q);
  const ScratchLayout l = scratch_layout(W, H, nProj, geom, pk);
  if (l.total > kScratchBytes)
    return fail(MEXCT_EINVAL,
                "the filter tables need %zu bytes; the allocation-free pre-filter has %zu "
                "(W=%lld H=%lld nProj=%lld%s)",
                l.total, kScratchBytes, (long long)W, (long long)H, (long long)nProj,
                pk ? ", Parker window" : "");
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  if (!sino || !out || !angles) return fail(MEXCT_EINVAL, "null array argument");
  DevGuard g(c->device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  unsigned char* base;
  RC(scratch_address_locked(c->device, &base));
  ScratchState& s = g_scratch[c->device];
  const bool same = s.valid && s.W == W && s.H == H && s.nProj == nProj && s.geom == geom &&
                    s.window == window && s.sad == sad && s.cutoff == cutoff &&
                    (pk ? s.q == parker_q : !parker_on(s.q)) && s.angles.size() == (size_t)nProj &&
                    memcmp(s.angles.data(), angles, (size_t)nProj * sizeof(float)) == 0;
  if (!same) {
    // The tables of mexct_prefilter_dev, copied from pageable host memory;
    // the stream is synchronised before the host copies go out of scope.
    s.valid = false;
    std::vector<double> full, taps;
    design_filter(window, (int)W, cutoff, full);
    conv_taps(full, (int)W, taps);
    HIPCHK(hipMemcpyAsync(base + l.taps, taps.data(), taps.size() * sizeof(double),
                          hipMemcpyHostToDevice, c->stream));
    std::vector<float> w, dA;
    if (geom != GEOM_PAR) {
      cosine_weights(W, H, geom, sad, w);
      HIPCHK(hipMemcpyAsync(base + l.wgt, w.data(), w.size() * sizeof(float),
                            hipMemcpyHostToDevice, c->stream));
    }
    angle_steps(nProj, angles, dA);
    HIPCHK(hipMemcpyAsync(base + l.dA, dA.data(), dA.size() * sizeof(float),
                          hipMemcpyHostToDevice, c->stream));
    std::vector<char> h;
    if (pk) {
      ParkerConsts k;
      parker_tables(W, nProj, nProj, geom, sad, angles, parker_q, h, k);
      HIPCHK(hipMemcpyAsync(base + l.pk, h.data(), h.size(), hipMemcpyHostToDevice, c->stream));
      HIPCHK(launch_parker((const float4*)(base + l.pk),
                           (const float*)(base + l.pk + parker_col_bytes(W)), k, (int)W,
                           (int)nProj, (float*)(base + l.win), c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    s.W = W;
    s.H = H;
    s.nProj = nProj;
    s.geom = geom;
    s.window = window;
    s.sad = sad;
    s.cutoff = cutoff;
    s.q = pk ? parker_q : MEXCT_NO_PARKER;
    s.angles.assign(angles, angles + nProj);
    s.valid = true;
  }
  FilterIn in{sino, nullptr, geom != GEOM_PAR ? (const float*)(base + l.wgt) : nullptr,
              pk ? (const float*)(base + l.win) : nullptr};
  HIPCHK(launch_prefilter(in, out, (const double*)(base + l.taps), (const float*)(base + l.dA),
                          (int)W, (int)H, (int)nProj, c->stream));
  return MEXCT_OK;
}

// ------------------------------------------------------------------ Parker weights
namespace {
int check_parker_args(int32_t geom, int64_t W, int64_t H, int64_t nProj, double sad, double q) {
  RC(check_geom(geom));
  if (W < 1 || H < 1 || nProj < 1) return fail(MEXCT_EINVAL, "bad dims");
  if (!parker_on(q)) return fail(MEXCT_EINVAL, "the Parker q must be >= 0");
  if (geom != GEOM_PAR && !(sad > 0.0)) return fail(MEXCT_EINVAL, "SAD must be > 0");
  return MEXCT_OK;
}
}  // namespace

int mexct_parker_weight_dev(mexct_ctx* c, const float* sino, int64_t W, int64_t H, int64_t nProj,
                            const float* angles, int32_t geom, double sad, double q, float* out) {
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  RC(check_parker_args(geom, W, H, nProj, sad, q));
  if (!sino || !out || !angles) return fail(MEXCT_EINVAL, "null array argument");
  DevGuard g(c->device);
  const float* win;
  RC(parker_window(c, W, nProj, nProj, geom, sad, angles, q, &win));
  HIPCHK(launch_apply_window(sino, win, out, (int)W, (int)H, (int)nProj, c->stream));
  return MEXCT_OK;
}

int mexct_parker_weight(const float* sino, int64_t W, int64_t H, int64_t nProj,
                        const float* angles, int64_t n_angles, int32_t geom, double sad, double q,
                        float* out) {
  if (n_angles < nProj) return fail(MEXCT_EINVAL, "nProj exceeds the number of angles");
  RC(check_parker_args(geom, W, H, nProj, sad, q));
  if (!sino || !out || !angles) return fail(MEXCT_EINVAL, "null array argument");
  mexct_ctx* c;
  RC(default_ctx(&c));
  DevGuard g(c->device);
  const size_t sb = (size_t)W * H * nProj * 4;
  void *ds, *dout;
  RC(c->get(WS_HOST_A, sb, &ds));
  RC(c->get(WS_HOST_OUT, sb, &dout));
  HIPCHK(hipMemcpyAsync(ds, sino, sb, hipMemcpyHostToDevice, c->stream));
  const float* win;
  RC(parker_window(c, W, nProj, n_angles, geom, sad, angles, q, &win));
  HIPCHK(launch_apply_window((const float*)ds, win, (float*)dout, (int)W, (int)H, (int)nProj,
                             c->stream));
  HIPCHK(hipMemcpyAsync(out, dout, sb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MEXCT_OK;
}

// ------------------------------------------------------------------ FBP
namespace {
int fbp_run(mexct_ctx* c, const float* pre, const float* post, int64_t W, int64_t H,
            int64_t nProj, int64_t n_all, const float* angles, int32_t geom, double sad,
            int32_t window, double cutoff, double parker_q, float* vol, const mexct_opts* opts) {
  void* fs;
  RC(c->get(WS_FILTERED, (size_t)W * H * nProj * 4, &fs));
  RC(prefilter_run(c, FilterIn{pre, post, nullptr, nullptr}, W, H, nProj, n_all, geom, sad,
                   angles, window, cutoff, parker_q, (float*)fs));
  // CUDAmex_BP(filteredSino, geom.SAD, scanAngles, geomFlag) (FBP_recon.m:32-38)
  return mexct_bp_dev(c, fs, MEXCT_F32, W, H, nProj, (float)sad, angles, geom, vol, opts);
}
}  // namespace

int mexct_fbp_recon_dev(mexct_ctx* c, const float* pre, const float* post, int64_t W, int64_t H,
                        int64_t nProj, const float* angles, int32_t geom, double sad,
                        int32_t window, double cutoff, double parker_q, float* vol,
                        const mexct_opts* opts) {
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  RC(check_prefilter_args(geom, W, H, nProj, window));
  if (!pre || !post || !vol || !angles) return fail(MEXCT_EINVAL, "null array argument");
  DevGuard g(c->device);
  return fbp_run(c, pre, post, W, H, nProj, nProj, angles, geom, sad, window, cutoff, parker_q,
                 vol, opts);
}

int mexct_fbp_recon(const float* pre, const float* post, int64_t W, int64_t H, int64_t nProj,
                    const float* angles, int64_t n_angles, int32_t geom, double sad,
                    int32_t window, double cutoff, double parker_q, float* vol,
                    const mexct_opts* opts) {
  if (n_angles < nProj) return fail(MEXCT_EINVAL, "nProj exceeds the number of angles");
  RC(check_prefilter_args(geom, W, H, nProj, window));
  if (!pre || !post || !vol || !angles) return fail(MEXCT_EINVAL, "null array argument");
  mexct_ctx* c;
  RC(default_ctx(&c));
  DevGuard g(c->device);
  const size_t sb = (size_t)W * H * nProj * 4, vb = (size_t)W * W * H * 4;
  void *dpre, *dpost, *dv;
  RC(c->get(WS_HOST_A, sb, &dpre));
  RC(c->get(WS_HOST_B, sb, &dpost));
  RC(c->get(WS_HOST_OUT, vb, &dv));
  HIPCHK(hipMemcpyAsync(dpre, pre, sb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dpost, post, sb, hipMemcpyHostToDevice, c->stream));
  RC(fbp_run(c, (const float*)dpre, (const float*)dpost, W, H, nProj, n_angles, angles, geom, sad,
             window, cutoff, parker_q, (float*)dv, opts));
  HIPCHK(hipMemcpyAsync(vol, dv, vb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MEXCT_OK;
}

// ------------------------------------------------------------------ measurement
int mexct_probe_gather_dev(mexct_ctx* c, int32_t storage, int64_t W, int64_t H, int32_t nviews,
                           int32_t pattern, int64_t* samples) {
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  RC(check_dtype(storage, "storage"));
  if (W < 2 || H < 1 || nviews < 1 || (pattern & ~0x31) != 0 || !samples)
    return fail(MEXCT_EINVAL, "bad probe arguments");
  const size_t es = (size_t)store_bytes(storage);
  const int kind = pattern & 1;
  const size_t bytes = kind != MEXCT_PROBE_FP ? (size_t)(W + 4) * (H + 3) * 2 * es * (size_t)nviews
                       : storage == MEXCT_F32  ? (size_t)(W + 3) * (W + 4) * (H + 4) * 2 * es
                                               : (size_t)(W + 3) * (W + 3) * (H + 4) * 4 * es;
  if (bytes >= ((size_t)1 << 32)) return fail(MEXCT_EINVAL, "probe buffer too large");
  DevGuard g(c->device);
  void *buf, *sink;
  RC(c->get(kind == MEXCT_PROBE_FP ? WS_VOL : WS_SINO, bytes, &buf));
  RC(c->get(WS_SCALARS, 4096, &sink));
  HIPCHK(launch_gather_probe(buf, storage, pattern, (int)W, (int)H, nviews, (float*)sink, samples,
                             c->stream));
  return MEXCT_OK;
}

// ------------------------------------------------------------------ OSC-TV driver
int mexct_calc_num_subsets(int32_t sInit, int32_t sFinal, double p, int32_t K, int32_t k) {
  // round((sInit - sFinal)/(K-1)^p * (K-k)^p + sFinal)  (OSC_TV_recon.m:323-325)
  const double v = (double)(sInit - sFinal) / pow((double)(K - 1), p) * pow((double)(K - k), p) +
                   (double)sFinal;
  return (int)round(v);
}

int mexct_calc_subset_order(int32_t n, int32_t first_pick, int32_t* picks) {
  // calcSubsetOrder (OSC_TV_recon.m:327-349): start at the drawn subset, then
  // repeatedly take the middle of the larger remaining side.
  if (n < 1 || !picks) return fail(MEXCT_EINVAL, "bad subset count");
  if (first_pick < 1 || first_pick > n) return fail(MEXCT_EINVAL, "first pick %d not in 1..%d", first_pick, n);
  std::vector<int> nums(n);
  for (int i = 0; i < n; ++i) nums[i] = i + 1;
  int cur = first_pick;
  for (int s = 1; s <= n; ++s) {
    if (s > 1) {
      std::vector<int> lo, hi;
      for (int v : nums) {
        if (v < cur) lo.push_back(v);
        if (v > cur) hi.push_back(v);
      }
      std::vector<int>& side = lo.size() > hi.size() ? lo : hi;
      cur = side[(side.size() + 1) / 2 - 1];
      side.erase(std::find(side.begin(), side.end(), cur));
      nums = lo;
      nums.insert(nums.end(), hi.begin(), hi.end());
    }
    picks[s - 1] = cur - 1;
  }
  return MEXCT_OK;
}

namespace {
__global__ void clamp_min_kernel(const float* __restrict__ in, float* __restrict__ out,
                                 float* __restrict__ out2, size_t n) {
  // pre(pre<epsilon) = epsilon, epsilon = 1e-6 (OSC_TV_recon.m:282-284);
  // out2 = copy with element 1 set to 0 (pre2(1) = 0, :288)
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    float v = in[i];
    if ((double)v < 1e-6) v = (float)1e-6;
    out[i] = v;
    out2[i] = i == 0 ? 0.0f : v;
  }
}
}  // namespace

namespace {
int osc_tv_args(int32_t geom, int64_t W, int64_t H, int64_t nProj, int32_t nIter, int32_t k_begin,
                 int32_t k_end) {
  RC(check_geom(geom));
  if (W < 2 || H < 1 || nProj < 1) return fail(MEXCT_EINVAL, "bad dims");
  if (nIter < 2) return fail(MEXCT_EINVAL, "nIter must be >= 2 (calcNumSubsets divides by (nIter-1)^p)");
  if (k_begin < 1 || k_end < k_begin || k_end > nIter)
    return fail(MEXCT_EINVAL, "iterations %d..%d outside 1..%d", k_begin, k_end, nIter);
  RC(check_padded(W, W, H, "volume"));
  return MEXCT_OK;
}

// Iterations k_begin..k_end of the nIter-iteration OSC_TV_recon.m loop on
// the device.  The state between iterations is mu alone (pre/post are
// re-derived from the inputs each iteration, the subset order from
// first_picks), so mu_in = the volume after iteration k_begin - 1 resumes
// the run exactly; k_begin = 1 starts from 1e-10 and ignores mu_in.
int osc_tv_run(mexct_ctx* c, const float* pre, const float* post, int64_t W, int64_t H,
               int64_t nProj, const float* angles, float sad, int32_t geom, int32_t nIter, float tv,
               int32_t sInit, int32_t sFinal, double p, const int32_t* first_picks,
               int32_t k_begin, int32_t k_end, const float* mu_in, float* mu_out,
               const mexct_opts* opts) {
  if (!c) return fail(MEXCT_EINVAL, "ctx is null");
  RC(osc_tv_args(geom, W, H, nProj, nIter, k_begin, k_end));
  if (!pre || !post || !angles || !first_picks || !mu_out || (k_begin > 1 && !mu_in))
    return fail(MEXCT_EINVAL, "null array argument");
  const int st_dt = storage_of(opts);
  RC(check_dtype(st_dt, "storage"));
  DevGuard g(c->device);
  hipStream_t st = c->stream;
  const size_t psz = (size_t)W * H * nProj, vsz = (size_t)W * W * H;
  const size_t padsz = (size_t)(W + 4) * (W + 4) * (H + 4);
  void *pre0, *post0, *pre2, *post2, *prew, *postw, *mp, *mo, *plain, *pp, *sp;
  RC(c->get(WS_PRE0, psz * 4, &pre0));
  RC(c->get(WS_POST0, psz * 4, &post0));
  RC(c->get(WS_PRE2, psz * 4, &pre2));
  RC(c->get(WS_POST2, psz * 4, &post2));
  RC(c->get(WS_PREW, psz * 4, &prew));
  RC(c->get(WS_POSTW, psz * 4, &postw));
  RC(c->get(WS_MUPAD, padsz * 4, &mp));
  RC(c->get(WS_MUOLD, padsz * 4, &mo));
  RC(c->get(WS_PLAIN, vsz * 4, &plain));
  RC(c->get(WS_PARTIALS, tv_partials_count((int)W, (int)W, (int)H) * 8, &pp));
  RC(c->get(WS_SCALARS, 256, &sp));
  const unsigned cg = (unsigned)std::min<size_t>((psz + 255) / 256, 65536);
  clamp_min_kernel<<<cg, 256, 0, st>>>(pre, (float*)pre0, (float*)pre2, psz);
  clamp_min_kernel<<<cg, 256, 0, st>>>(post, (float*)post0, (float*)post2, psz);
  HIPCHK(hipGetLastError());
  if (k_begin == 1) {
    // mu = ones(W,W,H,'single') .* 1e-10
    HIPCHK(launch_fill((float*)plain, (float)(1.0 * 1e-10), vsz, st));
    HIPCHK(launch_pad3(plain, MEXCT_F32, mp, MEXCT_F32, (int)W, (int)W, (int)H, ADDR_BORDER, st));
  } else {
    HIPCHK(launch_pad3(mu_in, MEXCT_F32, mp, MEXCT_F32, (int)W, (int)W, (int)H, ADDR_BORDER, st));
  }
  const float2* sc = nullptr;
  RC(c->sincos_table(angles, (int)nProj, &sc));
  const int64_t MPX = W + 4, MPXY = (W + 4) * (W + 4);
  const int64_t interior = 2 + 2 * MPX + 2 * MPXY;
  Vol3 mu_v{(float*)mp + interior, (int)W, (int)W, (int)H, MPX, MPXY};
  Vol3 old_v{(float*)mo + interior, (int)W, (int)W, (int)H, MPX, MPXY};
  double* sum = (double*)sp;
  float* cdev = (float*)((char*)sp + 64);
  float* adev = (float*)((char*)sp + 128);
  const int steps = tv_steps_of(opts);
  void* gp;
  RC(c->get(WS_TVG, vsz * 4, &gp));
  for (int k = k_begin; k <= k_end; ++k) {
    HIPCHK(hipMemcpyAsync(mo, mp, padsz * 4, hipMemcpyDeviceToDevice, st));
    const int n = mexct_calc_num_subsets(sInit, sFinal, p, nIter, k);
    if (n < 1) return fail(MEXCT_EINVAL, "iteration %d has %d subsets", k, n);
    std::vector<int32_t> picks(n);
    RC(mexct_calc_subset_order(n, first_picks[k - 1], picks.data()));
    float* wpre = (float*)(k == 1 ? pre0 : prew);
    float* wpost = (float*)(k == 1 ? post0 : postw);
    if (k > 1) {
      HIPCHK(hipMemcpyAsync(prew, pre2, psz * 4, hipMemcpyDeviceToDevice, st));
      HIPCHK(hipMemcpyAsync(postw, post2, psz * 4, hipMemcpyDeviceToDevice, st));
    }
    RC(osc_subsets(c, (float*)mp, (int)W, (int)H, wpre, wpost, (int)nProj, n, picks.data(), sc,
                   sad, geom, st_dt));
    if (k > 1) {
      HIPCHK(launch_diff_norm(old_v, mu_v, (double*)pp, sum, st));
      HIPCHK(launch_tv_const(sum, tv, cdev, st));
      HIPCHK(launch_tvmin(mu_v, (float*)gp, (double*)pp, adev, cdev, 0.f, steps, st));
      HIPCHK(launch_clean(mu_v, st));
    }
  }
  HIPCHK(launch_unpad3((const float*)mp, mu_out, (int)W, (int)W, (int)H, st));
  return MEXCT_OK;
}
}  // namespace

int mexct_osc_tv_recon_dev(mexct_ctx* c, const float* pre, const float* post, int64_t W,
                           int64_t H, int64_t nProj, const float* angles, float sad, int32_t geom,
                           int32_t nIter, float tv, int32_t sInit, int32_t sFinal, double p,
                           const int32_t* first_picks, float* mu_out, const mexct_opts* opts) {
  return osc_tv_run(c, pre, post, W, H, nProj, angles, sad, geom, nIter, tv, sInit, sFinal, p,
                    first_picks, 1, std::max<int32_t>(nIter, 1), nullptr, mu_out, opts);
}

int mexct_osc_tv_iterations_dev(mexct_ctx* c, const float* pre, const float* post, int64_t W,
                                int64_t H, int64_t nProj, const float* angles, float sad,
                                int32_t geom, int32_t nIter, float tv, int32_t sInit,
                                int32_t sFinal, double p, const int32_t* first_picks,
                                int32_t k_begin, int32_t k_end, float* mu, const mexct_opts* opts) {
  return osc_tv_run(c, pre, post, W, H, nProj, angles, sad, geom, nIter, tv, sInit, sFinal, p,
                    first_picks, k_begin, k_end, mu, mu, opts);
}

int mexct_osc_tv_iterations(const float* pre, const float* post, int64_t W, int64_t H,
                            int64_t nProj, const float* angles, int64_t n_angles, float sad,
                            int32_t geom, int32_t nIter, float tv, int32_t sInit, int32_t sFinal,
                            double p, const int32_t* first_picks, int32_t k_begin, int32_t k_end,
                            float* mu, const mexct_opts* opts) {
  if (n_angles < nProj) return fail(MEXCT_EINVAL, "nProj exceeds the number of angles");
  RC(osc_tv_args(geom, W, H, nProj, nIter, k_begin, k_end));
  if (!pre || !post || !mu || !angles || !first_picks) return fail(MEXCT_EINVAL, "null array argument");
  mexct_ctx* c;
  RC(default_ctx(&c));
  DevGuard g(c->device);
  const size_t pb = (size_t)W * H * nProj * 4, vb = (size_t)W * W * H * 4;
  void *dpre, *dpost, *dmu;
  RC(c->get(WS_HOST_A, pb, &dpre));
  RC(c->get(WS_HOST_B, pb, &dpost));
  RC(c->get(WS_HOST_OUT, vb, &dmu));
  HIPCHK(hipMemcpyAsync(dpre, pre, pb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dpost, post, pb, hipMemcpyHostToDevice, c->stream));
  if (k_begin > 1) HIPCHK(hipMemcpyAsync(dmu, mu, vb, hipMemcpyHostToDevice, c->stream));
  RC(mexct_osc_tv_iterations_dev(c, (const float*)dpre, (const float*)dpost, W, H, nProj, angles,
                                 sad, geom, nIter, tv, sInit, sFinal, p, first_picks, k_begin,
                                 k_end, (float*)dmu, opts));
  HIPCHK(hipMemcpyAsync(mu, dmu, vb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MEXCT_OK;
}

int mexct_osc_tv_recon(const float* pre, const float* post, int64_t W, int64_t H, int64_t nProj,
                       const float* angles, int64_t n_angles, float sad, int32_t geom,
                       int32_t nIter, float tv, int32_t sInit, int32_t sFinal, double p,
                       const int32_t* first_picks, float* mu_out, const mexct_opts* opts) {
  if (n_angles < nProj) return fail(MEXCT_EINVAL, "nProj exceeds the number of angles");
  if (W < 2 || H < 1 || nProj < 1) return fail(MEXCT_EINVAL, "bad dims");
  if (!pre || !post || !mu_out) return fail(MEXCT_EINVAL, "null array argument");
  mexct_ctx* c;
  RC(default_ctx(&c));
  DevGuard g(c->device);
  const size_t pb = (size_t)W * H * nProj * 4, vb = (size_t)W * W * H * 4;
  void *dpre, *dpost, *dout;
  RC(c->get(WS_HOST_A, pb, &dpre));
  RC(c->get(WS_HOST_B, pb, &dpost));
  RC(c->get(WS_HOST_OUT, vb, &dout));
  HIPCHK(hipMemcpyAsync(dpre, pre, pb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dpost, post, pb, hipMemcpyHostToDevice, c->stream));
  RC(mexct_osc_tv_recon_dev(c, (const float*)dpre, (const float*)dpost, W, H, nProj, angles, sad,
                            geom, nIter, tv, sInit, sFinal, p, first_picks, (float*)dout, opts));
  HIPCHK(hipMemcpyAsync(mu_out, dout, vb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MEXCT_OK;
}

}  // extern "C"
