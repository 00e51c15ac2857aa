// serve.hip -- continuous-batching request server on one GPU.
//
// There is no counterpart in the reference, where each MEX call owns the
// device for its whole duration.  A server owns `slots` workers.  Each
// worker has its own context (stream at default priority, grow-only
// workspace) and its own host thread.  Requests (FP, BP, FDK = prefilter +
// BP, OSC-TV) queue in arrival order.  A free worker admits the oldest
// request whose device footprint fits the HBM budget, so requests that
// arrive while others run start as soon as a worker and memory are free,
// and their kernels overlap on the device.  A worker that admits a request
// also takes up to max_group-1 queued requests with the same key (operator,
// geometry, sizes, angles, SAD, window, storage), as many as the budget
// holds.  FP, BP and FDK groups run as one batch: every kernel of the
// projector or backprojector is one launch for the whole group
// (mexct_fp_batch_dev / mexct_bp_batch_dev), each item computed exactly as
// alone; a batch of host-array requests runs in two halves whose copies
// overlap the other half's kernels.  OSC-TV groups run back to back with
// the tables resident.  Requests marked resident bring device arrays and
// skip the host copies.  The
// workers' contexts autotune the block order of a problem shape at its
// second call.
//
// Memory accounting: every worker reserves max(bytes it holds, footprint
// of the group it runs).  The sum of reservations never exceeds the
// budget.  When the oldest request cannot be admitted and releasing the
// memory held by idle workers would let it fit, the idle workers release
// it.  Later requests may overtake a blocked oldest request at most
// kMaxOvertake times; after that nothing is admitted before it.
//
// The operators are deterministic, so a request's result is the same
// whether it runs alone or among other requests.
#include "../../include/mexct.h"

#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include <hip/hip_runtime.h>

#include "internal.h"

namespace {

using clk = std::chrono::steady_clock;
constexpr int kMaxOvertake = 8;
constexpr size_t kSlack = (size_t)4 << 20;  // host-table caches and allocator rounding

double now_s() {
  return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

struct Req {
  uint64_t id = 0;
  mexct_request r{};
  std::vector<float> angles;
  std::vector<int32_t> picks;
  int64_t footprint = 0;
  double t_submit = 0, t_start = 0, t_done = 0;
  int status = 1;  // 1 queued, 2 running, 0 done, < 0 error code
  int overtaken = 0;
  std::string err;
};

// Upper estimate of the device bytes a request needs on a worker: the
// worker's staging arrays (run_one) plus the workspace the C ABI takes for
// that operator, from the ABI's own layouts (mexct::op_workspace_bytes).
// Per request; a batched group of G is reserved as G times this, which
// bounds the batch's staging and its workspace (only the haloed copies
// scale with G).
int64_t footprint_of(const mexct_request& r) {
  const size_t W = (size_t)r.W, H = (size_t)r.H, P = (size_t)r.nProj;
  const size_t v = W * W * H * 4, s = W * H * P * 4;
  size_t staging = 0;
  int min_subsets = 1;
  switch (r.op) {
    case MEXCT_OP_FP: staging = v + s; break;
    case MEXCT_OP_BP: staging = s + v; break;
    case MEXCT_OP_FDK: staging = 2 * s + v; break;
    default:
      staging = 2 * s + v;
      if (r.osc_iters >= 2) {
        min_subsets = INT32_MAX;
        for (int k = 1; k <= r.osc_iters; ++k)
          min_subsets = std::min(min_subsets, mexct_calc_num_subsets(r.osc_s_init, r.osc_s_final,
                                                                     r.osc_p, r.osc_iters, k));
        min_subsets = std::max(min_subsets, 1);
      }
      break;
  }
  const size_t ws = mexct::op_workspace_bytes(r.op, r.geom, r.storage, r.W, r.H, r.nProj, r.sad, 1,
                                              min_subsets);
  return (int64_t)(staging + ws + kSlack);
}

bool same_key(const Req& a, const Req& b) {
  const mexct_request &x = a.r, &y = b.r;
  return x.op == y.op && x.geom == y.geom && x.storage == y.storage && x.W == y.W &&
         x.H == y.H && x.nProj == y.nProj && x.sad == y.sad && x.window == y.window &&
         x.cutoff == y.cutoff && x.tv_steps == y.tv_steps && x.resident == y.resident &&
         a.angles == b.angles;
}

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

}  // namespace

struct mexct_server {
  struct Worker {
    mexct_ctx* ctx = nullptr;
    std::thread th;
    Buf a, b, o, x;  // staged input, second input, output, filtered sinogram
    int64_t reserved = 0;  // this worker's share of in_use
    bool running = false;
    // pinned host chunks for device -> host copies: a copy into pageable
    // memory runs at about a third of the DMA rate, so outputs go through
    // these in turn, one chunk copied out on the host while the next is in
    // flight (allocated at first use, freed when the server is destroyed)
    void* pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    // a batch of host requests runs in two halves: the copies of one half
    // go on this stream while the other half's kernels run on the context's
    hipStream_t cs = nullptr;
    hipEvent_t ev_in = nullptr, ev_k[2] = {nullptr, nullptr};
  };
  int64_t n_batched = 0, n_batches = 0;  // requests run in multi-item launches, such groups
  int device = 0;
  int64_t budget = 0;
  int64_t in_use = 0;
  int max_group = 1;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  uint64_t next_id = 1;
  bool stop = false;
  std::vector<std::unique_ptr<Worker>> workers;
  int64_t n_done = 0, n_groups = 0, n_grouped = 0, n_releases = 0, peak_in_use = 0;
};

namespace {

using Worker = mexct_server::Worker;

int grow(Buf& b, size_t bytes) {
  if (b.cap >= bytes) return MEXCT_OK;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return mexct::set_error(MEXCT_ENOMEM, "server: device allocation failed");
  }
  b.cap = bytes;
  return MEXCT_OK;
}

int64_t held_bytes(const Worker& w) {
  return (int64_t)(w.a.cap + w.b.cap + w.o.cap + w.x.cap) + mexct_ctx_workspace_bytes(w.ctx);
}

// Frees a worker's staging arrays and context workspace.  The worker's own
// thread calls this while the worker is idle; its stream is drained first.
int new_ctx(mexct_server* sv, Worker& w) {
  const int rc = mexct_ctx_create(&w.ctx, sv->device, nullptr);
  return rc ? rc : mexct_ctx_set_autotune(w.ctx, 2);  // tune shapes that recur
}

int release(mexct_server* sv, Worker& w) {
  (void)mexct_ctx_sync(w.ctx);
  for (Buf* b : {&w.a, &w.b, &w.o, &w.x}) {
    if (b->p) (void)hipFree(b->p);
    b->p = nullptr;
    b->cap = 0;
  }
  mexct_ctx_destroy(w.ctx);
  w.ctx = nullptr;
  return new_ctx(sv, w);
}

int copy_in(Worker& w, void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, mexct::ctx_stream(w.ctx));
  return e == hipSuccess ? MEXCT_OK : mexct::set_error(MEXCT_EHIP, hipGetErrorString(e));
}

constexpr size_t kPinChunk = (size_t)4 << 20;

bool pinned_ready(Worker& w) {
  if (w.pin[1] && w.ev[1]) return true;
  for (int k = 0; k < 2; ++k) {
    if (!w.pin[k] && hipHostMalloc(&w.pin[k], kPinChunk, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      w.pin[k] = nullptr;
      return false;
    }
    if (!w.ev[k] && hipEventCreateWithFlags(&w.ev[k], hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      w.ev[k] = nullptr;
      return false;
    }
  }
  return true;
}

// The copy stream (default priority) and the events of pipelined batches.
int worker_streams(int device, Worker& w) {
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&w.cs, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&w.ev_in, hipEventDisableTiming);
  for (int k = 0; k < 2 && e == hipSuccess; ++k)
    e = hipEventCreateWithFlags(&w.ev_k[k], hipEventDisableTiming);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return mexct::set_error(MEXCT_EHIP, hipGetErrorString(e));
  }
  return MEXCT_OK;
}

void free_streams(Worker& w) {
  if (w.cs) (void)hipStreamSynchronize(w.cs);
  for (hipEvent_t* e : {&w.ev_in, &w.ev_k[0], &w.ev_k[1]}) {
    if (*e) (void)hipEventDestroy(*e);
    *e = nullptr;
  }
  if (w.cs) (void)hipStreamDestroy(w.cs);
  w.cs = nullptr;
}

void free_pinned(Worker& w) {
  for (int k = 0; k < 2; ++k) {
    if (w.pin[k]) (void)hipHostFree(w.pin[k]);
    if (w.ev[k]) (void)hipEventDestroy(w.ev[k]);
    w.pin[k] = nullptr;
    w.ev[k] = nullptr;
  }
}

// Device -> host copies of n (dst, src, bytes) segments on stream st, in
// order after the work already on it.  Returns when every
// output is in host memory.
struct Seg {
  void* dst;
  const void* src;
  size_t bytes;
};
int copy_out_segs(Worker& w, hipStream_t st, const Seg* seg, int n) {
  hipError_t e = hipSuccess;
  size_t total = 0;
  for (int i = 0; i < n; ++i) total += seg[i].bytes;
  if (total < ((size_t)1 << 20) || !pinned_ready(w)) {
    for (int i = 0; i < n && e == hipSuccess; ++i)
      e = hipMemcpyAsync(seg[i].dst, seg[i].src, seg[i].bytes, hipMemcpyDeviceToHost, st);
  } else {
    // chunk c goes to pin[c & 1]; chunk c - 1 is copied out of the other
    // buffer on the host while chunk c is in flight
    char* prev_dst = nullptr;
    size_t prev_len = 0;
    int c = 0;
    for (int i = 0; i < n && e == hipSuccess; ++i) {
      for (size_t off = 0; off < seg[i].bytes && e == hipSuccess; off += kPinChunk, ++c) {
        const size_t len = std::min(kPinChunk, seg[i].bytes - off);
        e = hipMemcpyAsync(w.pin[c & 1], (const char*)seg[i].src + off, len,
                           hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(w.ev[c & 1], st);
        if (e == hipSuccess && prev_dst) {
          e = hipEventSynchronize(w.ev[(c - 1) & 1]);
          if (e == hipSuccess) memcpy(prev_dst, w.pin[(c - 1) & 1], prev_len);
        }
        prev_dst = (char*)seg[i].dst + off;
        prev_len = len;
      }
    }
    if (e == hipSuccess && prev_dst) {
      e = hipEventSynchronize(w.ev[(c - 1) & 1]);
      if (e == hipSuccess) memcpy(prev_dst, w.pin[(c - 1) & 1], prev_len);
    }
  }
  const hipError_t es = hipStreamSynchronize(st);
  if (e == hipSuccess) e = es;
  return e == hipSuccess ? MEXCT_OK : mexct::set_error(MEXCT_EHIP, hipGetErrorString(e));
}

int copy_out(Worker& w, void* dst, const void* src, size_t bytes) {
  const Seg s{dst, src, bytes};
  return copy_out_segs(w, mexct::ctx_stream(w.ctx), &s, 1);
}

#define TRY(x)          \
  do {                  \
    int r_ = (x);       \
    if (r_) return r_;  \
  } while (0)

#define HTRY(x)                                                                     \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) return mexct::set_error(MEXCT_EHIP, hipGetErrorString(e_)); \
  } while (0)

// A resident request's array: device memory of the server's device holding
// at least `bytes` from p on.
bool on_device(const void* p, size_t bytes, int device) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (at.type != hipMemoryTypeDevice || at.device != device) return false;
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return (const char*)p + bytes <= (const char*)base + size;
}

// Input and output bytes of a request (input2 has the input's size).
void io_bytes(const mexct_request& r, size_t* in, size_t* out) {
  const size_t vb = (size_t)r.W * r.W * r.H * 4, sb = (size_t)r.W * r.H * r.nProj * 4;
  *in = r.op == MEXCT_OP_FP ? vb : sb;
  *out = r.op == MEXCT_OP_FP ? sb : vb;
}

// A resident request: inputs and output in device memory, no staging.
int run_one_resident(Worker& w, Req& q) {
  const mexct_request& r = q.r;
  const mexct_opts o{r.storage, r.tv_steps, 0, 0, MEXCT_G3D_BRICKS};
  const size_t sb = (size_t)r.W * r.H * r.nProj * 4;
  const float* ang = q.angles.data();
  switch (r.op) {
    case MEXCT_OP_FP:
      TRY(mexct_fp_dev(w.ctx, r.input, MEXCT_F32, r.W, r.H, r.sad, ang, (int32_t)r.nProj, r.geom,
                       r.output, &o));
      break;
    case MEXCT_OP_BP:
    case MEXCT_OP_FDK: {
      const void* bp_in = r.input;
      if (r.op == MEXCT_OP_FDK) {
        TRY(grow(w.x, sb));
        TRY(mexct_prefilter_dev(w.ctx, r.input, r.W, r.H, r.nProj, r.geom, r.sad, ang, r.window,
                                r.cutoff, MEXCT_NO_PARKER, (float*)w.x.p));
        bp_in = w.x.p;
      }
      TRY(mexct_bp_dev(w.ctx, bp_in, MEXCT_F32, r.W, r.H, r.nProj, r.sad, ang, r.geom, r.output,
                       &o));
      break;
    }
    case MEXCT_OP_OSCTV:
      TRY(mexct_osc_tv_recon_dev(w.ctx, r.input, r.input2, r.W, r.H, r.nProj, ang, r.sad, r.geom,
                                 r.osc_iters, r.osc_tv, r.osc_s_init, r.osc_s_final, r.osc_p,
                                 q.picks.data(), r.output, &o));
      break;
    default: return mexct::set_error(MEXCT_EINVAL, "server: unknown operator");
  }
  HTRY(hipStreamSynchronize(mexct::ctx_stream(w.ctx)));
  return MEXCT_OK;
}

// Runs one request on a worker: host -> device, operator(s) on the worker's
// stream, device -> host.  Returns when the output is in host memory.
int run_one(Worker& w, Req& q) {
  const mexct_request& r = q.r;
  if (r.resident) return run_one_resident(w, q);
  const mexct_opts o{r.storage, r.tv_steps, 0, 0, MEXCT_G3D_BRICKS};
  const size_t vb = (size_t)r.W * r.W * r.H * 4, sb = (size_t)r.W * r.H * r.nProj * 4;
  const float* ang = q.angles.data();
  switch (r.op) {
    case MEXCT_OP_FP:
      TRY(grow(w.a, vb));
      TRY(grow(w.o, sb));
      TRY(copy_in(w, w.a.p, r.input, vb));
      TRY(mexct_fp_dev(w.ctx, w.a.p, MEXCT_F32, r.W, r.H, r.sad, ang, (int32_t)r.nProj, r.geom,
                       (float*)w.o.p, &o));
      return copy_out(w, r.output, w.o.p, sb);
    case MEXCT_OP_BP:
    case MEXCT_OP_FDK: {
      TRY(grow(w.a, sb));
      TRY(grow(w.o, vb));
      TRY(copy_in(w, w.a.p, r.input, sb));
      const void* bp_in = w.a.p;
      if (r.op == MEXCT_OP_FDK) {
        TRY(grow(w.x, sb));
        TRY(mexct_prefilter_dev(w.ctx, (const float*)w.a.p, r.W, r.H, r.nProj, r.geom, r.sad,
                                ang, r.window, r.cutoff, MEXCT_NO_PARKER, (float*)w.x.p));
        bp_in = w.x.p;
      }
      TRY(mexct_bp_dev(w.ctx, bp_in, MEXCT_F32, r.W, r.H, r.nProj, r.sad, ang, r.geom,
                       (float*)w.o.p, &o));
      return copy_out(w, r.output, w.o.p, vb);
    }
    case MEXCT_OP_OSCTV:
      TRY(grow(w.a, sb));
      TRY(grow(w.b, sb));
      TRY(grow(w.o, vb));
      TRY(copy_in(w, w.a.p, r.input, sb));
      TRY(copy_in(w, w.b.p, r.input2, sb));
      TRY(mexct_osc_tv_recon_dev(w.ctx, (const float*)w.a.p, (const float*)w.b.p, r.W, r.H,
                                 r.nProj, ang, r.sad, r.geom, r.osc_iters, r.osc_tv,
                                 r.osc_s_init, r.osc_s_final, r.osc_p, q.picks.data(),
                                 (float*)w.o.p, &o));
      return copy_out(w, r.output, w.o.p, vb);
    default: return mexct::set_error(MEXCT_EINVAL, "server: unknown operator");
  }
}

bool batchable(const mexct_request& r) {
  return r.op == MEXCT_OP_FP || r.op == MEXCT_OP_BP || r.op == MEXCT_OP_FDK;
}

// The batched operator over items [first, first + count) of a group whose
// inputs sit one after another in w.a; outputs go to w.o (FDK filters into
// w.x first).  Enqueued on the context's stream.
int launch_items(Worker& w, const mexct_request& r, const float* ang, int first, int count) {
  const mexct_opts o{r.storage, r.tv_steps, 0, 0, MEXCT_G3D_BRICKS};
  size_t in_b, out_b;
  io_bytes(r, &in_b, &out_b);
  const char* in = (const char*)w.a.p + first * in_b;
  float* out = (float*)((char*)w.o.p + first * out_b);
  if (r.op == MEXCT_OP_FP)
    return mexct_fp_batch_dev(w.ctx, in, MEXCT_F32, count, r.W, r.H, r.sad, ang, (int32_t)r.nProj,
                              r.geom, out, &o);
  const void* bp_in = in;
  if (r.op == MEXCT_OP_FDK) {
    char* x = (char*)w.x.p + first * in_b;
    for (int i = 0; i < count; ++i)
      TRY(mexct_prefilter_dev(w.ctx, (const float*)(in + i * in_b), r.W, r.H, r.nProj, r.geom,
                              r.sad, ang, r.window, r.cutoff, MEXCT_NO_PARKER,
                              (float*)(x + i * in_b)));
    bp_in = x;
  }
  return mexct_bp_batch_dev(w.ctx, bp_in, MEXCT_F32, count, r.W, r.H, r.nProj, r.sad, ang, r.geom,
                            out, &o);
}

// Outputs of items [first, first + count) to the host, on the copy stream
// once the kernels recorded by ev have run.
int copy_out_items(Worker& w, const std::vector<Req*>& g, int first, int count, hipEvent_t ev) {
  size_t in_b, out_b;
  io_bytes(g[0]->r, &in_b, &out_b);
  HTRY(hipStreamWaitEvent(w.cs, ev, 0));
  std::vector<Seg> segs((size_t)count);
  for (int i = 0; i < count; ++i)
    segs[i] = Seg{g[first + i]->r.output, (char*)w.o.p + (first + i) * out_b, out_b};
  return copy_out_segs(w, w.cs, segs.data(), count);
}

// Runs a same-key group of FP/BP/FDK requests as one batch, each item
// computed exactly as alone.
//  - resident requests: one batched launch that reads and writes every item
//    at its own address;
//  - host requests: in two halves, so that copies and kernels overlap:
//    inputs of half 0; inputs of half 1 while half 0 computes; outputs of
//    half 0 while half 1 computes; outputs of half 1.
int run_batch_impl(Worker& w, const std::vector<Req*>& g) {
  const mexct_request& r = g[0]->r;
  const int n = (int)g.size();
  size_t in_b, out_b;
  io_bytes(r, &in_b, &out_b);
  const float* ang = g[0]->angles.data();
  hipStream_t ks = mexct::ctx_stream(w.ctx);
  if (r.op == MEXCT_OP_FDK) TRY(grow(w.x, in_b * n));
  if (r.resident) {  // every item read and written in place
    const mexct_opts o{r.storage, r.tv_steps, 0, 0, MEXCT_G3D_BRICKS};
    std::vector<const void*> ins((size_t)n);
    std::vector<float*> outs((size_t)n);
    for (int i = 0; i < n; ++i) {
      ins[i] = g[i]->r.input;
      outs[i] = g[i]->r.output;
    }
    if (r.op == MEXCT_OP_FP) {
      TRY(mexct::fp_batch_ptrs(w.ctx, ins.data(), MEXCT_F32, n, r.W, r.H, r.sad, ang,
                               (int32_t)r.nProj, r.geom, outs.data(), &o));
    } else {
      if (r.op == MEXCT_OP_FDK)
        for (int i = 0; i < n; ++i) {
          float* x = (float*)((char*)w.x.p + i * in_b);
          TRY(mexct_prefilter_dev(w.ctx, (const float*)ins[i], r.W, r.H, r.nProj, r.geom, r.sad,
                                  ang, r.window, r.cutoff, MEXCT_NO_PARKER, x));
          ins[i] = x;
        }
      TRY(mexct::bp_batch_ptrs(w.ctx, ins.data(), MEXCT_F32, n, r.W, r.H, r.nProj, r.sad, ang,
                               r.geom, outs.data(), &o));
    }
    HTRY(hipStreamSynchronize(ks));
    return MEXCT_OK;
  }
  TRY(grow(w.a, in_b * n));
  TRY(grow(w.o, out_b * n));
  const int h = (n + 1) / 2;
  const int first[2] = {0, h}, count[2] = {h, n - h};
  for (int p = 0; p < 2; ++p) {
    if (count[p] > 0) {
      for (int i = first[p]; i < first[p] + count[p]; ++i)
        HTRY(hipMemcpyAsync((char*)w.a.p + i * in_b, g[i]->r.input, in_b, hipMemcpyHostToDevice,
                            w.cs));
      HTRY(hipEventRecord(w.ev_in, w.cs));
      HTRY(hipStreamWaitEvent(ks, w.ev_in, 0));
      TRY(launch_items(w, r, ang, first[p], count[p]));
      HTRY(hipEventRecord(w.ev_k[p], ks));
    }
    if (p == 1) TRY(copy_out_items(w, g, first[0], count[0], w.ev_k[0]));
  }
  if (count[1] > 0) TRY(copy_out_items(w, g, first[1], count[1], w.ev_k[1]));
  return MEXCT_OK;
}

int run_batch(Worker& w, const std::vector<Req*>& g) {
  const int rc = run_batch_impl(w, g);
  if (rc) {  // nothing of the batch may still run when its requests are retried alone
    (void)hipStreamSynchronize(w.cs);
    (void)hipStreamSynchronize(mexct::ctx_stream(w.ctx));
  }
  return rc;
}

enum class Pick { kWait, kRun, kRelease };

// Chooses an idle worker's next action (server lock held).
Pick choose(mexct_server* sv, Worker& w, std::vector<Req*>& group) {
  if (sv->queue.empty()) return Pick::kWait;
  Req* head = sv->queue.front();
  int64_t idle_reserved = 0;
  int idle = 0;  // workers free to take work now, this one included
  for (auto& o : sv->workers)
    if (!o->running) {
      idle_reserved += o->reserved;
      ++idle;
    }
  const int64_t others = sv->in_use - w.reserved;
  for (auto it = sv->queue.begin(); it != sv->queue.end(); ++it) {
    Req* q = *it;
    const int64_t need = std::max(w.reserved, q->footprint);
    if (others + need > sv->budget) {
      if (q != head) continue;
      // The oldest request does not fit.  Idle workers give their memory
      // back when that alone makes room for it.
      if (w.reserved > 0 && sv->in_use - idle_reserved + head->footprint <= sv->budget)
        return Pick::kRelease;
      if (head->overtaken >= kMaxOvertake) return Pick::kWait;
      continue;
    }
    if (q != head) ++head->overtaken;  // head was blocked to get here
    group.push_back(q);
    it = sv->queue.erase(it);
    // a batch holds every item's buffers at once; back-to-back runs reuse one set
    const bool batch = batchable(q->r);
    int64_t group_fp = q->footprint;
    // share the queued requests of this key evenly among the idle workers,
    // so that a burst keeps every one of them copying and computing (busy
    // workers take theirs when they finish)
    int same = 1;
    for (auto jt = it; jt != sv->queue.end(); ++jt) same += same_key(**jt, *q) ? 1 : 0;
    const int nw = std::max(idle, 1);
    const int cap = std::min(sv->max_group, (same + nw - 1) / nw);
    for (auto jt = it; jt != sv->queue.end() && (int)group.size() < cap;) {
      const int64_t more = batch ? group_fp + (*jt)->footprint : group_fp;
      if (same_key(**jt, *q) && others + std::max(w.reserved, more) <= sv->budget) {
        group.push_back(*jt);
        group_fp = more;
        jt = sv->queue.erase(jt);
      } else {
        ++jt;
      }
    }
    w.reserved = std::max(w.reserved, group_fp);
    sv->in_use = others + w.reserved;
    sv->peak_in_use = std::max(sv->peak_in_use, sv->in_use);
    return Pick::kRun;
  }
  return Pick::kWait;
}

void worker_loop(mexct_server* sv, Worker* w) {
  (void)hipSetDevice(sv->device);
  for (;;) {
    std::vector<Req*> group;
    Pick pick = Pick::kWait;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      for (;;) {
        if (sv->stop) return;
        pick = choose(sv, *w, group);
        if (pick != Pick::kWait) break;
        sv->cv_work.wait(lk);
      }
      if (pick == Pick::kRun) {
        w->running = true;
        const double t = now_s();
        for (Req* q : group) {
          q->status = 2;
          q->t_start = t;
        }
        ++sv->n_groups;
        sv->n_grouped += (int64_t)group.size();
      }
    }
    if (pick == Pick::kRelease) {
      const int rc = release(sv, *w);
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->in_use -= w->reserved;
      w->reserved = 0;
      ++sv->n_releases;
      if (rc) sv->stop = true;  // the context could not be recreated: take no more work
      sv->cv_work.notify_all();
      sv->cv_done.notify_all();
      continue;
    }
    bool batched = false;
    if (group.size() > 1 && batchable(group[0]->r) && run_batch(*w, group) == MEXCT_OK) {
      batched = true;
      std::lock_guard<std::mutex> lk(sv->mu);
      const double t = now_s();
      for (Req* q : group) {
        q->status = 0;
        q->t_done = t;
        ++sv->n_done;
      }
      sv->n_batched += (int64_t)group.size();
      ++sv->n_batches;
      sv->cv_done.notify_all();
    }
    // one by one: OSC-TV groups, single requests, and a batch that failed
    // (so that a failing request does not take the others with it)
    for (Req* q : group) {
      if (batched) break;
      const int rc = run_one(*w, *q);
      const std::string msg = rc ? std::string(mexct_last_error()) : std::string();
      std::lock_guard<std::mutex> lk(sv->mu);
      q->status = rc ? rc : 0;
      q->err = msg;
      q->t_done = now_s();
      ++sv->n_done;
      sv->cv_done.notify_all();
    }
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      w->running = false;
      const int64_t held = held_bytes(*w);
      sv->in_use += held - w->reserved;
      w->reserved = held;
      sv->peak_in_use = std::max(sv->peak_in_use, sv->in_use);
    }
    sv->cv_work.notify_all();
  }
}

}  // namespace

extern "C" {

int mexct_server_create(mexct_server** out, int device, int64_t hbm_budget, int32_t slots,
                        int32_t max_group) {
  if (!out) return mexct::set_error(MEXCT_EINVAL, "server is null");
  if (slots < 1 || slots > 16 || hbm_budget <= 0)
    return mexct::set_error(MEXCT_EINVAL, "server: need 1 <= slots <= 16 and a positive budget");
  std::unique_ptr<mexct_server> sv(new mexct_server());
  sv->device = device;
  sv->budget = hbm_budget;
  sv->max_group = max_group > 0 ? max_group : 1;
  int prev = 0;
  (void)hipGetDevice(&prev);
  for (int i = 0; i < slots; ++i) {
    std::unique_ptr<Worker> w(new Worker());
    int rc = new_ctx(sv.get(), *w);
    if (!rc) rc = worker_streams(sv->device, *w);
    if (rc) {
      free_streams(*w);
      if (w->ctx) mexct_ctx_destroy(w->ctx);
      for (auto& p : sv->workers) {
        free_streams(*p);
        mexct_ctx_destroy(p->ctx);
      }
      (void)hipSetDevice(prev);
      return rc;
    }
    sv->workers.push_back(std::move(w));
  }
  (void)hipSetDevice(prev);
  for (auto& w : sv->workers) w->th = std::thread(worker_loop, sv.get(), w.get());
  *out = sv.release();
  return MEXCT_OK;
}

int mexct_server_submit(mexct_server* sv, const mexct_request* r, uint64_t* ticket) {
  if (!sv || !r || !ticket) return mexct::set_error(MEXCT_EINVAL, "server: null argument");
  if (r->op < MEXCT_OP_FP || r->op > MEXCT_OP_OSCTV)
    return mexct::set_error(MEXCT_EINVAL, "server: unknown operator");
  if (r->geom < 0 || r->geom > 2) return mexct::set_error(MEXCT_EGEOM, "invalid geometry");
  if (r->W < 1 || r->H < 1 || r->nProj < 1 || r->nProj > INT32_MAX || !r->angles || !r->input ||
      !r->output ||
      (r->op == MEXCT_OP_OSCTV && (!r->input2 || !r->first_picks || r->osc_iters < 1)))
    return mexct::set_error(MEXCT_EINVAL, "server: bad request dims or null array");
  if (r->resident) {
    size_t in_b, out_b;
    io_bytes(*r, &in_b, &out_b);
    if (!on_device(r->input, in_b, sv->device) || !on_device(r->output, out_b, sv->device) ||
        (r->op == MEXCT_OP_OSCTV && !on_device(r->input2, in_b, sv->device)))
      return mexct::set_error(MEXCT_EINVAL,
                              "server: a resident request's arrays must be device memory of the "
                              "server's device, large enough for the request");
  }
  std::unique_ptr<Req> q(new Req());
  q->r = *r;
  q->angles.assign(r->angles, r->angles + r->nProj);
  if (r->op == MEXCT_OP_OSCTV) q->picks.assign(r->first_picks, r->first_picks + r->osc_iters);
  q->footprint = footprint_of(*r);
  if (q->footprint > sv->budget)
    return mexct::set_error(MEXCT_ENOMEM, "server: request needs more than the HBM budget");
  q->t_submit = now_s();
  std::lock_guard<std::mutex> lk(sv->mu);
  if (sv->stop) return mexct::set_error(MEXCT_EHIP, "server: stopped");
  q->id = sv->next_id++;
  *ticket = q->id;
  sv->queue.push_back(q.get());
  sv->reqs[q->id] = std::move(q);
  sv->cv_work.notify_all();
  return MEXCT_OK;
}

int64_t mexct_server_footprint(const mexct_request* r) {
  if (!r || r->op < MEXCT_OP_FP || r->op > MEXCT_OP_OSCTV || r->W < 1 || r->H < 1 || r->nProj < 1 ||
      r->storage < MEXCT_F32 || r->storage > MEXCT_BF16)
    return mexct::set_error(MEXCT_EINVAL, "server: bad request for a footprint");
  return footprint_of(*r);
}

int mexct_server_wait(mexct_server* sv, uint64_t ticket, int32_t* status, double* latency_s,
                      double* queue_s) {
  if (!sv) return mexct::set_error(MEXCT_EINVAL, "server is null");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return mexct::set_error(MEXCT_EINVAL, "server: unknown ticket");
  Req* q = it->second.get();
  sv->cv_done.wait(lk, [&] { return q->status <= 0 || (sv->stop && q->status == 1); });
  if (q->status == 1) return mexct::set_error(MEXCT_EHIP, "server: stopped before the request ran");
  if (status) *status = q->status;
  if (latency_s) *latency_s = q->t_done - q->t_submit;
  if (queue_s) *queue_s = q->t_start - q->t_submit;
  if (q->status) mexct::set_error(q->status, q->err.c_str());
  sv->reqs.erase(it);
  return MEXCT_OK;
}

int mexct_server_stats(mexct_server* sv, int64_t* stats, int32_t n) {
  if (!sv || !stats || n < 1) return mexct::set_error(MEXCT_EINVAL, "server: null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  const int64_t v[9] = {sv->n_done,   sv->n_groups,      sv->n_grouped,
                        (int64_t)sv->queue.size(), sv->in_use, sv->peak_in_use,
                        sv->n_releases, sv->n_batched, sv->n_batches};
  for (int i = 0; i < n && i < 9; ++i) stats[i] = v[i];
  return MEXCT_OK;
}

int mexct_server_destroy(mexct_server* sv) {
  if (!sv) return MEXCT_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  sv->cv_done.notify_all();
  for (auto& w : sv->workers)
    if (w->th.joinable()) w->th.join();
  (void)hipSetDevice(sv->device);
  for (auto& w : sv->workers) {
    if (w->ctx) (void)mexct_ctx_sync(w->ctx);
    for (Buf* b : {&w->a, &w->b, &w->o, &w->x})
      if (b->p) (void)hipFree(b->p);
    free_pinned(*w);
    free_streams(*w);
    mexct_ctx_destroy(w->ctx);
  }
  delete sv;
  return MEXCT_OK;
}

}  // extern "C"
