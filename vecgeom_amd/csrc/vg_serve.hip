// Serving loop: a persistent navigation server with continuous batching.
//
// Clients open baskets of tracks (positions and directions); the server keeps
// each basket's tracks -- position, direction, navigation path, boundary flag,
// last step, step count and the running hash of the (volume, boundary)
// sequence -- resident in a device arena sized by an HBM budget, and admits a
// basket only while the arena has a free run of slots for it (admission by
// footprint).  A request asks for one navigation step of every track of one
// basket: the body of the reference's TraceTrack loop
// (FindNextBoundaryAndStep, then p += d * (step + 1e-6);
// test/globalbenchmarks/TraceTrack.cpp:188-220).
//
// A host thread runs the loop.  Whenever a stream is free it takes every
// pending request whose basket is not already in flight (up to a merge
// limit), merges them into one segmented launch -- requests that arrive while
// others run join the next launch -- and records an event; with two streams
// two launches are in flight, so new work overlaps running work.  The thread
// polls the events of running launches (a dedicated host core while work is
// in flight) and sleeps on a condition variable when idle.  A request's
// latency runs from its submission to the loop's observation that its
// launch completed, on the host's steady clock.
//
// vg_server_replay is an in-process load generator (the transport client):
// baskets arrive at given times, wait for admission, take a fixed number of
// steps (each step one request, the next submitted when the previous
// completes) and close, freeing their slots for waiting baskets.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "vg_nav_impl.hpp"

namespace vga {

namespace {

using namespace nav;
using Clock = std::chrono::steady_clock;

constexpr int kServeBlock = 128;
// The world volume's daughter boxes are staged in LDS by the step kernel
// when they take at most this many bytes.
constexpr size_t kServeBoxLds = 8192;
// ... and, with the mixed navigator, their FP32 bound records (144 bytes
// each) when those take at most this many bytes more.
constexpr size_t kServeBoundLds = 9216;
constexpr int kMaxStreams = 2;
constexpr int kMaxSegs = 4096;  // requests merged into one launch at most

// Track arena, structure of arrays over slots.
struct Arena {
  double *x, *y, *z, *dx, *dy, *dz, *step;
  uint64_t *hash;
  int32_t *level, *boundary, *nsteps, *path;  // path[l * cap + slot]
  uint64_t cap;
  int32_t depth;
  int32_t pad;
};

// Bytes per slot: 7 doubles, the hash, 3 int32 and one int32 per depth.
uint64_t SlotBytes(int depth) { return 8 * 8 + 3 * 4 + 4 * (uint64_t)depth; }

struct Seg {
  uint64_t offset, n, tile_begin;
};

__device__ __forceinline__ int FindSeg(const Seg *segs, int nseg, uint64_t tile) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].tile_begin <= tile)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Locates new tracks (slots [offset, offset + n)) from the world.
template <int D, uint32_t M>
__global__ __launch_bounds__(kServeBlock) void ServeLocateKernel(const NavDev g, const Arena a, uint64_t offset,
                                                                 uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * kServeBlock + threadIdx.x;
  if (i >= n) return;
  const uint64_t s = offset + i;
  DPath<D> p;
  p.Clear();
  LocatePoint<D, M>(g, g.world, V3{a.x[s], a.y[s], a.z[s]}, p, true);
  a.level[s] = p.level;
  a.boundary[s] = 0;
  for (int l = 0; l < a.depth; ++l) a.path[(uint64_t)l * a.cap + s] = l < D ? p.Get(l) : -1;
  a.step[s] = 0;
  a.nsteps[s] = 0;
  a.hash[s] = 1469598103934665603ULL;
}

// One step of every track of every segment; a track outside the world
// (level 0) stays where it is with step 0.  Tiles of kServeBlock tracks are
// numbered across segments; each block finds its segment by binary search.
template <int D, uint32_t M, bool MX>
__global__ __launch_bounds__(kServeBlock) void ServeStepKernel(const NavDev g, const Arena a,
                                                               const Seg *__restrict__ segs, int nseg, uint64_t ntiles,
                                                               double pstep, int kind, int stage_lo, int stage_hi,
                                                               int stage_bounds) {
  // The world volume's daughter-slot boxes in LDS (as the trace kernel
  // stages them; none when stage_lo == stage_hi), and in the mixed kernel
  // their FP32 bound records after them.
  extern __shared__ __align__(16) float serve_boxes[];
  NavDev gl = g;
  if (stage_hi > stage_lo) {  // uniform over the grid
    const int nf = 8 * (stage_hi - stage_lo);
    const float *src = g.slot_box + 8 * (size_t)stage_lo;
    for (int j = threadIdx.x; j < nf; j += kServeBlock) serve_boxes[j] = src[j];
    __syncthreads();
    gl.lds_box = serve_boxes;
    gl.lds_lo = stage_lo;
    gl.lds_hi = stage_hi;
    if constexpr (MX) {
      if (stage_bounds) {  // uniform over the grid: the same daughters' FP32 bound records
        constexpr int kRec = (int)(sizeof(B32) / sizeof(uint32_t));
        uint32_t *dst = reinterpret_cast<uint32_t *>(serve_boxes + nf);
        const int nr = kRec * (stage_hi - stage_lo);
        for (int j = threadIdx.x; j < nr; j += kServeBlock) {
          const int slot = stage_lo + j / kRec;
          dst[j] = reinterpret_cast<const uint32_t *>(g.b32 + g.daughters[slot])[j % kRec];
        }
        __syncthreads();
        gl.lds_bound = reinterpret_cast<const B32 *>(dst);
      }
    }
  }
  const uint64_t tile = blockIdx.x;
  if (tile >= ntiles) return;
  const Seg sg = segs[FindSeg(segs, nseg, tile)];
  const uint64_t i = (tile - sg.tile_begin) * kServeBlock + threadIdx.x;
  if (i >= sg.n) return;
  const uint64_t s = sg.offset + i;
  DPath<D> cur, nst;
  cur.level = a.level[s];
  cur.boundary = a.boundary[s];
#pragma unroll
  for (int l = 0; l < D; ++l) cur.v[l] = (l < a.depth && l < cur.level) ? a.path[(uint64_t)l * a.cap + s] : -1;
  if (cur.level <= 0) {
    a.step[s] = 0;
    return;
  }
  V3 p{a.x[s], a.y[s], a.z[s]};
  const V3 d{a.dx[s], a.dy[s], a.dz[s]};
  double step;
  FindNextBoundaryAndStep<D, false, M, MX>(gl, p, d, cur, nst, pstep, step, kind);
  const double sp = step + 1E-6;
  p = V3{p.x + d.x * sp, p.y + d.y * sp, p.z + d.z * sp};
  a.x[s] = p.x;
  a.y[s] = p.y;
  a.z[s] = p.z;
  a.step[s] = step;
  a.level[s] = nst.level;
  a.boundary[s] = nst.boundary;
  for (int l = 0; l < a.depth; ++l) a.path[(uint64_t)l * a.cap + s] = l < D ? nst.Get(l) : -1;
  a.nsteps[s] += 1;
  const uint64_t v = (uint64_t)(uint32_t)(nst.Top() + 1) * 2ULL + (nst.boundary ? 1ULL : 0ULL);
  a.hash[s] = (a.hash[s] ^ v) * 1099511628211ULL;
}

}  // namespace
}  // namespace vga

using namespace vga;

struct vg_server {
  const vg_geometry *g = nullptr;
  NavDev dev{};
  int stage_lo = 0, stage_hi = 0;  // world daughter slots staged in LDS by the step kernel
  int stage_bounds = 0;            // 1: their FP32 bound records too (mixed navigator)
  size_t box_bytes = 0;            // LDS bytes of the step kernel
  Arena a{};
  int navigator = VG_NAV_ABBOX;
  double pstep = 0;
  int max_merge = kMaxSegs;
  std::atomic<int> nstreams{kMaxStreams};
  void *mem = nullptr;
  // arena allocation: free ranges (offset -> length), first fit
  std::map<uint64_t, uint64_t> free_ranges;
  struct Basket {
    uint64_t offset, n;
    bool busy;
  };
  std::map<int64_t, Basket> baskets;
  int64_t next_basket = 1;
  struct Request {
    int64_t basket;
    Clock::time_point submitted;
    bool done;
    double latency;
  };
  std::map<int64_t, Request> requests;
  std::deque<int64_t> pending;
  int64_t next_request = 1;
  // launches
  hipStream_t ustream = nullptr;  // basket uploads
  hipStream_t stream[kMaxStreams] = {nullptr, nullptr};
  hipEvent_t event[kMaxStreams] = {nullptr, nullptr};
  Seg *h_segs[kMaxStreams] = {nullptr, nullptr};  // pinned
  Seg *d_segs[kMaxStreams] = {nullptr, nullptr};
  std::vector<int64_t> inflight[kMaxStreams];
  // statistics
  uint64_t launches = 0, merged = 0, track_steps = 0;
  // threading
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::atomic<bool> stop_flag{false};   // stop, readable without the lock
  std::atomic<int64_t> npending{0};     // pending.size(), readable without the lock
  int error = VG_OK;
  std::string error_msg;
  std::thread loop;
};

namespace {

void Fail(vg_server *s, int rc) {
  if (s->error) return;
  s->error = rc;
  s->error_msg = LastError();
}

double Since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

// Takes pending requests for baskets not in flight into a launch on stream k.
// Called with the lock held.
void StartLaunch(vg_server *s, int k) {
  std::vector<int64_t> take;
  std::deque<int64_t> keep;
  for (int64_t r : s->pending) {
    auto &req = s->requests[r];
    auto &b = s->baskets[req.basket];  // a basket with queued requests cannot be closed
    if (b.busy || (int)take.size() >= s->max_merge) {
      keep.push_back(r);
      continue;
    }
    b.busy = true;
    take.push_back(r);
  }
  s->pending.swap(keep);
  s->npending.store((int64_t)s->pending.size(), std::memory_order_release);
  if (take.empty()) return;
  uint64_t tiles = 0;
  int nseg = 0;
  for (int64_t r : take) {
    const auto &b = s->baskets[s->requests[r].basket];
    if (b.n == 0) continue;
    s->h_segs[k][nseg] = Seg{b.offset, b.n, tiles};
    tiles += (b.n + kServeBlock - 1) / kServeBlock;
    s->track_steps += b.n;
    ++nseg;
  }
  int rc = VG_OK;
  if (nseg > 0) {
    rc = CheckHip(hipMemcpyAsync(s->d_segs[k], s->h_segs[k], nseg * sizeof(Seg), hipMemcpyHostToDevice, s->stream[k]),
                  "serve segment upload");
    if (!rc) {
      VGA_NAV_VARIANTS_MX(s->g, s->navigator == VG_NAV_ABBOX_MIXED, ServeStepKernel, dim3((unsigned)tiles),
                          dim3(kServeBlock), s->box_bytes, s->stream[k], s->dev, s->a, s->d_segs[k], nseg, tiles,
                          s->pstep, s->navigator, s->stage_lo, s->stage_hi, s->stage_bounds);
      rc = CheckHip(hipGetLastError(), "serve step launch");
    }
  }
  if (!rc) rc = CheckHip(hipEventRecord(s->event[k], s->stream[k]), "serve event");
  if (rc) Fail(s, rc);
  // The requests are in flight even on failure: the next pass completes
  // them, and their waiters see the error.
  s->inflight[k] = take;
  s->launches += 1;
  s->merged += take.size();
}

// Completes the requests of stream k's launch.  Called with the lock held.
void Complete(vg_server *s, int k) {
  const auto now = Clock::now();
  for (int64_t r : s->inflight[k]) {
    auto &req = s->requests[r];
    req.done = true;
    req.latency = std::chrono::duration<double>(now - req.submitted).count();
    s->baskets[req.basket].busy = false;
  }
  s->inflight[k].clear();
}

// The loop thread is the only writer of inflight[] and the only user of the
// launch streams and events, so it polls the events without the lock and
// takes the lock only to move requests.
void Loop(vg_server *s) {
  const int rc = CheckHip(hipSetDevice(s->g->device), "serve hipSetDevice");
  std::unique_lock<std::mutex> lk(s->mu);
  if (rc) Fail(s, rc);
  lk.unlock();
  while (true) {
    bool ready[kMaxStreams] = {false, false};
    hipError_t status[kMaxStreams] = {hipSuccess, hipSuccess};
    bool any_ready = false, any_idle = false, running = false;
    const int nstreams = s->nstreams.load(std::memory_order_acquire);
    for (int k = 0; k < kMaxStreams; ++k) {
      if (s->inflight[k].empty()) {
        any_idle |= k < nstreams;
        continue;
      }
      running = true;
      status[k] = hipEventQuery(s->event[k]);
      ready[k] = status[k] != hipErrorNotReady;
      any_ready |= ready[k];
    }
    // While launches run and nothing can change, poll again without the lock.
    if (running && !any_ready && !(any_idle && s->npending.load(std::memory_order_acquire) > 0) &&
        !s->stop_flag.load(std::memory_order_acquire)) {
      std::this_thread::yield();
      continue;
    }
    lk.lock();
    for (int k = 0; k < kMaxStreams; ++k) {
      if (!ready[k]) continue;
      if (status[k] != hipSuccess) Fail(s, CheckHip(status[k], "serve event query"));
      Complete(s, k);
    }
    if (any_ready) s->cv_done.notify_all();
    if (s->error) {
      // Running launches are waited for; queued requests complete too, and
      // their waiters see the error.
      for (int k = 0; k < kMaxStreams; ++k) {
        if (s->inflight[k].empty()) continue;
        (void)hipEventSynchronize(s->event[k]);
        Complete(s, k);
      }
      for (int64_t r : s->pending) s->requests[r].done = true;
      s->pending.clear();
      s->npending.store(0, std::memory_order_release);
      s->cv_done.notify_all();
      s->cv_work.wait(lk, [s] { return s->stop; });
      break;
    }
    for (int k = 0; k < nstreams && !s->pending.empty(); ++k)
      if (s->inflight[k].empty()) StartLaunch(s, k);
    bool any_inflight = false;
    for (int k = 0; k < kMaxStreams; ++k) any_inflight |= !s->inflight[k].empty();
    if (s->stop && !any_inflight && s->pending.empty()) break;
    if (!any_inflight) s->cv_work.wait(lk, [s] { return s->stop || !s->pending.empty(); });
    lk.unlock();
  }
}

bool Alloc(vg_server *s, uint64_t n, uint64_t *offset) {
  for (auto it = s->free_ranges.begin(); it != s->free_ranges.end(); ++it) {
    if (it->second < n) continue;
    *offset = it->first;
    const uint64_t rest = it->second - n, start = it->first + n;
    s->free_ranges.erase(it);
    if (rest) s->free_ranges[start] = rest;
    return true;
  }
  return false;
}

void Release(vg_server *s, uint64_t offset, uint64_t n) {
  if (n == 0) return;
  auto it = s->free_ranges.emplace(offset, n).first;
  auto next = std::next(it);
  if (next != s->free_ranges.end() && it->first + it->second == next->first) {
    it->second += next->second;
    s->free_ranges.erase(next);
  }
  if (it != s->free_ranges.begin()) {
    auto prev = std::prev(it);
    if (prev->first + prev->second == it->first) {
      prev->second += it->second;
      s->free_ranges.erase(it);
    }
  }
}

int ServerError(vg_server *s) { return SetError(s->error, "server failed: " + s->error_msg); }

int Open(vg_server *s, uint64_t n, const double *const src[6], int64_t *basket) {
  uint64_t off = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->error) return ServerError(s);
    if (n && !Alloc(s, n, &off)) return SetError(VG_ERR_CAPACITY, "HBM budget exhausted: basket not admitted");
  }
  int rc = EnsureDevice();
  double *dst[6] = {s->a.x, s->a.y, s->a.z, s->a.dx, s->a.dy, s->a.dz};
  for (int i = 0; i < 6 && !rc && n; ++i)
    rc = CheckHip(hipMemcpyAsync(dst[i] + off, src[i], n * 8, hipMemcpyHostToDevice, s->ustream), "basket upload");
  if (!rc && n) {
    const unsigned blocks = (unsigned)((n + kServeBlock - 1) / kServeBlock);
    VGA_NAV_VARIANTS(s->g, ServeLocateKernel, dim3(blocks), dim3(kServeBlock), 0, s->ustream, s->dev, s->a, off, n);
    rc = CheckHip(hipGetLastError(), "basket locate");
  }
  if (!rc) rc = CheckHip(hipStreamSynchronize(s->ustream), "basket open sync");
  std::lock_guard<std::mutex> lk(s->mu);
  if (rc) {
    Release(s, off, n);
    return rc;
  }
  const int64_t id = s->next_basket++;
  s->baskets[id] = vg_server::Basket{off, n, false};
  *basket = id;
  return VG_OK;
}

int Submit(vg_server *s, int64_t basket, int64_t *request) {
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->error) return ServerError(s);
  if (s->stop) return SetError(VG_ERR_STATE, "server is shutting down");
  if (!s->baskets.count(basket)) return SetError(VG_ERR_INVALID, "unknown basket");
  const int64_t id = s->next_request++;
  s->requests[id] = vg_server::Request{basket, Clock::now(), false, 0.};
  s->pending.push_back(id);
  s->npending.store((int64_t)s->pending.size(), std::memory_order_release);
  *request = id;
  s->cv_work.notify_one();
  return VG_OK;
}

// Returns 1 and the latency if the request completed (and releases it), 0 if
// not, or a negative status.  Called with the lock held.
int PollLocked(vg_server *s, int64_t request, double *latency_s) {
  auto it = s->requests.find(request);
  if (it == s->requests.end()) return SetError(VG_ERR_INVALID, "unknown request");
  if (!it->second.done) return 0;
  if (s->error) return ServerError(s);
  if (latency_s) *latency_s = it->second.latency;
  s->requests.erase(it);
  return 1;
}

int Close(vg_server *s, int64_t basket) {
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->baskets.find(basket);
  if (it == s->baskets.end()) return SetError(VG_ERR_INVALID, "unknown basket");
  if (it->second.busy) return SetError(VG_ERR_STATE, "basket has a step in flight");
  for (int64_t r : s->pending)
    if (s->requests[r].basket == basket) return SetError(VG_ERR_STATE, "basket has queued requests");
  Release(s, it->second.offset, it->second.n);
  s->baskets.erase(it);
  return VG_OK;
}

}  // namespace

extern "C" {

int vg_server_create(const vg_geometry *g, uint64_t hbm_budget_bytes, int32_t navigator, double pstep,
                     vg_server **out) {
  if (!g || !out) return SetError(VG_ERR_INVALID, "null argument");
  *out = nullptr;
  if (navigator != VG_NAV_ABBOX && navigator != VG_NAV_SIMPLE && navigator != VG_NAV_ABBOX_MIXED) return SetError(VG_ERR_INVALID, "unknown navigator");
  if (g->max_depth > VG_NAV_MAX_DEPTH) return SetError(VG_ERR_UNSUPPORTED, "geometry too deep for the navigator");
  if (g->nav_bad_type) return SetError(VG_ERR_UNSUPPORTED, "no device kernels for a shape of the volume tree");
  int rc = EnsureDevice();
  if (rc) return rc;
  const int depth = std::max(1, g->max_depth);
  const uint64_t cap = hbm_budget_bytes / SlotBytes(depth);
  if (cap == 0) return SetError(VG_ERR_CAPACITY, "HBM budget below one track");
  auto *s = new vg_server;
  s->g = g;
  s->navigator = navigator;
  s->pstep = pstep;
  s->dev = NavDev{g->d_placed, g->d_params, g->d_daughters, g->d_aabb, g->d_b32, g->d_b32side, g->d_slot_aabb,
                  g->world,    (int32_t)g->placed.size(), (int32_t)g->daughters.size(), 0};
  if (g->world >= 0) {
    const PlacedRec &w = g->placed[g->world];
    const size_t bytes = 8 * sizeof(float) * (size_t)(w.dend - w.dbegin);
    if (bytes > 0 && bytes <= kServeBoxLds) {
      s->stage_lo = w.dbegin;
      s->stage_hi = w.dend;
      s->box_bytes = bytes;
      const size_t bound_bytes = sizeof(B32) * (size_t)(w.dend - w.dbegin);
      if (navigator == VG_NAV_ABBOX_MIXED && g->d_b32 && bound_bytes <= kServeBoundLds) {
        s->stage_bounds = 1;
        s->box_bytes += bound_bytes;
      }
    }
  }
  rc = CheckHip(hipMalloc(&s->mem, cap * SlotBytes(depth)), "hipMalloc(serve arena)");
  if (!rc) {
    char *m = static_cast<char *>(s->mem);
    auto take = [&](uint64_t bytes) {
      char *p = m;
      m += bytes;
      return p;
    };
    Arena &a = s->a;
    a.cap = cap;
    a.depth = depth;
    double **d64[7] = {&a.x, &a.y, &a.z, &a.dx, &a.dy, &a.dz, &a.step};
    for (double **p : d64) *p = reinterpret_cast<double *>(take(cap * 8));
    a.hash = reinterpret_cast<uint64_t *>(take(cap * 8));
    a.level = reinterpret_cast<int32_t *>(take(cap * 4));
    a.boundary = reinterpret_cast<int32_t *>(take(cap * 4));
    a.nsteps = reinterpret_cast<int32_t *>(take(cap * 4));
    a.path = reinterpret_cast<int32_t *>(take(cap * 4 * (uint64_t)depth));
    s->free_ranges[0] = cap;
  }
  if (!rc) rc = CheckHip(hipStreamCreateWithFlags(&s->ustream, hipStreamNonBlocking), "serve upload stream");
  for (int k = 0; k < kMaxStreams && !rc; ++k) {
    rc = CheckHip(hipStreamCreateWithFlags(&s->stream[k], hipStreamNonBlocking), "serve stream");
    if (!rc) rc = CheckHip(hipEventCreateWithFlags(&s->event[k], hipEventDisableTiming), "serve event");
    if (!rc)
      rc = CheckHip(hipHostMalloc(reinterpret_cast<void **>(&s->h_segs[k]), kMaxSegs * sizeof(Seg)), "pinned segs");
    if (!rc) rc = CheckHip(hipMalloc(&s->d_segs[k], kMaxSegs * sizeof(Seg)), "hipMalloc(segs)");
  }
  if (rc) {
    const std::string msg = LastError();
    vg_server_destroy(s);
    return SetError(rc, msg);
  }
  s->loop = std::thread(Loop, s);
  *out = s;
  return VG_OK;
}

int vg_server_set_batching(vg_server *s, int32_t max_requests_per_launch, int32_t streams) {
  if (!s) return SetError(VG_ERR_INVALID, "null server");
  if (max_requests_per_launch < 1 || max_requests_per_launch > kMaxSegs)
    return SetError(VG_ERR_INVALID, "max_requests_per_launch must be in [1, 4096]");
  if (streams < 1 || streams > kMaxStreams) return SetError(VG_ERR_INVALID, "streams must be 1 or 2");
  std::lock_guard<std::mutex> lk(s->mu);
  bool busy = !s->pending.empty();
  for (int k = 0; k < kMaxStreams; ++k) busy |= !s->inflight[k].empty();
  if (busy) return SetError(VG_ERR_STATE, "requests are queued or in flight");
  s->max_merge = max_requests_per_launch;
  s->nstreams.store(streams, std::memory_order_release);
  return VG_OK;
}

int vg_server_capacity(const vg_server *s, uint64_t *tracks_total, uint64_t *tracks_free) {
  if (!s) return SetError(VG_ERR_INVALID, "null server");
  auto *m = const_cast<vg_server *>(s);
  std::lock_guard<std::mutex> lk(m->mu);
  if (tracks_total) *tracks_total = s->a.cap;
  if (tracks_free) {
    uint64_t f = 0;
    for (auto &r : s->free_ranges) f += r.second;
    *tracks_free = f;
  }
  return VG_OK;
}

int vg_server_open(vg_server *s, uint64_t n, const double *x, const double *y, const double *z, const double *dx,
                   const double *dy, const double *dz, int64_t *basket) {
  if (!s || !basket || (n && (!x || !y || !z || !dx || !dy || !dz))) return SetError(VG_ERR_INVALID, "null argument");
  const double *src[6] = {x, y, z, dx, dy, dz};
  return Open(s, n, src, basket);
}

int vg_server_submit(vg_server *s, int64_t basket, int64_t *request) {
  if (!s || !request) return SetError(VG_ERR_INVALID, "null argument");
  return Submit(s, basket, request);
}

int vg_server_wait(vg_server *s, int64_t request, double *latency_s) {
  if (!s) return SetError(VG_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(request);
  if (it == s->requests.end()) return SetError(VG_ERR_INVALID, "unknown request");
  s->cv_done.wait(lk, [&] { return it->second.done; });
  const int r = PollLocked(s, request, latency_s);
  return r < 0 ? r : VG_OK;
}

int vg_server_poll(vg_server *s, int64_t request, int32_t *done, double *latency_s) {
  if (!s || !done) return SetError(VG_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  const int r = PollLocked(s, request, latency_s);
  if (r < 0) return r;
  *done = r;
  return VG_OK;
}

int vg_server_basket_size(vg_server *s, int64_t basket, uint64_t *n) {
  if (!s || !n) return SetError(VG_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->baskets.find(basket);
  if (it == s->baskets.end()) return SetError(VG_ERR_INVALID, "unknown basket");
  *n = it->second.n;
  return VG_OK;
}

int vg_server_read(vg_server *s, int64_t basket, uint64_t capacity, double *x, double *y, double *z, int32_t *level,
                   int32_t *boundary, int32_t *nsteps, uint64_t *hash, double *last_step) {
  if (!s) return SetError(VG_ERR_INVALID, "null server");
  uint64_t off, n;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->baskets.find(basket);
    if (it == s->baskets.end()) return SetError(VG_ERR_INVALID, "unknown basket");
    if (it->second.busy) return SetError(VG_ERR_STATE, "basket has a step in flight");
    off = it->second.offset;
    n = it->second.n;
  }
  if (n > capacity)
    return SetError(VG_ERR_CAPACITY, "basket holds " + std::to_string(n) + " tracks, the arrays " +
                                         std::to_string(capacity));
  int rc = EnsureDevice();
  struct Out {
    void *dst;
    const void *src;
    uint64_t w;
  } outs[8] = {{x, s->a.x, 8},          {y, s->a.y, 8},           {z, s->a.z, 8},       {level, s->a.level, 4},
               {boundary, s->a.boundary, 4}, {nsteps, s->a.nsteps, 4}, {hash, s->a.hash, 8}, {last_step, s->a.step, 8}};
  for (auto &o : outs) {
    if (rc || !o.dst || !n) continue;
    rc = CheckHip(hipMemcpy(o.dst, static_cast<const char *>(o.src) + off * o.w, n * o.w, hipMemcpyDeviceToHost),
                  "basket read");
  }
  return rc;
}

int vg_server_close(vg_server *s, int64_t basket) {
  if (!s) return SetError(VG_ERR_INVALID, "null server");
  return Close(s, basket);
}

int vg_server_stats(vg_server *s, uint64_t *launches, uint64_t *requests_launched, uint64_t *track_steps) {
  if (!s) return SetError(VG_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (launches) *launches = s->launches;
  if (requests_launched) *requests_launched = s->merged;
  if (track_steps) *track_steps = s->track_steps;
  return VG_OK;
}

int vg_server_replay(vg_server *s, uint64_t nbaskets, const double *arrival_s, const uint64_t *sizes,
                     int32_t steps_per_basket, const double *x, const double *y, const double *z, const double *dx,
                     const double *dy, const double *dz, double *latency_s, double *admit_wait_s, double *wall_s) {
  if (!s || (nbaskets && (!arrival_s || !sizes || !x || !y || !z || !dx || !dy || !dz || !latency_s)))
    return SetError(VG_ERR_INVALID, "null argument");
  if (steps_per_basket < 1) return SetError(VG_ERR_INVALID, "steps_per_basket must be >= 1");
  std::vector<uint64_t> off(nbaskets + 1, 0);
  for (uint64_t i = 0; i < nbaskets; ++i) {
    if (sizes[i] > s->a.cap) return SetError(VG_ERR_CAPACITY, "a basket is larger than the whole arena");
    if (i && arrival_s[i] < arrival_s[i - 1]) return SetError(VG_ERR_INVALID, "arrival times must be sorted");
    off[i + 1] = off[i] + sizes[i];
  }
  // Stage the tracks in pinned host memory, so that basket uploads are
  // asynchronous copies (pageable copies would stall the client thread).
  const uint64_t total = off[nbaskets];
  double *pinned = nullptr;
  if (total) {
    int prc = CheckHip(hipHostMalloc(reinterpret_cast<void **>(&pinned), total * 6 * sizeof(double)),
                       "pinned replay staging");
    if (prc) return prc;
    const double *srcs[6] = {x, y, z, dx, dy, dz};
    for (int c = 0; c < 6; ++c) std::memcpy(pinned + c * total, srcs[c], total * sizeof(double));
  }
  struct Live {
    uint64_t idx;
    int64_t basket, request;
    int32_t done;
  };
  std::vector<Live> live;
  std::deque<uint64_t> admitq;
  uint64_t next = 0, finished = 0;
  const auto t0 = Clock::now();
  int rc = VG_OK;
  while (finished < nbaskets && !rc) {
    bool progress = false;
    const double now = Since(t0);
    while (next < nbaskets && arrival_s[next] <= now) admitq.push_back(next++);
    while (!admitq.empty() && !rc) {
      const uint64_t i = admitq.front();
      const double *src[6];
      for (int c = 0; c < 6; ++c) src[c] = pinned + c * total + off[i];
      Live l{i, 0, 0, 0};
      rc = Open(s, sizes[i], src, &l.basket);
      if (rc == VG_ERR_CAPACITY && !live.empty()) {
        rc = VG_OK;  // waits until closing baskets free slots
        break;
      }
      if (rc) break;
      if (admit_wait_s) admit_wait_s[i] = Since(t0) - arrival_s[i];
      admitq.pop_front();
      live.push_back(l);
      rc = Submit(s, l.basket, &live.back().request);
      progress = true;
    }
    for (size_t j = 0; j < live.size() && !rc;) {
      Live &l = live[j];
      double lat = 0;
      int r;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        r = PollLocked(s, l.request, &lat);
      }
      if (r < 0) {
        rc = r;
        break;
      }
      if (r == 0) {
        ++j;
        continue;
      }
      l.request = 0;
      progress = true;
      latency_s[l.idx * steps_per_basket + l.done] = lat;
      if (++l.done < steps_per_basket) {
        rc = Submit(s, l.basket, &l.request);
        ++j;
        continue;
      }
      rc = Close(s, l.basket);
      ++finished;
      live[j] = live.back();
      live.pop_back();
    }
    if (!progress && !rc) {
      if (live.empty() && admitq.empty() && next < nbaskets) {
        const double wait = arrival_s[next] - Since(t0);
        if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
      } else {
        std::this_thread::yield();
      }
    }
  }
  if (wall_s) *wall_s = Since(t0);
  if (pinned) (void)hipHostFree(pinned);
  if (rc) {
    // Leave no basket behind: wait for what is queued or running, then close.
    for (auto &l : live) {
      double lat;
      if (l.request) (void)vg_server_wait(s, l.request, &lat);
      (void)Close(s, l.basket);
    }
    return rc;
  }
  return VG_OK;
}

int vg_server_destroy(vg_server *s) {
  if (!s) return SetError(VG_ERR_INVALID, "null server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->stop_flag.store(true, std::memory_order_release);
    s->cv_work.notify_all();
  }
  if (s->loop.joinable()) s->loop.join();
  if (s->ustream) (void)hipStreamSynchronize(s->ustream);
  if (s->ustream) (void)hipStreamDestroy(s->ustream);
  for (int k = 0; k < kMaxStreams; ++k) {
    if (s->stream[k]) (void)hipStreamSynchronize(s->stream[k]);
    if (s->stream[k]) (void)hipStreamDestroy(s->stream[k]);
    if (s->event[k]) (void)hipEventDestroy(s->event[k]);
    if (s->h_segs[k]) (void)hipHostFree(s->h_segs[k]);
    if (s->d_segs[k]) (void)hipFree(s->d_segs[k]);
  }
  if (s->mem) (void)hipFree(s->mem);
  const int err = s->error;
  const std::string msg = s->error_msg;
  delete s;
  return err ? SetError(err, "server had failed: " + msg) : VG_OK;
}

}  // extern "C"
