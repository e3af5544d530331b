// Serving loop: continuous batching of triangle-count requests over a
// device-resident CSR pool.  One server thread owns all device work:
//   retire  -- finished launches: deliver counts, free their pool extents;
//   admit   -- waiting requests whose CSR fits the pool (arrival order, with
//              bounded look-ahead past a blocked head) get their extent; a
//              heavy one (heavy_ids DAG ids or more), or one that did not
//              fit the pinned staging area, is uploaded by DMA on a copy
//              stream right away, so it is resident before its launch;
//   launch  -- a free slot takes every admitted light request (up to the
//              batch limits); a heavy request goes alone, on any slot but
//              the last when there are several, so light launches never
//              wait for one and always keep a slot.  On its own stream the slot runs one staging-copy kernel
//              that moves the light requests' CSRs from pinned host memory
//              into the pool, partner build -> zero -> plan -> per-bin count
//              kernels, and the copy-back.
// Launch slots run concurrently on separate streams, so requests arriving
// while others are on the GPU overlap with them on the device.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <atomic>
#include <deque>
#include <thread>

#include "tc_internal.hpp"

using namespace tcr;

namespace {

constexpr uint64_t kAlign = 256;
constexpr int64_t kDefaultBatchIds = (int64_t)32 << 20;  // max_batch_ids = 0
constexpr int64_t kDefaultHeavyIds = (int64_t)4 << 20;   // heavy_ids = 0
constexpr int32_t kDefaultFillGraphs = 256;               // fill_graphs = 0
constexpr int32_t kDefaultFillWaitUs = 2000;              // fill_wait_us = 0
inline uint64_t align_up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// A request's extent in the HBM pool: rowptr | colidx | partner list
// references (each 256-byte aligned).  Its partner rows live in the
// launching slot's scratch.
struct Layout {
  uint64_t ci, pci, total;
  Layout(int32_t n, int64_t nnz) {
    ci = align_up(sizeof(int64_t) * ((uint64_t)n + 1));
    pci = ci + align_up(sizeof(int32_t) * (uint64_t)nnz);
    total = pci + align_up(sizeof(tck::ListRef) * (uint64_t)nnz);
  }
};

// A request's CSR is copied on the submitting thread (into the pinned
// staging area, or a pageable buffer when that is full) by the same pass
// that checks its shape: rowptr non-decreasing with out-degrees below
// kMaxOutDegree, ids in [0, n).  One read of the caller's arrays, in plain
// loops -- an OpenMP region per call, from several client threads at once,
// costs far more than the loops themselves for graphs of a few thousand ids.
// A large request is split over a few helper threads, since one thread
// moves ~9 GB/s and an RMAT-20 request holds ~71 MB of CSR.
constexpr size_t kCopySplit = (size_t)8 << 20;  // bytes per helper thread at least
constexpr int kCopyThreads = 4;
constexpr int64_t kMaxOutDegree = (int64_t)1 << 25;  // d * 64 < 2^31 in the plan's wedge-work field

struct ShapeFlags {
  uint32_t down = 0, wide = 0, id = 0;  // rowptr decreases / degree too large / id outside [0, n)
};

void check_copy(const tc_csr_view& g, int64_t* rp_dst, int32_t* ci_dst, int64_t v0, int64_t v1, int64_t e0,
                int64_t e1, ShapeFlags& f) {
  const int64_t* rp = g.rowptr;
  uint32_t down = 0, wide = 0;
  for (int64_t v = v0; v < v1; ++v) {
    const int64_t d = rp[v + 1] - rp[v];
    down |= d < 0;
    wide |= d >= kMaxOutDegree;
    rp_dst[v + 1] = rp[v + 1];
  }
  const int32_t* ci = g.colidx;
  const uint32_t n = (uint32_t)g.n_rows;
  uint32_t id = 0;
  for (int64_t e = e0; e < e1; ++e) {
    const int32_t c = ci[e];
    id |= (uint32_t)c >= n;
    ci_dst[e] = c;
  }
  f = ShapeFlags{down, wide, id};
}

// Copies g (rowptr[0] == 0 and rowptr[n] == nnz already checked) to rp_dst /
// ci_dst; returns null, or what is wrong with its shape.
const char* stage_checked(const tc_csr_view& g, int64_t* rp_dst, int32_t* ci_dst) {
  const int64_t n = g.n_rows, m = g.nnz;
  rp_dst[0] = 0;
  const size_t bytes = sizeof(int64_t) * (size_t)n + sizeof(int32_t) * (size_t)m;
  const int nt = (int)std::max<size_t>(1, std::min<size_t>(kCopyThreads, bytes / kCopySplit));
  ShapeFlags f[kCopyThreads];
  auto part = [&](int t) {
    check_copy(g, rp_dst, ci_dst, n * t / nt, n * (t + 1) / nt, m * t / nt, m * (t + 1) / nt, f[t]);
  };
  std::vector<std::thread> helpers;
  for (int t = 1; t < nt; ++t) {
    try {
      helpers.emplace_back(part, t);
    } catch (...) {  // no thread to be had: this part runs here
      part(t);
    }
  }
  part(0);
  for (auto& h : helpers) h.join();
  ShapeFlags all;
  for (int t = 0; t < nt; ++t) {
    all.down |= f[t].down;
    all.wide |= f[t].wide;
    all.id |= f[t].id;
  }
  if (all.down) return "rowptr is not monotone";
  if (all.wide) return "DAG out-degree too large";
  if (all.id) return "neighbour id out of range [0, n)";
  return nullptr;
}

// One admission's host-to-device copies: launches of its requests wait on
// the event; the last request to finish returns the event to the pool.
struct Upload {
  hipEvent_t ev = nullptr;
  int refs = 0;
};

struct Req {
  uint64_t tag = 0;
  int32_t n = 0;
  int32_t nv_max = 0;
  int64_t nnz = 0;
  std::vector<int64_t> rp;  // pageable copy (used when the staging area is full)
  std::vector<int32_t> ci;
  char* staged = nullptr;   // or: copy in the pinned staging area
  uint64_t stage_off = 0;
  uint64_t stage_bytes = 0;  // rowptr | pad | colidx, laid out as in the pool
  bool dma = false;          // uploaded at admission (heavy, or not staged)
  uint64_t off = 0, bytes = 0;
  Upload* upload = nullptr;
  double t_submit = 0, t_admit = 0, t_launch = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t start = nullptr;        // timing events around a launch's device work
  hipEvent_t done = nullptr;
  void* plan = nullptr;              // device: items | bin_of | blk | part | counters | outs | descs | nv |
                                     //         partner rows | row offsets
  uint64_t* pinned = nullptr;        // host: outs[G], then descs[k] and stage copies[k] of a launch
  tck::PlanBufs bufs{};
  tck::GraphDesc* d_descs = nullptr;
  int32_t* d_nv = nullptr;
  tck::VRow* d_vrow = nullptr;
  int64_t* d_vrp = nullptr;
  BuildScratch build;
  std::vector<Req*> reqs;
  bool busy = false;
};

}  // namespace

struct tc_server {
  tc_ctx* ctx = nullptr;
  tc_server_config cfg{};
  char* pool = nullptr;
  tch_arena* arena = nullptr;
  char* staging = nullptr;         // pinned host memory
  const char* staging_dev = nullptr;  // the same memory, mapped for the device
  tch_arena* stage_arena = nullptr;
  hipStream_t copy_stream = nullptr;
  std::vector<hipEvent_t> free_events;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> waiting, ready;
  std::deque<tc_request_result> results;
  int64_t outstanding = 0, submitted = 0, completed = 0, rejected = 0, launches = 0, launched_graphs = 0;
  int64_t max_ids = 0;  // DAG ids per launch (cfg.max_batch_ids or its default)
  int64_t heavy_ids = 0;  // cfg.heavy_ids or its default
  int32_t fill_graphs = 0, fill_wait_us = 0;  // cfg's, or their defaults
  // time accounting (ns): submit (callers, atomic); the server thread's
  // phases, read by get_stats (atomic)
  std::atomic<int64_t> t_submit{0}, t_retire{0}, t_admit{0}, t_launch{0}, t_wait_busy{0}, t_wait_idle{0},
      t_device{0};
  bool stop = false;
  std::thread th;

  void run();
  bool retire();
  bool admit();
  bool launch();
  void finish(Req* r, uint64_t count, tc_status st, int batch, double t_done);
  void release();
};

void tc_server::finish(Req* r, uint64_t count, tc_status st, int batch, double t_done) {
  // The request's copies are complete by now (its launch waited on them, or
  // they failed): its staging and the upload event can be reused.
  if (r->staged) tch_arena_free(stage_arena, r->stage_off);
  if (r->upload && --r->upload->refs == 0) {
    free_events.push_back(r->upload->ev);
    delete r->upload;
  }
  tc_request_result res;
  res.tag = r->tag;
  res.count = count;
  res.status = st;
  res.batch_graphs = batch;
  res.t_submit_ms = r->t_submit;
  res.t_admit_ms = r->t_admit;
  res.t_launch_ms = r->t_launch;
  res.t_done_ms = t_done;
  if (r->bytes) tch_arena_free(arena, r->off);
  delete r;
  std::lock_guard<std::mutex> lk(mu);
  results.push_back(res);
  --outstanding;
  ++completed;
  cv_done.notify_all();
}

bool tc_server::retire() {
  bool any = false;
  for (auto& s : slots) {
    if (!s.busy) continue;
    hipError_t e = hipEventQuery(s.done);
    if (e == hipErrorNotReady) continue;
    const double t = now_ms();
    const int k = (int)s.reqs.size();
    float dev_ms = 0;
    if (e == hipSuccess && hipEventElapsedTime(&dev_ms, s.start, s.done) == hipSuccess)
      t_device += (int64_t)((double)dev_ms * 1e6);
    for (int i = 0; i < k; ++i)
      finish(s.reqs[i], e == hipSuccess ? s.pinned[i] : 0, e == hipSuccess ? TC_OK : TC_EDEVICE, k, t);
    if (e != hipSuccess) (void)hipGetLastError();
    s.reqs.clear();
    s.busy = false;
    any = true;
  }
  return any;
}

bool tc_server::admit() {
  std::vector<Req*> take;
  {
    std::lock_guard<std::mutex> lk(mu);
    int passed = 0;
    for (auto it = waiting.begin(); it != waiting.end() && passed <= cfg.lookahead;) {
      Req* r = *it;
      uint64_t off;
      if (tch_arena_alloc(arena, r->bytes, &off) == TC_OK) {
        r->off = off;
        take.push_back(r);
        it = waiting.erase(it);
      } else {
        ++passed;
        ++it;
      }
    }
  }
  if (take.empty()) return false;
  // Light staged requests are copied by their launch's staging kernel.
  // Heavy or pageable ones go out by DMA on the copy stream now, without
  // waiting: one event marks their end, and their launches wait on it on
  // the device.
  hipError_t e = hipSuccess;
  bool dma = false;
  for (Req* r : take) {
    if (r->staged && r->nnz < heavy_ids) continue;
    dma = r->dma = true;
    const Layout L(r->n, r->nnz);
    char* base = pool + r->off;
    if (r->staged) {
      e = hipMemcpyAsync(base, r->staged, r->stage_bytes, hipMemcpyHostToDevice, copy_stream);
    } else {
      e = hipMemcpyAsync(base, r->rp.data(), sizeof(int64_t) * ((size_t)r->n + 1), hipMemcpyHostToDevice,
                         copy_stream);
      if (e == hipSuccess && r->nnz > 0)
        e = hipMemcpyAsync(base + L.ci, r->ci.data(), sizeof(int32_t) * (size_t)r->nnz, hipMemcpyHostToDevice,
                           copy_stream);
    }
    if (e != hipSuccess) break;
  }
  Upload* up = nullptr;
  if (e == hipSuccess && dma) {
    up = new Upload;
    if (!free_events.empty()) {
      up->ev = free_events.back();
      free_events.pop_back();
    } else {
      e = hipEventCreateWithFlags(&up->ev, hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventRecord(up->ev, copy_stream);
    if (e != hipSuccess) {
      if (up->ev) free_events.push_back(up->ev);
      delete up;
      up = nullptr;
    }
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamSynchronize(copy_stream);  // nothing may still read the staging
  }
  const double t = now_ms();
  for (Req* r : take) {
    if (e != hipSuccess) {
      finish(r, 0, TC_EDEVICE, 0, t);
      continue;
    }
    if (r->dma) {
      r->upload = up;
      ++up->refs;
    }
    r->t_admit = t;
    ready.push_back(r);
  }
  return true;
}

bool tc_server::launch() {
  bool any = false;
  for (size_t si = 0; si < slots.size(); ++si) {
    Slot& s = slots[si];
    if (s.busy || ready.empty()) continue;
    // every slot but the last of several takes the oldest heavy request
    // first (alone), so light requests always keep one slot; every slot
    // then takes light requests in order, up to the batch limits
    const bool heavy_slot = si + 1 < slots.size() || slots.size() == 1;
    if (heavy_slot) {
      for (auto it = ready.begin(); it != ready.end(); ++it) {
        if ((*it)->nnz < heavy_ids) continue;
        if (slots.size() > 1 || it == ready.begin()) {  // one slot: in arrival order
          s.reqs.push_back(*it);
          ready.erase(it);
        }
        break;
      }
    }
    if (s.reqs.empty() && fill_graphs > 1) {
      // sharing the GPU with another launch: wait for a full batch, or for
      // the oldest light request to have waited fill_wait_us
      bool others = false;
      for (size_t sj = 0; sj < slots.size(); ++sj) others |= sj != si && slots[sj].busy;
      if (others) {
        int32_t lights = 0;
        double oldest = 0;
        for (const Req* r : ready)
          if (r->nnz < heavy_ids && lights++ == 0) oldest = r->t_submit;
        if (lights < fill_graphs && (lights == 0 || now_ms() - oldest < fill_wait_us * 1e-3)) continue;
      }
    }
    if (s.reqs.empty()) {
      int64_t rows = 0, ids = 0;
      for (auto it = ready.begin(); it != ready.end() && (int)s.reqs.size() < cfg.max_batch_graphs;) {
        Req* r = *it;
        if (r->nnz >= heavy_ids) {
          if (slots.size() == 1) break;  // one slot: a heavy request waits its turn
          ++it;
          continue;
        }
        if (rows + r->nv_max + 1 > cfg.max_batch_rows || ids + r->nnz > max_ids) break;
        rows += r->nv_max + 1;
        ids += r->nnz;
        s.reqs.push_back(r);
        it = ready.erase(it);
      }
    }
    if (s.reqs.empty()) continue;
    const int k = (int)s.reqs.size();
    // descs[k] then the staging copies, in pinned memory and (one copy) on the device
    tck::GraphDesc* hd = (tck::GraphDesc*)(s.pinned + cfg.max_batch_graphs);
    const uint64_t o_copies = align_up(sizeof(tck::GraphDesc) * k);
    tck::StageCopy* hc = (tck::StageCopy*)((char*)hd + o_copies);
    int32_t ncopies = 0;
    int64_t pb = 0, rb = 0, eb = 0, cblocks = 0;
    for (int i = 0; i < k; ++i) {
      Req* r = s.reqs[i];
      const Layout L(r->n, r->nnz);
      hd[i] = tck::GraphDesc{(int64_t)(r->off / sizeof(int64_t)), (int64_t)((r->off + L.ci) / sizeof(int32_t)),
                             (int64_t)((r->off + L.pci) / sizeof(tck::ListRef)), rb, pb, r->n, r->nv_max, 0, r->n, eb};
      pb += (int64_t)r->nv_max + 1;
      rb += (int64_t)r->n + 1;
      eb += r->nnz;
      if (r->staged && !r->dma) {
        hc[ncopies++] = tck::StageCopy{staging_dev + r->stage_off, pool + r->off, (int64_t)r->stage_bytes, cblocks};
        cblocks += tck::stage_blocks((int64_t)r->stage_bytes);
      }
    }
    tck::LaunchArgs a{};
    a.descs = s.d_descs;
    a.ngraphs = k;
    a.P = pb;
    a.rp = (const int64_t*)pool;
    a.ci = (const int32_t*)pool;
    a.pci = (const tck::ListRef*)pool;
    a.vrp = s.d_vrp;
    a.vrow = s.d_vrow;
    a.nv = s.d_nv;
    {
      std::lock_guard<std::recursive_mutex> lk(ctx->mu);
      a.tune = to_kernel_tune(ctx->tune);
    }
    a.bufs = s.bufs;
    for (const Req* r : s.reqs) a.wide_ids |= r->n > tck::kNarrowIds ? 1 : 0;
    const double t = now_ms();
    for (Req* r : s.reqs) r->t_launch = t;
    hipError_t e = hipSuccess;
    for (int i = 0; i < k && e == hipSuccess; ++i) {  // this launch's uploads must have landed
      Upload* up = s.reqs[i]->upload;
      bool seen = false;
      for (int j = 0; j < i && !seen; ++j) seen = s.reqs[j]->upload == up;
      if (up && !seen) e = hipStreamWaitEvent(s.stream, up->ev, 0);
    }
    if (e == hipSuccess) e = hipEventRecord(s.start, s.stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s.d_descs, hd, o_copies + sizeof(tck::StageCopy) * ncopies, hipMemcpyHostToDevice,
                         s.stream);
    if (e == hipSuccess)
      e = tck::launch_stage_copy((const tck::StageCopy*)((char*)s.d_descs + o_copies), ncopies, cblocks,
                                 s.stream);
    // Partner structures of this launch's graphs, then plan and count.
    if (e == hipSuccess)
      e = s.build.ensure(rb, max_ids) == TC_OK ? tck::launch_partners(a, rb, eb, s.build.b, s.stream)
                                              : hipErrorOutOfMemory;
    // A batch of one graph would take the single-graph reduction path.
    if (e == hipSuccess) e = tck::launch_zero(a.bufs.counters, tck::kNumCounters, s.stream);
    if (e == hipSuccess) e = tck::launch_zero(a.bufs.outs, 2 * (int64_t)k, s.stream);
    if (e == hipSuccess) e = tck::launch_plan(a, s.stream);
    if (e == hipSuccess) e = tck::launch_count(a, ctx->cus, ctx->tune.grid, s.stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s.pinned, a.bufs.outs, sizeof(uint64_t) * k, hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      (void)hipStreamSynchronize(s.stream);
      (void)hipStreamSynchronize(copy_stream);  // pageable copies may still read their buffers
      const double tf = now_ms();
      for (Req* r : s.reqs) finish(r, 0, TC_EDEVICE, k, tf);
      s.reqs.clear();
      continue;
    }
    s.busy = true;
    any = true;
    std::lock_guard<std::mutex> lk(mu);
    ++launches;
    launched_graphs += k;
  }
  return any;
}

static int64_t ns_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
}

void tc_server::run() {
  (void)hipSetDevice(ctx->device);
  using clk = std::chrono::steady_clock;
  for (;;) {
    auto t = clk::now();
    bool progressed = retire();
    t_retire += ns_since(t);
    t = clk::now();
    progressed |= admit();
    t_admit += ns_since(t);
    t = clk::now();
    progressed |= launch();
    t_launch += ns_since(t);
    if (progressed) continue;
    t = clk::now();
    std::unique_lock<std::mutex> lk(mu);
    const bool idle = waiting.empty() && ready.empty() &&
                      std::none_of(slots.begin(), slots.end(), [](const Slot& s) { return s.busy; });
    if (stop && idle) break;
    if (idle) {
      cv_work.wait_for(lk, std::chrono::milliseconds(50));
      t_wait_idle += ns_since(t);
    } else {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));  // launches in flight
      t_wait_busy += ns_since(t);
    }
  }
}

void tc_server::release() {
  for (auto& s : slots) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.start) (void)hipEventDestroy(s.start);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    dev_free(s.plan);
    s.build.release();
    if (s.pinned) (void)hipHostFree(s.pinned);
  }
  slots.clear();
  if (copy_stream) (void)hipStreamSynchronize(copy_stream);
  for (hipEvent_t ev : free_events) (void)hipEventDestroy(ev);
  free_events.clear();
  if (copy_stream) (void)hipStreamDestroy(copy_stream);
  dev_free(pool);
  if (arena) tch_arena_destroy(arena);
  if (staging) (void)hipHostFree(staging);
  if (stage_arena) tch_arena_destroy(stage_arena);
  staging = nullptr;
  stage_arena = nullptr;
  for (Req* r : waiting) delete r;
  for (Req* r : ready) {
    if (r->upload && --r->upload->refs == 0) {
      (void)hipEventDestroy(r->upload->ev);
      delete r->upload;
    }
    delete r;
  }
  waiting.clear();
  ready.clear();
}

extern "C" {

double tc_clock_ms(void) { return now_ms(); }

uint64_t tc_request_footprint(int32_t n_rows, int64_t nnz) {
  if (n_rows < 0 || nnz < 0) return 0;
  return Layout(n_rows, nnz).total;
}

tc_status tc_server_create(tc_ctx* ctx, const tc_server_config* cfg, tc_server** out) try {
  if (!ctx || !cfg || !out) return fail(TC_EINVAL, "null argument");
  if (cfg->slots < 1 || cfg->slots > 8) return fail(TC_EINVAL, "slots must be in [1, 8]");
  if (cfg->max_batch_graphs < 1 || cfg->max_batch_rows < 1 || cfg->max_queue < 1 || cfg->lookahead < 0)
    return fail(TC_EINVAL, "batch limits and queue size must be positive");
  if (cfg->max_batch_rows >= ((int64_t)1 << 31)) return fail(TC_ERANGE, "max_batch_rows must be < 2^31");
  if (cfg->hbm_budget < kAlign) return fail(TC_EINVAL, "hbm_budget too small");
  if (cfg->max_batch_ids < 0) return fail(TC_EINVAL, "max_batch_ids must be >= 0");
  if (cfg->heavy_ids < 0) return fail(TC_EINVAL, "heavy_ids must be >= 0");
  if (cfg->fill_graphs < 0 || cfg->fill_wait_us < 0) return fail(TC_EINVAL, "fill_graphs and fill_wait_us must be >= 0");
  std::unique_ptr<tc_server> s(new tc_server);
  s->ctx = ctx;
  s->cfg = *cfg;
  s->max_ids = cfg->max_batch_ids > 0 ? cfg->max_batch_ids : kDefaultBatchIds;
  s->heavy_ids = cfg->heavy_ids > 0 ? cfg->heavy_ids : kDefaultHeavyIds;
  s->fill_graphs = cfg->fill_graphs > 0 ? cfg->fill_graphs : kDefaultFillGraphs;
  s->fill_wait_us = cfg->fill_wait_us > 0 ? cfg->fill_wait_us : kDefaultFillWaitUs;
  DeviceGuard guard(ctx->device);
  tc_status st = dev_alloc((void**)&s->pool, cfg->hbm_budget, "server pool");
  if (st != TC_OK) return st;
  st = tch_arena_create(cfg->hbm_budget, kAlign, &s->arena);
  if (st != TC_OK) {
    s->release();
    return fail(st, tch_last_error());
  }
  if (hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess) {
    s->release();
    return fail(TC_EDEVICE, "copy stream creation failed");
  }
  if (cfg->host_staging >= kAlign) {
    void* dev = nullptr;
    if (hipHostMalloc((void**)&s->staging, cfg->host_staging, hipHostMallocDefault) != hipSuccess ||
        hipHostGetDevicePointer(&dev, s->staging, 0) != hipSuccess ||
        tch_arena_create(cfg->host_staging, kAlign, &s->stage_arena) != TC_OK) {
      (void)hipGetLastError();
      s->release();
      return fail(TC_ENOMEM, "pinned staging allocation failed");
    }
    s->staging_dev = (const char*)dev;
  }
  const int64_t R = cfg->max_batch_rows, G = cfg->max_batch_graphs;
  const int64_t nb = tck::kNumBins * tck::plan_blocks(R) + 1;
  const uint64_t o_items = 0, o_bin = o_items + align_up(8 * R), o_blk = o_bin + align_up(2 * R),
                 o_part = o_blk + align_up(4 * nb), o_cnt = o_part + align_up(8 * tck::scan_partials(nb)),
                 o_outs = o_cnt + align_up(4 * 64), o_desc = o_outs + align_up(8 * G),
                 o_nv = o_desc + align_up(sizeof(tck::GraphDesc) * G) + align_up(sizeof(tck::StageCopy) * G),
                 o_vrow = o_nv + align_up(4 * G),
                 o_vrp = o_vrow + align_up(sizeof(tck::VRow) * R), total = o_vrp + align_up(8 * (R + 1));
  s->slots.resize(cfg->slots);
  for (auto& sl : s->slots) {
    hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&sl.start);
    if (e == hipSuccess) e = hipEventCreate(&sl.done);
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&sl.pinned,
                        8 * G + align_up(sizeof(tck::GraphDesc) * G) + sizeof(tck::StageCopy) * G,
                        hipHostMallocDefault);
    if (e != hipSuccess) {
      s->release();
      return fail(TC_EDEVICE, std::string("server slot: ") + hipGetErrorString(e));
    }
    st = dev_alloc(&sl.plan, total, "slot plan scratch");
    if (st == TC_OK) st = sl.build.ensure(R, s->max_ids);
    if (st != TC_OK) {
      s->release();
      return st;
    }
    char* p = (char*)sl.plan;
    sl.bufs = tck::PlanBufs{(uint16_t*)(p + o_bin), (uint64_t*)(p + o_items), (uint32_t*)(p + o_blk),
                            (int64_t*)(p + o_part), (uint32_t*)(p + o_cnt), (uint64_t*)(p + o_outs)};
    sl.d_descs = (tck::GraphDesc*)(p + o_desc);
    sl.d_nv = (int32_t*)(p + o_nv);
    sl.d_vrow = (tck::VRow*)(p + o_vrow);
    sl.d_vrp = (int64_t*)(p + o_vrp);
  }
  tc_server* raw = s.release();
  raw->th = std::thread([raw]() { raw->run(); });
  *out = raw;
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

static tc_status submit_request(tc_server* s, const tc_csr_view* dag, uint64_t tag, uint32_t check);

tc_status tc_server_submit(tc_server* s, const tc_csr_view* dag, uint64_t tag, uint32_t check) try {
  if (!s || !dag) return fail(TC_EINVAL, "null argument");
  const auto t0 = std::chrono::steady_clock::now();
  const tc_status st = submit_request(s, dag, tag, check);
  s->t_submit += ns_since(t0);
  return st;
} catch (...) {
  return tcr::caught();
}

static tc_status submit_request(tc_server* s, const tc_csr_view* dag, uint64_t tag, uint32_t check) {
  auto reject = [&](tc_status code, const std::string& msg) {
    std::lock_guard<std::mutex> lk(s->mu);
    ++s->rejected;
    return fail(code, msg);
  };
  const tc_csr_view& g = *dag;
  // constant-time shape checks here; the O(n + nnz) ones run with the copy
  if (g.n_rows < 0 || g.nnz < 0) return reject(TC_EINVAL, "negative n_rows or nnz");
  if (!g.rowptr) return reject(TC_EINVAL, "null rowptr");
  if (g.nnz > 0 && !g.colidx) return reject(TC_EINVAL, "null colidx");
  if (g.rowptr[0] != 0) return reject(TC_EINVAL, "rowptr[0] != 0");
  if (g.rowptr[g.n_rows] != g.nnz) return reject(TC_EINVAL, "rowptr[n] != nnz");
  if (check & ~(uint32_t)TC_CHECK_SHAPE) {  // row order / self-loop checks: the host library's pass
    const tc_status st = tch_validate(dag, check | TC_CHECK_SHAPE);
    if (st != TC_OK) return reject(st, tch_last_error());
  }
  const uint64_t fp = tc_request_footprint(g.n_rows, g.nnz);
  tch_arena_stats as;
  tch_arena_stats_get(s->arena, &as);
  if (fp > as.capacity) return reject(TC_ENOMEM, "request footprint exceeds the server's HBM pool");
  const int32_t nvm = tck::partner_rows_max(g.n_rows, g.nnz);
  if ((int64_t)nvm + 1 > s->cfg.max_batch_rows)
    return reject(TC_ERANGE, "request needs more plan rows than max_batch_rows (rows + nnz/1024 + 2)");
  if (g.nnz > s->max_ids) return reject(TC_ERANGE, "request has more DAG ids than max_batch_ids");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(TC_ESTATE, "server is shutting down");
    if (s->outstanding >= s->cfg.max_queue) {
      ++s->rejected;
      return fail(TC_EBUSY, "server queue is full");
    }
    ++s->outstanding;  // reserve the queue place before the copy
  }
  std::unique_ptr<Req> r(new Req);
  r->tag = tag;
  r->n = g.n_rows;
  r->nv_max = nvm;
  r->nnz = g.nnz;
  r->bytes = fp;
  uint64_t so = 0;
  int64_t* rp_dst;
  int32_t* ci_dst;
  if (s->stage_arena && tch_arena_alloc(s->stage_arena, Layout(g.n_rows, g.nnz).pci, &so) == TC_OK) {
    r->stage_off = so;
    r->staged = s->staging + so;
    r->stage_bytes = Layout(g.n_rows, g.nnz).pci;
    rp_dst = (int64_t*)r->staged;
    ci_dst = (int32_t*)(r->staged + align_up(sizeof(int64_t) * ((uint64_t)g.n_rows + 1)));
  } else {
    r->rp.resize((size_t)g.n_rows + 1);
    r->ci.resize((size_t)g.nnz);
    rp_dst = r->rp.data();
    ci_dst = r->ci.data();
  }
  if (const char* bad = stage_checked(g, rp_dst, ci_dst)) {
    if (r->staged) tch_arena_free(s->stage_arena, r->stage_off);
    std::lock_guard<std::mutex> lk(s->mu);
    --s->outstanding;
    ++s->rejected;
    s->cv_done.notify_all();
    return fail(TC_EINVAL, bad);
  }
  r->t_submit = now_ms();
  std::lock_guard<std::mutex> lk(s->mu);
  s->waiting.push_back(r.release());
  ++s->submitted;
  s->cv_work.notify_one();
  return TC_OK;
}

tc_status tc_server_submit_at(tc_server* s, const tc_csr_view* dag, uint64_t tag, uint32_t check,
                              double t_ms) try {
  for (;;) {
    const double left = t_ms - now_ms();
    if (left <= 0) break;
    if (left > 2.0) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((left - 1.5) * 1000)));
    else std::this_thread::yield();
  }
  return tc_server_submit(s, dag, tag, check);
} catch (...) {
  return tcr::caught();
}

tc_status tc_server_poll(tc_server* s, tc_request_result* out, int32_t max, int32_t* n_out, int32_t timeout_ms) try {
  if (!s || !n_out || (max > 0 && !out)) return fail(TC_EINVAL, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->results.empty() && timeout_ms > 0)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !s->results.empty(); });
  int32_t k = 0;
  while (k < max && !s->results.empty()) {
    out[k++] = s->results.front();
    s->results.pop_front();
  }
  *n_out = k;
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

tc_status tc_server_drain(tc_server* s, int32_t timeout_ms) try {
  if (!s) return fail(TC_EINVAL, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_work.notify_one();
  if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(std::max(timeout_ms, 0)),
                           [&] { return s->outstanding == 0; }))
    return fail(TC_ETIMEDOUT, "requests still unfinished after the drain timeout");
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

tc_status tc_server_get_stats(tc_server* s, tc_server_stats* o) try {
  if (!s || !o) return fail(TC_EINVAL, "null argument");
  tch_arena_stats as, ss{};
  tch_arena_stats_get(s->arena, &as);
  if (s->stage_arena) tch_arena_stats_get(s->stage_arena, &ss);
  std::lock_guard<std::mutex> lk(s->mu);
  o->submitted = s->submitted;
  o->completed = s->completed;
  o->rejected = s->rejected;
  o->launches = s->launches;
  o->queued = (int64_t)s->waiting.size();
  o->in_flight = s->outstanding - (int64_t)s->waiting.size();
  o->mean_batch_graphs = s->launches ? (double)s->launched_graphs / s->launches : 0.0;
  o->pool_capacity = as.capacity;
  o->pool_used = as.used;
  o->pool_peak = as.peak_used;
  o->submit_ms = s->t_submit.load() / 1e6;
  o->retire_ms = s->t_retire.load() / 1e6;
  o->admit_ms = s->t_admit.load() / 1e6;
  o->launch_ms = s->t_launch.load() / 1e6;
  o->wait_busy_ms = s->t_wait_busy.load() / 1e6;
  o->wait_idle_ms = s->t_wait_idle.load() / 1e6;
  o->device_ms = s->t_device.load() / 1e6;
  o->staging_used = ss.used;
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

tc_status tc_server_destroy(tc_server* s) try {
  if (!s) return TC_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  if (s->th.joinable()) s->th.join();
  DeviceGuard guard(s->ctx->device);
  s->release();
  delete s;
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

}  // extern "C"
