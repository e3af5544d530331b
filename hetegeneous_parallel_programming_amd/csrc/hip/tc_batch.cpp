// Batches of graphs: all CSRs resident in two HBM pools, counted by one
// hipGraph-captured sequence (zero counters -> bin pivots of every graph ->
// one persistent kernel pulling (graph, degree-bin) tiles from a device
// queue -> copy the per-graph counts back).  The baseline path launches the
// plain kernel once per graph, as the reference's multi-graph driver does
// per subgraph (multigpu.cu:190-214: launch, copy back, per graph).
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "tc_internal.hpp"

using namespace tcr;

struct tc_batch {
  tc_ctx* ctx = nullptr;
  int32_t n = 0;
  int64_t* rp = nullptr;
  int32_t* ci = nullptr;
  tck::GraphDesc* d_descs = nullptr;
  std::vector<tck::GraphDesc> descs;
  std::vector<int32_t> rows;
  int64_t P = 0;        // plan indices (partner-row slots over all graphs)
  int64_t RP = 0;       // rowptr pool slots
  int64_t cio = 0;      // colidx pool slots
  Partners partners;
  int64_t nv = 0;       // partner rows in use over all graphs
  double build_ms = 0;
  PlanScratch scratch;
  uint64_t* pinned = nullptr;  // n counts, then counters
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  tc_tune captured{};
  int64_t bytes = 0;

  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    dev_free(rp);
    dev_free(ci);
    dev_free(d_descs);
    partners.release();
    scratch.release();
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
  }
};

static tc_status capture(tc_ctx* ctx, tc_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  tck::LaunchArgs a{};
  a.descs = b->d_descs;
  a.ngraphs = b->n;
  a.P = b->P;
  a.rp = b->rp;
  a.ci = b->ci;
  b->partners.into(a);
  a.tune = to_kernel_tune(ctx->tune);
  a.bufs = b->scratch.bufs();
  for (int32_t r : b->rows) a.wide_ids |= r > tck::kNarrowIds ? 1 : 0;
  hipStream_t s = ctx->stream;
  TC_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  // Zero fills are kernels, not memset nodes.
  hipError_t e = tck::launch_zero(a.bufs.counters, tck::kNumCounters, s);
  if (e == hipSuccess) e = tck::launch_zero(a.bufs.outs, 2 * (int64_t)b->n, s);
  if (e == hipSuccess) e = tck::launch_plan(a, s);
  // The bin groups (tck::kBinOrder): group 0 on the capture stream, the
  // others on forked branches, whose workgroups fill the CUs the other
  // groups' tails free.
  static_assert(tck::kBinGroups == tc_ctx::kSides + 1, "one side stream per bin group past the first");
  const tck::BinGroups& G = tck::kBinOrder;
  if (e == hipSuccess) e = hipEventRecord(ctx->fork, s);
  for (int i = 1; i < G.n && e == hipSuccess; ++i) e = hipStreamWaitEvent(ctx->side[i - 1], ctx->fork, 0);
  for (int i = 0; i < G.n && e == hipSuccess; ++i)
    e = tck::launch_count_order(a, ctx->cus, ctx->tune.grid, G.bin[i], G.len[i], nullptr,
                                i == 0 ? s : ctx->side[i - 1]);
  for (int i = 1; i < G.n && e == hipSuccess; ++i) {
    e = hipEventRecord(ctx->join[i - 1], ctx->side[i - 1]);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, ctx->join[i - 1], 0);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->pinned, a.bufs.outs, sizeof(uint64_t) * (size_t)b->n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->pinned + b->n, a.bufs.counters, sizeof(uint32_t) * tck::kNumCounters,
                       hipMemcpyDeviceToHost, s);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(s, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(TC_EDEVICE, std::string("batch capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
  }
  b->graph = g;
  TC_HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->captured = ctx->tune;
  return TC_OK;
}

extern "C" {

tc_status tc_batch_create(tc_ctx* ctx, const tc_csr_view* dags, int32_t n_graphs, uint32_t check,
                          tc_batch** out) try {
  if (!ctx || !out || n_graphs < 0 || (n_graphs > 0 && !dags)) return fail(TC_EINVAL, "bad arguments");
  for (int32_t i = 0; i < n_graphs; ++i) {
    const tc_csr_view& d = dags[i];
    tc_status st = tch_validate(&d, check | TC_CHECK_SHAPE);
    if (st != TC_OK) return fail(st, "graph " + std::to_string(i) + ": " + tch_last_error());
    for (int64_t v = 0; v < d.n_rows; ++v)
      if ((d.rowptr[v + 1] - d.rowptr[v]) * 64 >= ((int64_t)1 << 31))
        return fail(TC_ERANGE, "graph " + std::to_string(i) + ": DAG out-degree too large");
  }
  int64_t tot_slots = 0;
  for (int32_t i = 0; i < n_graphs; ++i) tot_slots += (int64_t)tck::partner_rows_max(dags[i].n_rows, dags[i].nnz) + 1;
  if (tot_slots >= ((int64_t)1 << 31)) return fail(TC_ERANGE, "batch has 2^31 or more pivots");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  std::unique_ptr<tc_batch> b(new tc_batch);
  b->ctx = ctx;
  b->n = n_graphs;
  b->descs.resize(n_graphs);
  b->rows.resize(n_graphs);
  int64_t rpo = 0, cio = 0, pro = 0;
  for (int32_t i = 0; i < n_graphs; ++i) {
    const int32_t nvm = tck::partner_rows_max(dags[i].n_rows, dags[i].nnz);
    b->descs[i] = tck::GraphDesc{rpo, cio, cio, rpo, pro, dags[i].n_rows, nvm, 0, dags[i].n_rows, cio};
    b->rows[i] = dags[i].n_rows;
    rpo += dags[i].n_rows + 1;
    cio += dags[i].nnz;
    pro += (int64_t)nvm + 1;
  }
  b->P = pro;
  b->RP = rpo;
  b->cio = cio;
  tc_status st = dev_alloc((void**)&b->rp, sizeof(int64_t) * (size_t)std::max<int64_t>(rpo, 1), "batch rowptr");
  if (st == TC_OK) st = dev_alloc((void**)&b->ci, sizeof(int32_t) * (size_t)std::max<int64_t>(cio, 1), "batch colidx");
  if (st == TC_OK)
    st = dev_alloc((void**)&b->d_descs, sizeof(tck::GraphDesc) * (size_t)std::max(n_graphs, 1), "batch descs");
  if (st == TC_OK) st = b->partners.alloc(cio, pro, n_graphs);
  if (st == TC_OK) st = b->scratch.ensure(b->P, n_graphs);
  if (st != TC_OK) {
    b->release();
    return st;
  }
  hipError_t e = hipHostMalloc((void**)&b->pinned, sizeof(uint64_t) * ((size_t)n_graphs + 16), hipHostMallocDefault);
  if (e != hipSuccess) {
    b->pinned = nullptr;
    b->release();
    return fail(TC_ENOMEM, std::string("pinned counts: ") + hipGetErrorString(e));
  }
  // Pack through a pinned staging buffer, copies in flight on the ctx stream.
  const size_t kStage = (size_t)64 << 20;
  char* stage[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  for (int k = 0; k < 2 && e == hipSuccess; ++k) {
    e = hipHostMalloc((void**)&stage[k], kStage, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreate(&done[k]);
  }
  int cur = 0;
  size_t used = 0;
  struct Pending { void* dst; size_t off, bytes; };
  std::vector<Pending> pend;
  auto flush = [&]() -> hipError_t {
    hipError_t r = hipSuccess;
    for (auto& p : pend) {
      r = hipMemcpyAsync(p.dst, stage[cur] + p.off, p.bytes, hipMemcpyHostToDevice, ctx->stream);
      if (r != hipSuccess) return r;
    }
    pend.clear();
    r = hipEventRecord(done[cur], ctx->stream);
    if (r != hipSuccess) return r;
    cur ^= 1;
    used = 0;
    return hipEventSynchronize(done[cur]);  // the buffer about to be refilled is free again
  };
  auto put = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
    const char* s = (const char*)src;
    while (bytes > 0) {
      if (used == kStage) {
        hipError_t r = flush();
        if (r != hipSuccess) return r;
      }
      size_t chunk = std::min(bytes, kStage - used);
      std::memcpy(stage[cur] + used, s, chunk);
      pend.push_back({dst, used, chunk});
      used += chunk;
      s += chunk;
      dst = (char*)dst + chunk;
      bytes -= chunk;
    }
    return hipSuccess;
  };
  for (int32_t i = 0; i < n_graphs && e == hipSuccess; ++i) {
    e = put(b->rp + b->descs[i].rp_off, dags[i].rowptr, sizeof(int64_t) * ((size_t)dags[i].n_rows + 1));
    if (e == hipSuccess && dags[i].nnz > 0)
      e = put(b->ci + b->descs[i].ci_off, dags[i].colidx, sizeof(int32_t) * (size_t)dags[i].nnz);
  }
  if (e == hipSuccess && !pend.empty()) e = flush();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && n_graphs > 0)
    e = hipMemcpy(b->d_descs, b->descs.data(), sizeof(tck::GraphDesc) * (size_t)n_graphs, hipMemcpyHostToDevice);
  for (int k = 0; k < 2; ++k) {
    if (stage[k]) (void)hipHostFree(stage[k]);
    if (done[k]) (void)hipEventDestroy(done[k]);
  }
  if (e != hipSuccess) {
    b->release();
    return fail(TC_EDEVICE, std::string("batch upload: ") + hipGetErrorString(e));
  }
  // Partner structures of all graphs.  Pairs never cross graphs, so the
  // build goes by segments of whole graphs, each within the radix build's
  // limits (row slots, ids; at most kSegIds ids, which also bounds the
  // scratch), over descriptor copies rebased to the segment's first graph.
  if (n_graphs > 0) {
    struct Seg {
      int32_t g0, g1;
      int64_t rows, ids;
    };
    // TC_BATCH_SEGMENT_IDS lowers the ids per segment (tests build small
    // batches in many segments with it)
    int64_t kSegIds = (int64_t)1 << 27;
    if (const char* env = std::getenv("TC_BATCH_SEGMENT_IDS")) {
      const long long v = std::atoll(env);
      if (v > 0 && v < kSegIds) kSegIds = v;
    }
    std::vector<Seg> segs;
    {
      Seg cur{0, 0, 0, 0};
      for (int32_t i = 0; i < n_graphs; ++i) {
        const tck::GraphDesc& d = b->descs[i];
        const int64_t r = (int64_t)d.n_rows + 1, m = (i + 1 < n_graphs ? b->descs[i + 1].ci_off : cio) - d.ci_off;
        tck::RadixGeom x;
        if (i > cur.g0 && (cur.ids + m > kSegIds || !tck::radix_geom(cur.rows + r, cur.ids + m, x))) {
          cur.g1 = i;
          segs.push_back(cur);
          cur = Seg{i, i, 0, 0};
        }
        cur.rows += r;
        cur.ids += m;
      }
      cur.g1 = n_graphs;
      segs.push_back(cur);
    }
    std::vector<tck::GraphDesc> rebased(b->descs);
    int64_t max_rows = 0, max_ids = 0;
    for (const Seg& sg : segs) {
      const tck::GraphDesc& f = b->descs[sg.g0];
      for (int32_t i = sg.g0; i < sg.g1; ++i) {
        rebased[i].row_base -= f.row_base;
        rebased[i].pr_base -= f.pr_base;
        rebased[i].eb -= f.eb;
      }
      max_rows = std::max(max_rows, sg.rows);
      max_ids = std::max(max_ids, sg.ids);
    }
    tck::GraphDesc* d_seg = nullptr;
    st = dev_alloc((void**)&d_seg, sizeof(tck::GraphDesc) * (size_t)n_graphs, "batch build descs");
    if (st != TC_OK) {
      b->release();
      return st;
    }
    BuildScratch bs;
    e = hipMemcpy(d_seg, rebased.data(), sizeof(tck::GraphDesc) * (size_t)n_graphs, hipMemcpyHostToDevice);
    if (e == hipSuccess) st = bs.ensure(max_rows, max_ids);  // once: the segments reuse it in stream order
    if (e == hipSuccess && st == TC_OK) e = hipEventRecord(ctx->ev[0], ctx->stream);
    for (const Seg& sg : segs) {
      if (e != hipSuccess || st != TC_OK) break;
      const tck::GraphDesc& f = b->descs[sg.g0];
      const int64_t pr1 = sg.g1 < n_graphs ? b->descs[sg.g1].pr_base : b->P;
      tck::LaunchArgs a{};
      a.descs = d_seg + sg.g0;
      a.ngraphs = sg.g1 - sg.g0;
      a.rp = b->rp;
      a.ci = b->ci;
      a.P = pr1 - f.pr_base;
      b->partners.into(a);
      a.vrp += f.pr_base;
      a.vrow += f.pr_base;
      a.nv += sg.g0;
      a.gbase = sg.g0;
      st = launch_build(a, sg.rows, sg.ids, bs, ctx->stream);
    }
    if (st == TC_OK && e == hipSuccess) e = hipEventRecord(ctx->ev[1], ctx->stream);
    std::vector<int32_t> nv((size_t)n_graphs);
    if (st == TC_OK && e == hipSuccess)
      e = hipMemcpyAsync(nv.data(), b->partners.nv, sizeof(int32_t) * nv.size(), hipMemcpyDeviceToHost,
                         ctx->stream);
    if (st == TC_OK && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (st != TC_OK || e != hipSuccess) (void)hipStreamSynchronize(ctx->stream);  // nothing may still use the scratch
    bs.release();
    dev_free(d_seg);
    if (st == TC_OK && e != hipSuccess) st = fail(TC_EDEVICE, std::string("partner build: ") + hipGetErrorString(e));
    if (st != TC_OK) {
      b->release();
      return st;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    b->build_ms = ms;
    for (int32_t v : nv) b->nv += v;
  }
  b->bytes = sizeof(int64_t) * rpo + sizeof(int32_t) * cio + (int64_t)sizeof(tck::GraphDesc) * n_graphs +
             b->partners.bytes + PlanScratch::bytes_for(b->P) + 8 * (int64_t)n_graphs;
  *out = b.release();
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

tc_status tc_batch_count(tc_ctx* ctx, tc_batch* b, uint64_t* counts, tc_stats* stats) try {
  if (!ctx || !b || (!counts && b->n > 0)) return fail(TC_EINVAL, "null argument");
  if (b->ctx != ctx) return fail(TC_EINVAL, "batch belongs to another context");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  const double t0 = now_ms();
  if (!b->exec || std::memcmp(&b->captured, &ctx->tune, sizeof(tc_tune)) != 0) {
    tc_status st = capture(ctx, b);
    if (st != TC_OK) return st;
  }
  TC_HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream));
  TC_HIP_TRY(hipGraphLaunch(b->exec, ctx->stream));
  TC_HIP_TRY(hipEventRecord(ctx->ev[2], ctx->stream));
  TC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (b->n > 0) std::memcpy(counts, b->pinned, sizeof(uint64_t) * (size_t)b->n);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->total_ms = now_ms() - t0;
    float ms = 0;
    TC_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
    stats->count_ms = ms;
    stats->pivots = b->nv;
    const uint32_t* c = (const uint32_t*)(b->pinned + b->n);
    for (int k = 0; k < TC_NUM_BINS; ++k) stats->bin_pivots[k] = c[tck::kCntBinCount + k];
    stats->hash_fallbacks = c[tck::kCntFallback];
  }
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

tc_status tc_batch_count_each(tc_ctx* ctx, tc_batch* b, uint64_t* counts, double* ms_each,
                              tc_stats* stats) try {
  if (!ctx || !b || (!counts && b->n > 0)) return fail(TC_EINVAL, "null argument");
  if (b->ctx != ctx) return fail(TC_EINVAL, "batch belongs to another context");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  const double t0 = now_ms();
  hipStream_t s = ctx->stream;
  uint64_t* d_out = b->scratch.outs;
  TC_HIP_TRY(hipEventRecord(ctx->ev[1], s));
  for (int32_t i = 0; i < b->n; ++i) {
    const double ti = now_ms();
    const tck::GraphDesc& d = b->descs[i];
    const int64_t P = b->rows[i];
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(ctx->plain_grid, (P + tck::kWaves - 1) / tck::kWaves));
    TC_HIP_TRY(hipMemsetAsync(d_out + i, 0, sizeof(uint64_t), s));
    TC_HIP_TRY(tck::launch_plain(b->rp + d.rp_off, b->ci + d.ci_off, 0, b->rows[i], nullptr, d_out + i, grid, s));
    TC_HIP_TRY(hipMemcpyAsync(b->pinned + i, d_out + i, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    TC_HIP_TRY(hipStreamSynchronize(s));
    counts[i] = b->pinned[i];
    if (ms_each) ms_each[i] = now_ms() - ti;
  }
  TC_HIP_TRY(hipEventRecord(ctx->ev[2], s));
  TC_HIP_TRY(hipEventSynchronize(ctx->ev[2]));
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->total_ms = now_ms() - t0;
    float ms = 0;
    TC_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
    stats->count_ms = ms;
    stats->pivots = (int64_t)b->RP - b->n;
  }
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

tc_status tc_batch_free(tc_ctx* ctx, tc_batch* b) try {
  if (!b) return TC_OK;
  if (ctx && b->ctx != ctx) return fail(TC_EINVAL, "batch belongs to another context");
  std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
  DeviceGuard guard(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  b->release();
  delete b;
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

int64_t tc_batch_bytes(const tc_batch* b) { return b ? b->bytes : 0; }
int32_t tc_batch_graphs(const tc_batch* b) { return b ? b->n : 0; }

tc_status tc_batch_get_info(const tc_batch* b, tc_graph_info* out) try {
  if (!b || !out) return fail(TC_EINVAL, "null argument");
  out->graphs = b->n;
  out->rows = b->RP - b->n;
  int64_t nnz = 0;
  for (int32_t i = 0; i < b->n; ++i)
    nnz += (i + 1 < b->n ? b->descs[i + 1].ci_off : b->cio) - b->descs[i].ci_off;
  out->nnz = nnz;
  out->partner_rows = b->nv;
  out->build_ms = b->build_ms;
  out->bytes = b->bytes;
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

}  // extern "C"
