// Device runtime of the triangle-counting library: contexts, graphs in HBM
// and the single-graph count entry points of tc.h.
//
// Reference counterparts: TCSolver (successful_push_pull/multigpu_nvshmem_pushpull.cu:29-85:
// device check, full-residency grid, device graph, zeroed count, launch,
// copy back) and GraphGPU::init (successful_push_pull/graph_gpu.h:84-137).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>

#include "tc_internal.hpp"

namespace tcr {

static thread_local std::string g_err;
// ctx->pinned layout (uint64 words): [0] count, [8..) plan counters,
// [kPinnedFallback] the count's hash-fallback counter.
constexpr int kPinnedFallback = 24;

tc_status fail(tc_status st, const std::string& msg) {
  g_err = msg;
  return st;
}

tc_status caught() noexcept {
  try {
    try {
      throw;
    } catch (const std::bad_alloc&) {
      return fail(TC_ENOMEM, "host allocation failed");
    } catch (const std::length_error&) {
      return fail(TC_ENOMEM, "host allocation too large");
    } catch (const std::exception& e) {
      return fail(TC_EINVAL, std::string("unexpected error: ") + e.what());
    }
  } catch (...) {
    // the error text itself could not be stored
  }
  return TC_ENOMEM;
}

tc_status dev_alloc(void** p, size_t bytes, const char* what) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(e == hipErrorOutOfMemory ? TC_ENOMEM : TC_EDEVICE,
                std::string("hipMalloc(") + what + ", " + std::to_string(bytes) + " B): " +
                    hipGetErrorString(e));
  }
  return TC_OK;
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

tc_status PlanScratch::ensure(int64_t pivots, int64_t nouts) {
  if (!counters) {
    tc_status st = dev_alloc((void**)&counters, sizeof(uint32_t) * 64, "plan counters");
    if (st != TC_OK) return st;
  }
  if (pivots > cap) {
    dev_free(bin_of);
    dev_free(items);
    dev_free(blk);
    dev_free(part);
    bin_of = nullptr;
    items = nullptr;
    blk = nullptr;
    part = nullptr;
    cap = 0;
    int64_t c = std::max<int64_t>(pivots, 1024);
    const int64_t nb = tck::kNumBins * tck::plan_blocks(c) + 1;
    tc_status st = dev_alloc((void**)&bin_of, sizeof(uint16_t) * (size_t)c, "pivot bins");
    if (st == TC_OK) st = dev_alloc((void**)&items, sizeof(uint64_t) * (size_t)c, "pivot items");
    if (st == TC_OK) st = dev_alloc((void**)&blk, sizeof(uint32_t) * (size_t)nb, "plan blocks");
    if (st == TC_OK) st = dev_alloc((void**)&part, sizeof(int64_t) * (size_t)tck::scan_partials(nb), "plan scan");
    if (st != TC_OK) return st;
    cap = c;
  }
  if (nouts > outs_cap) {
    dev_free(outs);
    outs = nullptr;
    outs_cap = 0;
    int64_t c = std::max<int64_t>(nouts, 16);
    tc_status st = dev_alloc((void**)&outs, sizeof(uint64_t) * (size_t)c, "counts");
    if (st != TC_OK) return st;
    outs_cap = c;
  }
  return TC_OK;
}

void PlanScratch::release() {
  dev_free(bin_of);
  dev_free(items);
  dev_free(blk);
  dev_free(part);
  dev_free(counters);
  dev_free(outs);
  *this = PlanScratch();
}

int64_t PlanScratch::bytes_for(int64_t pivots) {
  const int64_t nb = tck::kNumBins * tck::plan_blocks(pivots) + 1;
  return 10 * pivots + 4 * nb + 8 * tck::scan_partials(nb);
}

tc_status Partners::alloc(int64_t ci_slots, int64_t row_slots, int32_t ngraphs) {
  tc_status st = dev_alloc((void**)&pci, sizeof(tck::ListRef) * (size_t)std::max<int64_t>(ci_slots, 1), "partner lists");
  if (st == TC_OK) st = dev_alloc((void**)&vrp, sizeof(int64_t) * (size_t)(row_slots + 1), "partner rows");
  if (st == TC_OK) st = dev_alloc((void**)&vrow, sizeof(tck::VRow) * (size_t)std::max<int64_t>(row_slots, 1), "partner owners");
  if (st == TC_OK) st = dev_alloc((void**)&nv, sizeof(int32_t) * (size_t)std::max(ngraphs, 1), "partner row counts");
  if (st != TC_OK) {
    release();
    return st;
  }
  bytes = bytes_for(ci_slots, row_slots, ngraphs);
  return TC_OK;
}

void Partners::release() {
  dev_free(pci);
  dev_free(vrp);
  dev_free(vrow);
  dev_free(nv);
  *this = Partners();
}

int64_t Partners::bytes_for(int64_t ci_slots, int64_t row_slots, int32_t ngraphs) {
  return (int64_t)sizeof(tck::ListRef) * ci_slots + (8 + (int64_t)sizeof(tck::VRow)) * row_slots + 8 +
         4 * (int64_t)ngraphs;
}

// Piece sizes of a build scratch, in allocation order (256-byte pieces).
static int64_t up256(int64_t b) { return (b + 255) & ~(int64_t)255; }
// Values the scans of a build may cover: the row slots, and with the radix
// part the level-1 and level-2 count arrays.
static int64_t scan_values(int64_t n, int64_t ids) {
  int64_t v = n;
  if (ids > 0) {
    v = std::max<int64_t>(v, (int64_t)tck::kCoarse * tck::radix_split_blocks(n, ids) + 1);
    v = std::max<int64_t>(v, (int64_t)tck::kFine * tck::radix_tiles2(ids) + 1);
  }
  return v;
}
static int64_t build_bytes(int64_t n, int64_t ids) {
  // per row: c1, kc (int32), c2 (int32 per XCD stripe), cur (int64 per
  // stripe), off, voff (int64), scan partials; with the radix part also per
  // row gid, erow, lref, deg16, per id key, pkey (int32), ref, pref (list
  // references), and the
  // merge-path, count and tile arrays
  int64_t b = up256(4 * n) * 2 + up256(4 * tck::kXcds * n) + up256(8 * tck::kXcds * n) + up256(8 * n) * 2 +
              up256(8 * tck::scan_partials(scan_values(n, ids)));
  if (ids > 0) {
    const int64_t nmp = tck::radix_split_blocks(n, ids), nt2 = tck::radix_tiles2(ids);
    b += up256(4 * n) + up256(8 * (n + 1)) + up256(8 * n) + up256(2 * n) + (up256(4 * ids) + up256(8 * ids)) * 2 +
         up256(8 * (nmp + 1)) + up256(4 * (tck::kCoarse * nmp + 1)) + up256(4 * (tck::kFine * nt2 + 1)) +
         up256(8 * (tck::kCoarse + 1)) + up256(8 * (tck::kMaxFineBuckets + 1)) * 2;
  }
  return b;
}

int64_t BuildScratch::bytes_for(int64_t rp_slots, int64_t ids) {
  return build_bytes(std::max<int64_t>(rp_slots, 1), ids);
}

tc_status BuildScratch::ensure(int64_t rp_slots, int64_t ids) {
  const int64_t n = std::max<int64_t>(rp_slots, 1);
  const int64_t need = build_bytes(n, ids);
  if (need > cap) {
    dev_free(mem);
    mem = nullptr;
    cap = 0;
    tc_status st = dev_alloc((void**)&mem, (size_t)need, "partner build scratch");
    if (st != TC_OK) return st;
    cap = need;
  }
  char* p = mem;
  b = tck::BuildBufs{};
  b.c1 = (int32_t*)p;  p += up256(4 * n);
  b.c2 = (int32_t*)p;  p += up256(4 * tck::kXcds * n);
  b.cur = (unsigned long long*)p; p += up256(8 * tck::kXcds * n);
  b.kc = (int32_t*)p;  p += up256(4 * n);
  b.off = (int64_t*)p; p += up256(8 * n);
  b.voff = (int64_t*)p; p += up256(8 * n);
  b.part = (int64_t*)p; p += up256(8 * tck::scan_partials(scan_values(n, ids)));
  b.rows = n;
  if (ids > 0) {
    const int64_t nmp = tck::radix_split_blocks(n, ids), nt2 = tck::radix_tiles2(ids);
    b.gid = (int32_t*)p; p += up256(4 * n);
    b.erow = (int64_t*)p; p += up256(8 * (n + 1));
    b.lref = (tck::ListRef*)p; p += up256(8 * n);
    b.deg16 = (uint16_t*)p; p += up256(2 * n);
    b.key = (int32_t*)p; p += up256(4 * ids);
    b.ref = (tck::ListRef*)p; p += up256(8 * ids);
    b.pkey = (int32_t*)p; p += up256(4 * ids);
    b.pref = (tck::ListRef*)p; p += up256(8 * ids);
    b.bq = (int64_t*)p; p += up256(8 * (nmp + 1));
    b.cnt1 = (uint32_t*)p; p += up256(4 * (tck::kCoarse * nmp + 1));
    b.cnt2 = (uint32_t*)p; p += up256(4 * (tck::kFine * nt2 + 1));
    b.tstart = (int64_t*)p; p += up256(8 * (tck::kCoarse + 1));
    b.chstart = (int64_t*)p; p += up256(8 * (tck::kMaxFineBuckets + 1));
    b.fst = (int64_t*)p;
    b.ids = ids;
  }
  return TC_OK;
}

tc_status BuildScratch::ensure_edges(int64_t ids) {
  const int64_t need = 2 * up256(4 * std::max<int64_t>(ids, 1));
  if (need > ecap) {
    dev_free(emem);
    emem = nullptr;
    ecap = 0;
    tc_status st = dev_alloc((void**)&emem, (size_t)need, "partner build id positions");
    if (st != TC_OK) return st;
    ecap = need;
  }
  b.e1 = (uint32_t*)emem;
  b.e2 = (uint32_t*)(emem + need / 2);
  return TC_OK;
}

void BuildScratch::release() {
  dev_free(mem);
  dev_free(emem);
  *this = BuildScratch();
}

tc_status launch_build(const tck::LaunchArgs& a, int64_t RP, int64_t M, BuildScratch& bs, hipStream_t s,
                       const LocalBufs* loc, int32_t* pown) {
  // The local rebuild also writes each entry's edge position and partner
  // vertex: the radix path carries the positions along (one graph, its
  // edges' rows in loc->esrc, filled before); the per-edge atomic path
  // writes them directly where the radix geometry does not fit.
  tc_status st = bs.ensure(RP, M);
  if (st == TC_OK && loc) st = bs.ensure_edges(M);
  if (st != TC_OK) return st;
  bs.b.pedge = loc ? loc->pedge : nullptr;
  bs.b.pvid = loc ? loc->pvid : nullptr;
  bs.b.pown = pown;  // one graph: each partner row's owner
  bs.b.esrc = loc ? loc->esrc : nullptr;
  TC_HIP_TRY(tck::launch_partners(a, RP, M, bs.b, s));
  return TC_OK;
}

void LocalBufs::release() {
  dev_free(pedge);
  dev_free(pvid);
  dev_free(ecount);
  dev_free(vcount);
  dev_free(pe);
  dev_free(bounds);
  dev_free(esrc);
  *this = LocalBufs();
}

tc_tune default_tune() {
  tc_tune t;
  t.t_g8 = 256;
  t.t_g16 = 1024;
  t.t_g32 = 4096;
  t.t_wave = 16384;
  t.rounds = 4;
  t.grid = 0;
  t.pivot_set = 0;
  t.hash_kicks = 128;
  return t;
}

tc_status check_tune(const tc_tune& t) {
  if (t.t_g8 < 0 || t.t_g16 < t.t_g8 || t.t_g32 < t.t_g16 || t.t_wave < t.t_g32)
    return fail(TC_EINVAL, "tune thresholds must satisfy 0 <= t_g8 <= t_g16 <= t_g32 <= t_wave");
  if (t.rounds < 1 || t.rounds > 64) return fail(TC_EINVAL, "tune rounds must be in [1, 64]");
  if (t.grid < 0 || t.grid > (1 << 20)) return fail(TC_EINVAL, "tune grid must be in [0, 2^20]");
  if (t.pivot_set < 0 || t.pivot_set > 2) return fail(TC_EINVAL, "tune pivot_set must be 0, 1 or 2");
  if (t.hash_kicks < 1 || t.hash_kicks > 1024) return fail(TC_EINVAL, "tune hash_kicks must be in [1, 1024]");
  return TC_OK;
}

tck::Tune to_kernel_tune(const tc_tune& t) {
  return tck::Tune{t.t_g8, t.t_g16, t.t_g32, t.t_wave, t.rounds, t.pivot_set, t.hash_kicks};
}

// Host-side checks shared by every upload: shape (and more when asked), and
// the per-batch int32 element total (64 lists of at most max_out ids).
static tc_status check_dag(const tc_csr_view* dag, uint32_t check, int32_t* max_out) {
  if (!dag) return fail(TC_EINVAL, "null csr view");
  tc_status st = tch_validate(dag, check | TC_CHECK_SHAPE);
  if (st != TC_OK) return fail(st, tch_last_error());
  int64_t mo = 0;
  for (int64_t v = 0; v < dag->n_rows; ++v) mo = std::max<int64_t>(mo, dag->rowptr[v + 1] - dag->rowptr[v]);
  if (mo * 64 >= ((int64_t)1 << 31))
    return fail(TC_ERANGE, "DAG out-degree too large (max_out * 64 must fit in int32)");
  *max_out = (int32_t)mo;
  return TC_OK;
}

static tc_status upload_csr(tc_ctx* ctx, const tc_csr_view* dag, int32_t row_begin, int32_t row_end,
                            const uint8_t* owned, int32_t max_out, tc_graph** out) {
  std::unique_ptr<tc_graph> g(new tc_graph);
  g->ctx = ctx;
  g->n = dag->n_rows;
  g->nnz = dag->nnz;
  g->row_begin = row_begin;
  g->row_end = row_end;
  g->edge_begin = dag->rowptr[row_begin];
  g->edge_end = dag->rowptr[row_end];
  g->max_out = max_out;
  g->nv_max = tck::partner_rows_max(g->n, g->nnz);
  auto cleanup = [&]() {
    dev_free(g->rp);
    dev_free(g->ci);
    dev_free(g->d_desc);
    dev_free(g->own);
    dev_free(g->pown);
    g->partners.release();
  };
  tc_status st = dev_alloc((void**)&g->rp, sizeof(int64_t) * ((size_t)g->n + 1), "rowptr");
  if (st == TC_OK && owned)
    st = dev_alloc((void**)&g->own, std::max<size_t>((size_t)g->n, 1), "owned-row mask");
  if (st == TC_OK) st = dev_alloc((void**)&g->ci, sizeof(int32_t) * (size_t)std::max<int64_t>(g->nnz, 1), "colidx");
  if (st == TC_OK) st = dev_alloc((void**)&g->d_desc, sizeof(tck::GraphDesc), "graph desc");
  if (st == TC_OK) st = g->partners.alloc(g->nnz, (int64_t)g->nv_max + 1, 1);
  if (st == TC_OK)
    st = dev_alloc((void**)&g->pown, sizeof(int32_t) * ((size_t)g->nv_max + 1), "partner row owners");
  if (st != TC_OK) {
    cleanup();
    return st;
  }
  const tck::GraphDesc d{0, 0, 0, 0, 0, g->n, g->nv_max, row_begin, row_end, 0};
  hipError_t e = hipMemcpy(g->rp, dag->rowptr, sizeof(int64_t) * ((size_t)g->n + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess && g->nnz > 0)
    e = hipMemcpy(g->ci, dag->colidx, sizeof(int32_t) * (size_t)g->nnz, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->d_desc, &d, sizeof(d), hipMemcpyHostToDevice);
  if (e == hipSuccess && owned && g->n > 0) e = hipMemcpy(g->own, owned, (size_t)g->n, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    cleanup();
    return fail(TC_EDEVICE, std::string("graph upload: ") + hipGetErrorString(e));
  }
  // Partner structure, built on the device once per upload.
  tck::LaunchArgs a{};
  a.descs = g->d_desc;
  a.ngraphs = 1;
  a.P = (int64_t)g->nv_max + 1;
  a.rp = g->rp;
  a.ci = g->ci;
  a.own = g->own;
  g->partners.into(a);
  BuildScratch bs;
  hipStream_t s = ctx->stream;
  // the scratch is allocated before the timed span: build_ms is the device's build
  st = bs.ensure((int64_t)g->n + 1, g->nnz);
  if (st != TC_OK) {
    cleanup();
    return st;
  }
  e = hipEventRecord(ctx->ev[0], s);
  if (e == hipSuccess) {
    st = launch_build(a, (int64_t)g->n + 1, g->nnz, bs, s, nullptr, g->pown);
    if (st == TC_OK) e = hipEventRecord(ctx->ev[1], s);
    if (st == TC_OK && e == hipSuccess)
      e = hipMemcpyAsync(ctx->pinned, g->partners.nv, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (st == TC_OK && e == hipSuccess) e = hipStreamSynchronize(s);
  }
  bs.release();
  if (st == TC_OK && e != hipSuccess) st = fail(TC_EDEVICE, std::string("partner build: ") + hipGetErrorString(e));
  if (st != TC_OK) {
    cleanup();
    return st;
  }
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
  g->build_ms = ms;
  g->nv = *(const int32_t*)ctx->pinned;
  g->bytes = sizeof(int64_t) * ((int64_t)g->n + 1) + sizeof(int32_t) * g->nnz + sizeof(tck::GraphDesc) +
             g->partners.bytes + (owned ? (int64_t)g->n : 0) + sizeof(int32_t) * ((int64_t)g->nv_max + 1);
  *out = g.release();
  return TC_OK;
}

// plan_ms = ev[0] -> ev[plan_end]; count_ms = ev[1] -> ev[2].
// fallbacks: the hash builds that fell back to the tree (-1: not read).
static tc_status finish_stats(tc_ctx* ctx, tc_stats* st, double t0, bool planned, int64_t pivots,
                              int plan_end = 1, int64_t fallbacks = 0) {
  if (!st) return TC_OK;
  std::memset(st, 0, sizeof(*st));
  st->hash_fallbacks = fallbacks;
  st->total_ms = now_ms() - t0;
  float ms = 0;
  if (planned) {
    TC_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[plan_end]));
    st->plan_ms = ms;
    const uint32_t* c = (const uint32_t*)(ctx->pinned + 8);
    for (int b = 0; b < TC_NUM_BINS; ++b) st->bin_pivots[b] = c[tck::kCntBinCount + b];
  }
  TC_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
  st->count_ms = ms;
  st->pivots = pivots;
  return TC_OK;
}

}  // namespace tcr

using namespace tcr;

extern "C" {

const char* tc_last_error(const tc_ctx*) { return tcr::g_err.c_str(); }

tc_status tc_ctx_create(int32_t device, tc_ctx** out) try {
  if (!out) return fail(TC_EINVAL, "null out");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return fail(TC_EDEVICE, std::string("no HIP device available: ") + hipGetErrorString(e));
  if (device < 0 || device >= ndev)
    return fail(TC_EDEVICE, "device " + std::to_string(device) + " not present (" + std::to_string(ndev) + " visible)");
  std::unique_ptr<tc_ctx> c(new tc_ctx);
  c->device = device;
  DeviceGuard guard(device);
  hipDeviceProp_t prop;
  TC_HIP_TRY(hipGetDeviceProperties(&prop, device));
  c->cus = prop.multiProcessorCount;
  const std::string marketing = prop.name[0] ? std::string(prop.name) : std::string("AMD GPU");
  c->name = marketing + " (" + prop.gcnArchName + ", " + std::to_string(c->cus) + " CUs)";
  TC_HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int i = 0; i < tc_ctx::kSides; ++i) {
    TC_HIP_TRY(hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking));
    TC_HIP_TRY(hipEventCreateWithFlags(&c->join[i], hipEventDisableTiming));
  }
  TC_HIP_TRY(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
  for (auto& ev : c->ev) TC_HIP_TRY(hipEventCreate(&ev));
  TC_HIP_TRY(hipHostMalloc((void**)&c->pinned, 4096, hipHostMallocDefault));
  {
    tc_status st = dev_alloc((void**)&c->plain_out, 256, "plain count word");
    if (st != TC_OK) return st;
    TC_HIP_TRY(hipMemset(c->plain_out, 0, 256));
  }
  // device address of the allocation-free count's static scratch (this device's copy)
  TC_HIP_TRY(tck::noalloc_scratch_address(&c->noalloc));
  c->count_grid = tck::count_blocks_per_cu() * c->cus;
  c->plain_grid = tck::plain_blocks_per_cu() * c->cus;
  c->tune = default_tune();
  *out = c.release();
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

void tc_ctx_destroy(tc_ctx* ctx) {
  if (!ctx) return;
  {
    DeviceGuard guard(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->scratch.release();
    dev_free(ctx->plain_out);
    dev_free(ctx->once_dev);
    for (auto& ev : ctx->ev)
      if (ev) (void)hipEventDestroy(ev);
    for (hipStream_t x : ctx->side)
      if (x) (void)hipStreamSynchronize(x);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    for (hipEvent_t x : ctx->join)
      if (x) (void)hipEventDestroy(x);
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    for (hipStream_t x : ctx->side)
      if (x) (void)hipStreamDestroy(x);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
}

tc_status tc_get_tune(tc_ctx* ctx, tc_tune* out) try {
  if (!ctx || !out) return fail(TC_EINVAL, "null argument");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  *out = ctx->tune;
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

tc_status tc_set_tune(tc_ctx* ctx, const tc_tune* t) try {
  if (!ctx || !t) return fail(TC_EINVAL, "null argument");
  tc_status st = check_tune(*t);
  if (st != TC_OK) return st;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  ctx->tune = *t;
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

tc_status tc_device_name(tc_ctx* ctx, char* buf, int32_t len) try {
  if (!ctx || !buf || len <= 0) return fail(TC_EINVAL, "null argument");
  std::snprintf(buf, (size_t)len, "%s", ctx->name.c_str());
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

tc_status tc_graph_upload(tc_ctx* ctx, const tc_csr_view* dag, uint32_t check, tc_graph** out) try {
  if (!ctx || !out) return fail(TC_EINVAL, "null argument");
  int32_t max_out = 0;
  tc_status st = check_dag(dag, check, &max_out);
  if (st != TC_OK) return st;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  return upload_csr(ctx, dag, 0, dag->n_rows, nullptr, max_out, out);
} catch (...) {
  return tcr::caught();
}

tc_status tc_graph_upload_rows(tc_ctx* ctx, const tc_csr_view* dag, const uint8_t* owned, uint32_t check,
                               tc_graph** out) try {
  if (!ctx || !out) return fail(TC_EINVAL, "null argument");
  int32_t max_out = 0;
  tc_status st = check_dag(dag, check, &max_out);
  if (st != TC_OK) return st;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  return upload_csr(ctx, dag, 0, dag->n_rows, owned, max_out, out);
} catch (...) {
  return tcr::caught();
}

tc_status tc_partition_upload(tc_ctx* ctx, const tch_partition_view* part, uint32_t check,
                              tc_graph** out) try {
  if (!ctx || !part || !out) return fail(TC_EINVAL, "null argument");
  tch_csr* full = nullptr;
  tc_status st = tch_partition_assemble(part, &full);
  if (st != TC_OK) return fail(st, tch_last_error());
  tc_csr_view v;
  tch_csr_view(full, &v);
  int32_t max_out = 0;
  st = check_dag(&v, check, &max_out);
  if (st == TC_OK) {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DeviceGuard guard(ctx->device);
    st = upload_csr(ctx, &v, part->row_begin, part->row_end, nullptr, max_out, out);
  }
  tch_csr_free(full);
  return st;
} catch (...) {
  return tcr::caught();
}

tc_status tc_graph_free(tc_ctx* ctx, tc_graph* g) try {
  if (!g) return TC_OK;
  if (ctx && g->ctx != ctx) return fail(TC_EINVAL, "graph belongs to another context");
  tc_ctx* c = g->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DeviceGuard guard(c->device);
  (void)hipStreamSynchronize(c->stream);
  dev_free(g->rp);
  dev_free(g->ci);
  dev_free(g->d_desc);
  dev_free(g->own);
  dev_free(g->pown);
  g->partners.release();
  g->local.release();
  g->plan.release();
  delete g;
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

int64_t tc_graph_bytes(const tc_graph* g) { return g ? g->bytes + g->local.bytes : 0; }

tc_status tc_graph_get_info(const tc_graph* g, tc_graph_info* out) try {
  if (!g || !out) return fail(TC_EINVAL, "null argument");
  out->graphs = 1;
  out->rows = g->n;
  out->nnz = g->nnz;
  out->partner_rows = g->nv;
  out->build_ms = g->build_ms;
  out->bytes = g->bytes + g->local.bytes;
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

// Fills the LaunchArgs of a count of g (descs, pools, partners, tune, plan
// scratch).
static tck::LaunchArgs count_args(tc_ctx* ctx, tc_graph* g) {
  tck::LaunchArgs a{};
  a.descs = g->d_desc;
  a.ngraphs = 1;
  a.P = (int64_t)g->nv_max + 1;
  a.rp = g->rp;
  a.ci = g->ci;
  g->partners.into(a);
  a.tune = to_kernel_tune(ctx->tune);
  a.bufs = ctx->scratch.bufs();
  a.own = g->own;
  a.wide_ids = g->n > tck::kNarrowIds ? 1 : 0;
  return a;
}

}  // extern "C"

namespace tcr {

// Index of the first group with work (it runs on the library stream).
static int first_group(const bool* any, int n) {
  for (int gi = 0; gi < n; ++gi)
    if (any[gi]) return gi;
  return 0;
}

// The binned count over plan buffers `bufs`.  kept: the plan in bufs is
// already built for this graph and tune (its rows per bin in bins[], tiles
// per bin in tiles[]): re-arm the tile queues and count.  Otherwise plan
// first (one host round trip for the rows per bin, which size the launches)
// and fill bins[] and tiles[].
static tc_status binned_count(tc_ctx* ctx, tc_graph* g, const tck::PlanBufs& bufs, bool kept, int64_t* bins,
                              int64_t* tiles, uint64_t* total, tc_stats* stats) {
  const double t0 = now_ms();
  tck::LaunchArgs a = count_args(ctx, g);
  a.bufs = bufs;
  hipStream_t s = ctx->stream;
  TC_HIP_TRY(hipEventRecord(ctx->ev[0], s));
  if (kept) {
    TC_HIP_TRY(tck::launch_rearm(a.bufs.counters, a.bufs.outs, 2, s));
    TC_HIP_TRY(hipEventRecord(ctx->ev[3], s));
  } else {
    TC_HIP_TRY(tck::launch_zero(a.bufs.counters, tck::kNumCounters, s));
    TC_HIP_TRY(tck::launch_zero(a.bufs.outs, 2, s));
    TC_HIP_TRY(tck::launch_plan(a, s));
    TC_HIP_TRY(hipEventRecord(ctx->ev[3], s));
    // The plan's rows per bin size each bin's launch: empty bins are skipped
    // and small ones get no more workgroups than tiles (a small graph would
    // otherwise pay for six full-residency grids).
    TC_HIP_TRY(hipMemcpyAsync(ctx->pinned + 8, a.bufs.counters, sizeof(uint32_t) * tck::kNumCounters,
                              hipMemcpyDeviceToHost, s));
    TC_HIP_TRY(hipStreamSynchronize(s));
    const uint32_t* c = (const uint32_t*)(ctx->pinned + 8);
    for (int b = 0; b < tck::kNumBins; ++b) {
      const int64_t piv = tck::tile_pivots(b, ctx->tune.rounds > 0 ? ctx->tune.rounds : 1);
      bins[b] = c[tck::kCntBinCount + b];
      tiles[b] = (bins[b] + piv - 1) / piv;
    }
  }
  TC_HIP_TRY(hipEventRecord(ctx->ev[1], s));
  // The bin groups (tck::kBinOrder) run on separate streams (default
  // priority): each group's persistent workgroups hold their CUs until its
  // bins drain; the dispatcher interleaves the groups' workgroups from the
  // start and the others take over the CUs a group's tail frees.  Groups
  // with no tiles are skipped; one non-empty group runs on the library
  // stream alone.
  const tck::BinGroups& G = tck::kBinOrder;
  bool any[tck::kBinGroups] = {};
  for (int gi = 0; gi < G.n; ++gi)
    for (int j = 0; j < G.len[gi]; ++j) any[gi] = any[gi] || tiles[G.bin[gi][j]] > 0;
  const int g0 = first_group(any, G.n);
  int side = 0;
  for (int gi = 0; gi < G.n; ++gi) {
    if (!any[gi]) continue;
    hipStream_t gs = s;
    if (gi != g0) {
      gs = ctx->side[side++];
      TC_HIP_TRY(hipStreamWaitEvent(gs, ctx->ev[1], 0));  // timing starts before any group
    }
    TC_HIP_TRY(tck::launch_count_order(a, ctx->cus, ctx->tune.grid, G.bin[gi], G.len[gi], tiles, gs));
  }
  for (int i = 0; i < side; ++i) {
    TC_HIP_TRY(hipEventRecord(ctx->join[i], ctx->side[i]));
    TC_HIP_TRY(hipStreamWaitEvent(s, ctx->join[i], 0));
  }
  TC_HIP_TRY(hipEventRecord(ctx->ev[2], s));
  TC_HIP_TRY(hipMemcpyAsync(ctx->pinned, a.bufs.outs, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  TC_HIP_TRY(hipMemcpyAsync(ctx->pinned + kPinnedFallback, a.bufs.counters + tck::kCntFallback, sizeof(uint32_t),
                            hipMemcpyDeviceToHost, s));
  TC_HIP_TRY(hipStreamSynchronize(s));
  *total = ctx->pinned[0];
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->total_ms = now_ms() - t0;
    float ms = 0;
    TC_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]));
    stats->plan_ms = ms;
    TC_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
    stats->count_ms = ms;
    stats->pivots = g->nv;
    for (int b = 0; b < TC_NUM_BINS; ++b) stats->bin_pivots[b] = bins[b];
    stats->hash_fallbacks = *(const uint32_t*)(ctx->pinned + kPinnedFallback);
    stats->plan_kept = kept ? 1 : 0;
  }
  return TC_OK;
}

// The binned count with the plan kept in the graph.  The plan (rows per bin,
// bin-major item lists) depends only on the graph's partner rows and the
// launch setting, so it is built on the first call: later calls re-arm the
// tile queues and launch the sized bins directly -- no planning kernels and
// no host round trip before the count.
static tc_status count_binned(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_stats* stats) {
  const int64_t P = (int64_t)g->nv_max + 1;
  const bool kept = g->plan_valid && std::memcmp(&g->plan_tune, &ctx->tune, sizeof(tc_tune)) == 0;
  if (!kept) {
    const int64_t had = g->plan.cap;
    tc_status st = g->plan.ensure(P, 1);
    if (st != TC_OK) return st;
    if (g->plan.cap != had) g->bytes += PlanScratch::bytes_for(g->plan.cap) - (had ? PlanScratch::bytes_for(had) : 0);
    g->plan_valid = false;
  }
  tc_status st = binned_count(ctx, g, g->plan.bufs(), kept, g->plan_bins, g->plan_tiles, total, stats);
  if (st == TC_OK && !kept) {
    g->plan_tune = ctx->tune;
    g->plan_valid = true;
  }
  return st;
}

// The plain kernel's single launch, its u64 sum at `out` on the device.
// kept_zero: `out` is the context's plain-count word, zero between calls:
// no zero launch before the kernel; it is cleared again after the result is
// read, off the caller's critical path (the next launch on the stream waits
// for it).
static tc_status plain_count(tc_ctx* ctx, tc_graph* g, uint64_t* out, uint64_t* total, tc_stats* stats,
                             bool kept_zero = false) {
  const double t0 = now_ms();
  const int64_t P = (int64_t)g->row_end - g->row_begin;
  hipStream_t s = ctx->stream;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(ctx->plain_grid, (P + tck::kWaves - 1) / tck::kWaves));
  // kept_zero: out is the context's pair of words, one per call in turn
  uint64_t* word = out;
  uint64_t* next = nullptr;
  if (kept_zero) {
    if (ctx->plain_dirty) TC_HIP_TRY(tck::launch_zero(out, 4, s));
    ctx->plain_dirty = true;
    word = out + ctx->plain_turn;
    next = out + (1 - ctx->plain_turn);
  } else {
    TC_HIP_TRY(tck::launch_zero(out, 2, s));
  }
  TC_HIP_TRY(hipEventRecord(ctx->ev[1], s));
  TC_HIP_TRY(tck::launch_plain(g->rp, g->ci, g->row_begin, g->row_end, g->own, word, grid, s, next));
  TC_HIP_TRY(hipEventRecord(ctx->ev[2], s));
  TC_HIP_TRY(hipMemcpyAsync(ctx->pinned, word, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  TC_HIP_TRY(hipStreamSynchronize(s));
  *total = ctx->pinned[0];
  if (kept_zero) {
    ctx->plain_turn = 1 - ctx->plain_turn;
    ctx->plain_dirty = false;
  }
  tc_status st = finish_stats(ctx, stats, t0, false, P);
  if (st == TC_OK && stats) stats->single_launch = 1;
  return st;
}

static tc_status count_plain(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_stats* stats) {
  return plain_count(ctx, g, ctx->plain_out, total, stats, true);
}

// Graphs that tc_count counts by the single plain launch (every list must
// fit a wave's LDS tree): those whose counted rows hold at most kPlainMaxIds
// DAG ids, where one launch beats planning and the binned kernels' launches;
// and flat graphs up to kPlainFlatMaxIds ids -- at most kPlainFlatMaxRows
// rows and out-lists of at most kPlainFlatMaxOut, which bounds the plain
// kernel's per-edge work -- where it still does.  Measured on MI355X (plain
// against binned, resident): ER n=65536 d=32 0.086 / 0.200 ms, ER n=131072
// d=32 0.136 / 0.211 ms, ER n=131072 d=16 0.102 / 0.126 ms; against that,
// ER n=262144 d=8 (more rows) 0.150 / 0.114 ms and RMAT n=32768 d=16 (lists
// to 189) 0.392 / 0.263 ms.
constexpr int64_t kPlainMaxIds = (int64_t)1 << 18;
constexpr int64_t kPlainFlatMaxIds = (int64_t)1 << 21;
constexpr int64_t kPlainFlatMaxRows = (int64_t)1 << 17;
constexpr int32_t kPlainFlatMaxOut = 64;

static bool plain_fits(int64_t rows, int64_t ids, int32_t max_out) {
  if (max_out >= tck::kCapWave) return false;
  return ids <= kPlainMaxIds ||
         (ids <= kPlainFlatMaxIds && rows <= kPlainFlatMaxRows && max_out <= kPlainFlatMaxOut);
}

static bool small_graph(const tc_graph* g) {
  return plain_fits((int64_t)g->row_end - g->row_begin, g->edge_end - g->edge_begin, g->max_out);
}

}  // namespace tcr

extern "C" {

tc_status tc_count(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_stats* stats) try {
  if (!ctx || !g || !total) return fail(TC_EINVAL, "null argument");
  if (g->ctx != ctx) return fail(TC_EINVAL, "graph belongs to another context");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  return small_graph(g) ? count_plain(ctx, g, total, stats) : count_binned(ctx, g, total, stats);
} catch (...) {
  return tcr::caught();
}

tc_status tc_count_csr(tc_ctx* ctx, const tc_csr_view* dag, uint32_t check, uint64_t* total, tc_stats* stats) try {
  if (!ctx || !total) return fail(TC_EINVAL, "null argument");
  if (!dag) return fail(TC_EINVAL, "null csr view");
  int32_t max_out = 0;
  tc_status st;
  auto whole = [&]() {  // upload + tc_count + free
    tc_graph* g = nullptr;
    tc_status s2 = upload_csr(ctx, dag, 0, dag->n_rows, nullptr, max_out, &g);
    if (s2 != TC_OK) return s2;
    s2 = tc_count(ctx, g, total, stats);
    tc_graph_free(ctx, g);
    return s2;
  };
  // A graph that tc_count would count by the single plain launch needs no
  // partner structure: it is copied to the context's grow-only buffer and
  // counted there.  One of at most kPlainMaxIds ids is copied first and
  // checked on the host while the copies run (the launch follows the
  // check); a larger one is checked first, so that one that turns out to
  // need the whole-graph path is not copied for nothing.  (Staging the
  // arrays in pinned memory for one copy measured no faster on MI355X.)
  const bool small = dag->n_rows >= 0 && dag->nnz >= 0 && dag->nnz <= tcr::kPlainMaxIds && dag->rowptr &&
                     (dag->nnz == 0 || dag->colidx);
  if (!small) {
    st = check_dag(dag, check, &max_out);
    if (st != TC_OK) return st;
  }
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  if (!small && !tcr::plain_fits(dag->n_rows, dag->nnz, max_out)) return whole();
  const size_t rp_bytes = sizeof(int64_t) * ((size_t)dag->n_rows + 1);
  const size_t ci_off = (rp_bytes + 255) & ~(size_t)255;
  const size_t bytes = ci_off + sizeof(int32_t) * (size_t)std::max<int64_t>(dag->nnz, 1);
  if (bytes > ctx->once_cap) {
    dev_free(ctx->once_dev);
    ctx->once_dev = nullptr;
    ctx->once_cap = 0;
    st = dev_alloc((void**)&ctx->once_dev, bytes, "one-shot graph");
    if (st != TC_OK) return st;
    ctx->once_cap = bytes;
  }
  hipStream_t s = ctx->stream;
  hipError_t ce = hipMemcpyAsync(ctx->once_dev, dag->rowptr, rp_bytes, hipMemcpyHostToDevice, s);
  if (ce == hipSuccess && dag->nnz > 0)
    ce = hipMemcpyAsync(ctx->once_dev + ci_off, dag->colidx, sizeof(int32_t) * (size_t)dag->nnz,
                        hipMemcpyHostToDevice, s);
  if (ce != hipSuccess) {
    (void)hipStreamSynchronize(s);  // a copy already queued may still read the caller's arrays
    return fail(TC_EDEVICE, std::string("one-shot copy: ") + hipGetErrorString(ce));
  }
  if (small) {
    st = check_dag(dag, check, &max_out);
    if (st != TC_OK || !tcr::plain_fits(dag->n_rows, dag->nnz, max_out)) {
      TC_HIP_TRY(hipStreamSynchronize(s));  // the copies read the caller's arrays
      return st != TC_OK ? st : whole();
    }
  }
  tc_graph g;  // a view of the copy: no partners, nothing to free
  g.ctx = ctx;
  g.rp = reinterpret_cast<int64_t*>(ctx->once_dev);
  g.ci = reinterpret_cast<int32_t*>(ctx->once_dev + ci_off);
  g.n = dag->n_rows;
  g.nnz = dag->nnz;
  g.row_begin = 0;
  g.row_end = dag->n_rows;
  g.edge_end = dag->nnz;
  g.max_out = max_out;
  return tcr::plain_count(ctx, &g, ctx->plain_out, total, stats, true);
} catch (...) {
  return tcr::caught();
}

tc_status tc_count_binned(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_stats* stats) try {
  if (!ctx || !g || !total) return fail(TC_EINVAL, "null argument");
  if (g->ctx != ctx) return fail(TC_EINVAL, "graph belongs to another context");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  return count_binned(ctx, g, total, stats);
} catch (...) {
  return tcr::caught();
}

tc_status tc_count_plain(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_stats* stats) try {
  if (!ctx || !g || !total) return fail(TC_EINVAL, "null argument");
  if (g->ctx != ctx) return fail(TC_EINVAL, "graph belongs to another context");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  return count_plain(ctx, g, total, stats);
} catch (...) {
  return tcr::caught();
}

tc_status tc_count_noalloc(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_stats* stats) try {
  if (!ctx || !g || !total) return fail(TC_EINVAL, "null argument");
  if (g->ctx != ctx) return fail(TC_EINVAL, "graph belongs to another context");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  const tck::PlanBufs bufs = tck::noalloc_plan_bufs(ctx->noalloc);
  if (small_graph(g)) return plain_count(ctx, g, bufs.outs, total, stats);
  const int64_t P = (int64_t)g->nv_max + 1;
  if (P > tck::kNoAllocSlots)
    return fail(TC_ERANGE, "tc_count_noalloc: the graph needs " + std::to_string(P) +
                               " plan slots (rows + ids/1024 + 2); the static scratch holds " +
                               std::to_string(tck::kNoAllocSlots));
  int64_t bins[tck::kNumBins], tiles[tck::kNumBins];
  return binned_count(ctx, g, bufs, false, bins, tiles, total, stats);
} catch (...) {
  return tcr::caught();
}

// First local count of g: its buffers, and a partner rebuild that also
// records every partner entry's DAG edge position.
// The split points and per-part sums of a local count (what a call with
// part sums alone needs beside the graph's partner-row owners).
constexpr int64_t kSplitBytes = 8 * tck::kMaxParts + 4 * (tck::kMaxParts + 1);
static tc_status ensure_split(LocalBufs& L) {
  if (L.pe) return TC_OK;
  tc_status st = dev_alloc((void**)&L.bounds, sizeof(int32_t) * (tck::kMaxParts + 1), "split points");
  if (st == TC_OK) st = dev_alloc((void**)&L.pe, sizeof(unsigned long long) * tck::kMaxParts, "part counts");
  if (st != TC_OK) {
    dev_free(L.bounds);
    L.bounds = nullptr;
    return st;
  }
  L.bytes = std::max(L.bytes, kSplitBytes);
  return TC_OK;
}

static tc_status ensure_local(tc_ctx* ctx, tc_graph* g) {
  LocalBufs& L = g->local;
  if (L.pedge) return TC_OK;
  const int64_t m = std::max<int64_t>(g->nnz, 1);
  tc_status st = dev_alloc((void**)&L.pedge, sizeof(int64_t) * (size_t)m, "partner edge positions");
  if (st == TC_OK) st = dev_alloc((void**)&L.pvid, sizeof(int32_t) * (size_t)m, "partner vertices");
  if (st == TC_OK) st = dev_alloc((void**)&L.ecount, sizeof(uint32_t) * (size_t)m, "edge counts");
  if (st == TC_OK)
    st = dev_alloc((void**)&L.vcount, sizeof(unsigned long long) * ((size_t)g->n + 1), "vertex counts");
  if (st == TC_OK) st = ensure_split(L);
  if (st == TC_OK) st = dev_alloc((void**)&L.esrc, sizeof(int32_t) * (size_t)m, "edge rows");
  if (st == TC_OK) {
    tck::LaunchArgs a = count_args(ctx, g);
    BuildScratch bs;
    hipError_t e = tck::launch_local_prep(g->rp, g->n, L.esrc, ctx->stream);  // the rebuild reads and marks esrc
    if (e == hipSuccess) st = launch_build(a, (int64_t)g->n + 1, g->nnz, bs, ctx->stream, &L, g->pown);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    bs.release();
    if (st == TC_OK && e != hipSuccess)
      st = fail(TC_EDEVICE, std::string("partner rebuild: ") + hipGetErrorString(e));
  }
  if (st != TC_OK) {
    L.release();
    return st;
  }
  g->plan_valid = false;  // the rebuild rewrote the partner rows the kept plan lists
  L.bytes = 20 * m + 8 * ((int64_t)g->n + 1) + kSplitBytes;
  return TC_OK;
}

tc_status tc_count_local(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_local* out, tc_stats* stats) try {
  if (!ctx || !g || !total || !out) return fail(TC_EINVAL, "null argument");
  if (g->ctx != ctx) return fail(TC_EINVAL, "graph belongs to another context");
  const int32_t k = out->parts;
  if (k < 0 || k > tck::kMaxParts) return fail(TC_EINVAL, "parts must be in [0, 1024]");
  if (k > 0) {
    if (!out->bounds || !out->part_counts) return fail(TC_EINVAL, "parts > 0 needs bounds and part_counts");
    if (out->bounds[0] != 0 || out->bounds[k] != g->n)
      return fail(TC_EINVAL, "split points must start at 0 and end at the graph's row count");
    for (int32_t p = 0; p < k; ++p)
      if (out->bounds[p] > out->bounds[p + 1]) return fail(TC_EINVAL, "split points must not decrease");
  }
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  const double t0 = now_ms();
  const int64_t P = (int64_t)g->nv_max + 1;
  // Without vertex_counts the kernels attribute to edges only; with part
  // sums alone, each partner row's hits go straight to its owner's part
  // (the part that computes its edges under the reference's push/pull rule):
  // no per-edge counts, per-part pass or local rebuild.
  const bool vertices = out->vertex_counts != nullptr;
  const bool parts_only = k > 0 && !vertices && !out->edge_counts;
  const int mode = vertices ? tck::kLocalFull : parts_only ? tck::kLocalParts : tck::kLocalEdges;
  tc_status st = ctx->scratch.ensure(P, 1);
  if (st == TC_OK) st = parts_only ? ensure_split(g->local) : ensure_local(ctx, g);
  if (st != TC_OK) return st;
  LocalBufs& L = g->local;
  tck::LaunchArgs a = count_args(ctx, g);
  a.local = tck::LocalOut{L.pedge, L.pvid, g->pown, L.ecount, L.vcount};
  a.local.bounds = L.bounds;
  a.local.parts = k;
  a.local.pe = reinterpret_cast<uint64_t*>(L.pe);
  hipStream_t s = ctx->stream;
  if (k > 0)
    TC_HIP_TRY(hipMemcpyAsync(L.bounds, out->bounds, sizeof(int32_t) * ((size_t)k + 1), hipMemcpyHostToDevice, s));
  TC_HIP_TRY(hipEventRecord(ctx->ev[0], s));
  TC_HIP_TRY(tck::launch_zero(a.bufs.counters, tck::kNumCounters, s));
  TC_HIP_TRY(tck::launch_zero(a.bufs.outs, 2, s));
  if (!parts_only) TC_HIP_TRY(tck::launch_zero(L.ecount, std::max<int64_t>(g->nnz, 1), s));
  if (vertices) TC_HIP_TRY(tck::launch_zero(L.vcount, 2 * ((int64_t)g->n + 1), s));
  TC_HIP_TRY(tck::launch_zero(L.pe, 2 * (int64_t)tck::kMaxParts, s));
  TC_HIP_TRY(tck::launch_plan(a, s));
  TC_HIP_TRY(hipEventRecord(ctx->ev[1], s));
  {  // bin groups on the library stream and the side streams, as in tc_count
    const tck::BinGroups& G = mode == tck::kLocalEdges ? tck::kEdgesBinOrder : tck::kBinOrder;
    for (int gi = 0; gi < G.n; ++gi) {
      hipStream_t gs = gi == 0 ? s : ctx->side[gi - 1];
      if (gi > 0) TC_HIP_TRY(hipStreamWaitEvent(gs, ctx->ev[1], 0));
      for (int j = 0; j < G.len[gi]; ++j)
        TC_HIP_TRY(tck::launch_count_local(a, ctx->cus, G.bin[gi][j], G.bin[gi][j], mode, gs));
      if (gi > 0) TC_HIP_TRY(hipEventRecord(ctx->join[gi - 1], gs));
    }
    for (int gi = 1; gi < G.n; ++gi) TC_HIP_TRY(hipStreamWaitEvent(s, ctx->join[gi - 1], 0));
  }
  if (k > 0 && !parts_only)
    TC_HIP_TRY(tck::launch_local_parts(g->ci, L.esrc, g->n, g->edge_begin, g->edge_end, L.ecount, L.bounds, k, L.pe,
                                       ctx->cus, s));
  TC_HIP_TRY(hipEventRecord(ctx->ev[2], s));
  TC_HIP_TRY(hipMemcpyAsync(ctx->pinned, a.bufs.outs, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  TC_HIP_TRY(hipMemcpyAsync(ctx->pinned + 8, a.bufs.counters, sizeof(uint32_t) * tck::kNumCounters,
                            hipMemcpyDeviceToHost, s));
  TC_HIP_TRY(hipStreamSynchronize(s));
  *total = ctx->pinned[0];
  const int64_t fallbacks = ((const uint32_t*)(ctx->pinned + 8))[tck::kCntFallback];
  // pageable destinations: synchronous copies after the stream is done
  if (out->edge_counts && g->edge_end > g->edge_begin)
    TC_HIP_TRY(hipMemcpy(out->edge_counts, L.ecount + g->edge_begin,
                         sizeof(uint32_t) * (size_t)(g->edge_end - g->edge_begin), hipMemcpyDeviceToHost));
  if (out->vertex_counts && g->n > 0)
    TC_HIP_TRY(hipMemcpy(out->vertex_counts, L.vcount, sizeof(uint64_t) * (size_t)g->n, hipMemcpyDeviceToHost));
  if (k > 0)
    TC_HIP_TRY(hipMemcpy(out->part_counts, L.pe, sizeof(uint64_t) * (size_t)k, hipMemcpyDeviceToHost));
  if (parts_only) {  // every row's hits went to a part
    uint64_t sum = 0;
    for (int32_t p = 0; p < k; ++p) sum += out->part_counts[p];
    *total = sum;
  }
  return finish_stats(ctx, stats, t0, true, g->nv, 1, fallbacks);
} catch (...) {
  return tcr::caught();
}

tc_status tc_probe_copy(tc_ctx* ctx, int64_t bytes, int32_t iters, double* ms_per_iter) try {
  if (!ctx || !ms_per_iter || bytes <= 0 || bytes % 16 != 0 || iters < 1)
    return fail(TC_EINVAL, "need bytes > 0, a multiple of 16, and iters >= 1");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  void *src = nullptr, *dst = nullptr;
  tc_status st = dev_alloc(&src, (size_t)bytes, "copy src");
  if (st == TC_OK) st = dev_alloc(&dst, (size_t)bytes, "copy dst");
  if (st != TC_OK) {
    dev_free(src);
    return st;
  }
  hipStream_t s = ctx->stream;
  hipError_t e = hipMemsetAsync(src, 1, (size_t)bytes, s);
  const int grid = ctx->cus * 8;
  if (e == hipSuccess) e = tck::launch_copy(src, dst, bytes, grid, s);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(ctx->ev[1], s);
  for (int i = 0; i < iters && e == hipSuccess; ++i) e = tck::launch_copy(src, dst, bytes, grid, s);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev[2], s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]);
  dev_free(src);
  dev_free(dst);
  if (e != hipSuccess) return fail(TC_EDEVICE, std::string("copy probe: ") + hipGetErrorString(e));
  *ms_per_iter = ms / iters;
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

tc_status tc_probe_read(tc_ctx* ctx, int64_t bytes, int32_t iters, double* ms_per_iter) try {
  if (!ctx || !ms_per_iter || bytes <= 0 || bytes % 16 != 0 || iters < 1)
    return fail(TC_EINVAL, "need bytes > 0, a multiple of 16, and iters >= 1");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DeviceGuard guard(ctx->device);
  const int grid = ctx->cus * 32;
  void *src = nullptr, *sink = nullptr;
  tc_status st = dev_alloc(&src, (size_t)bytes, "read src");
  if (st == TC_OK) st = dev_alloc(&sink, sizeof(uint32_t) * (size_t)grid * tck::kBlock, "read sink");
  if (st != TC_OK) {
    dev_free(src);
    return st;
  }
  hipStream_t s = ctx->stream;
  hipError_t e = hipMemsetAsync(src, 1, (size_t)bytes, s);
  uint32_t* k = (uint32_t*)sink;
  if (e == hipSuccess) e = tck::launch_read(src, bytes, k, grid, s);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(ctx->ev[1], s);
  for (int i = 0; i < iters && e == hipSuccess; ++i) e = tck::launch_read(src, bytes, k, grid, s);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev[2], s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]);
  dev_free(src);
  dev_free(sink);
  if (e != hipSuccess) return fail(TC_EDEVICE, std::string("read probe: ") + hipGetErrorString(e));
  *ms_per_iter = ms / iters;
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

}  // extern "C"
