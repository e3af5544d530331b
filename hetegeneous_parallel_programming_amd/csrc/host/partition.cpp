// Partitioning of a DAG for one GPU's share of the multi-GPU count, and the
// algorithmic work model of a count.
//
// Reference behaviour followed here:
//   contiguous ranges        successful_push_pull/graph_partition.cc:85-107
//                            (chunk = ceil(n/k); [p*chunk, min(n, (p+1)*chunk)))
//   induced subgraph         successful_push_pull/graph_partition.cc:27-82, 108-131
//   partition-local pivots   successful_push_pull/bs_warp_vertex_nvshmem.cuh:27-40
//   pulled remote lists      successful_push_pull/bs_warp_vertex_nvshmem.cuh:110-136
// The pull buffer / graph_cache.h role becomes a halo cache that holds every
// remote list the owned rows reach, staged once instead of fetched per edge.
#include <algorithm>

#include "host_common.hpp"

struct tch_partition {
  int32_t n_global = 0, row_begin = 0, row_end = 0;
  tch::Csr local;  // n = owned rows, global ids in colidx
  int32_t n_halo = 0;
  std::unique_ptr<int32_t[]> halo_ids;
  tch::Csr halo;   // n = n_halo
};

namespace tch {

static void range_of(int32_t n, int32_t k, int32_t part, int32_t& b, int32_t& e) {
  int64_t chunk = n / k;
  if (n % k != 0) ++chunk;  // graph_partition.cc:88-89
  int64_t lo = chunk * part, hi = lo + chunk;
  if (hi > n) hi = n;       // graph_partition.cc:105-107
  if (lo > n) lo = n;
  b = (int32_t)lo;
  e = (int32_t)hi;
}

// Marks then lists, in ascending order, the vertices with mark[v] != 0.
static int32_t compact_marked(const uint8_t* mark, int32_t n, std::unique_ptr<int32_t[]>& ids) {
  std::unique_ptr<int64_t[]> pos(new int64_t[(size_t)n + 1]);
  exclusive_scan(mark, n, pos.get());
  int32_t cnt = (int32_t)pos[n];
  ids.reset(new int32_t[cnt > 0 ? cnt : 1]);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < n; ++v)
    if (mark[v]) ids[pos[v]] = (int32_t)v;
  return cnt;
}

}  // namespace tch

using namespace tch;

extern "C" {

tc_status tch_partition_range(const tc_csr_view* dag, int32_t k, int32_t part, tch_partition** out) try {
  if (!dag || !out) return fail(TC_EINVAL, "null argument");
  if (k < 1 || part < 0 || part >= k) return fail(TC_EINVAL, "need k >= 1 and 0 <= part < k");
  int32_t b, e;
  range_of(dag->n_rows, k, part, b, e);
  return tch_partition_span(dag, b, e, out);
} catch (...) {
  return tch::caught();
}

// PARTNER weights, in streamed-id equivalents scaled by kWeightOne (fixed
// point, so that every platform finds the same split): per DAG edge the ids
// the partner scheme streams, min(d+(u), d+(v)), plus kPartnerEdgeIds for
// its list reference, and its share of its owner's partner rows
// (ceil(c / TCH_PARTNERS_PER_ROW) rows over the owner's c edges), each
// row costing kPartnerRowIds (the row record, the owner's set built in LDS,
// the plan entry).  The two constants are a least-squares fit of the count
// kernel times of the eight parts of RMAT scale 24's 4-way vertex and
// partner splits on MI355X (0.52 ms per G streamed ids, 0.78 ms per M
// partner rows; largest error 0.34 ms); they took the slowest part of the
// partner split from 10.95 to 10.52 ms against edge weights of 2 ids and no
// row term.
constexpr int64_t kWeightOne = 1 << 16;
constexpr int64_t kPartnerEdgeIds = 6;
constexpr int64_t kPartnerRowIds = 1503;

tc_status tch_partition_bounds(const tc_csr_view* dag, int32_t k, int32_t balance, int32_t* bounds) try {
  if (!dag || !bounds) return fail(TC_EINVAL, "null argument");
  if (k < 1) return fail(TC_EINVAL, "need k >= 1");
  if (balance < TCH_BALANCE_VERTICES || balance > TCH_BALANCE_PARTNER)
    return fail(TC_EINVAL, "unknown balance criterion");
  const int32_t n = dag->n_rows;
  if (balance == TCH_BALANCE_VERTICES) {
    for (int32_t i = 0; i < k; ++i) range_of(n, k, i, bounds[i], bounds[i + 1]);
    return TC_OK;
  }
  const int64_t* rp = dag->rowptr;
  const int32_t* ci = dag->colidx;
  // PARTNER: the edges each vertex owns (the endpoint with the longer
  // out-list, the row on a tie), for each edge's share of its owner's
  // partner rows.
  std::unique_ptr<int32_t[]> own;
  if (balance == TCH_BALANCE_PARTNER) {
    own.reset(new int32_t[(size_t)n + 1]());
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t u = 0; u < n; ++u) {
      const int64_t du = rp[u + 1] - rp[u];
      for (int64_t x = rp[u]; x < rp[u + 1]; ++x) {
        const int32_t v = ci[x];
        const int32_t o = rp[v + 1] - rp[v] > du ? v : (int32_t)u;
        __atomic_fetch_add(&own[o], 1, __ATOMIC_RELAXED);
      }
    }
  }
  std::unique_ptr<int64_t[]> w(new int64_t[(size_t)n + 1]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < n; ++u) {
    const int64_t du = rp[u + 1] - rp[u];
    int64_t c = du;
    if (balance == TCH_BALANCE_WEDGES) {
      for (int64_t x = rp[u]; x < rp[u + 1]; ++x) c += rp[ci[x] + 1] - rp[ci[x]];
    } else if (balance == TCH_BALANCE_PARTNER) {
      c = 0;
      for (int64_t x = rp[u]; x < rp[u + 1]; ++x) {
        const int32_t v = ci[x];
        const int64_t dv = rp[v + 1] - rp[v];
        const int64_t oc = own[dv > du ? v : (int32_t)u];
        c += (std::min(du, dv) + kPartnerEdgeIds) * kWeightOne +
             kPartnerRowIds * kWeightOne * ((oc + TCH_PARTNERS_PER_ROW - 1) / TCH_PARTNERS_PER_ROW) / oc;
      }
    }
    w[u] = c;
  }
  own.reset();
  std::unique_ptr<int64_t[]> pre(new int64_t[(size_t)n + 1]);
  exclusive_scan(w.get(), n, pre.get());
  w.reset();
  const int64_t total = pre[n];
  bounds[0] = 0;
  bounds[k] = n;
  for (int32_t i = 1; i < k; ++i) {
    // split where the prefix is nearest to i/k of the total
    const int64_t t = (int64_t)((__int128)total * i / k);
    int64_t r = std::lower_bound(pre.get(), pre.get() + n + 1, t) - pre.get();
    if (r > 0 && t - pre[r - 1] < pre[r] - t) --r;
    bounds[i] = std::max(bounds[i - 1], (int32_t)r);
  }
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_partition_span(const tc_csr_view* dag, int32_t row_begin, int32_t row_end,
                             tch_partition** out) try {
  if (!dag || !out) return fail(TC_EINVAL, "null argument");
  const int32_t n = dag->n_rows;
  if (row_begin < 0 || row_end > n || row_begin > row_end)
    return fail(TC_EINVAL, "owned range outside [0, n]");
  std::unique_ptr<tch_partition> p(new tch_partition);
  p->n_global = n;
  p->row_begin = row_begin;
  p->row_end = row_end;
  const int32_t b = p->row_begin, e = p->row_end, nl = e - b;
  const int64_t e0 = dag->rowptr[b];
  p->local.allocate(nl, dag->rowptr[e] - e0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i <= nl; ++i) p->local.rowptr[i] = dag->rowptr[b + i] - e0;
  std::memcpy(p->local.colidx.get(), dag->colidx + e0, sizeof(int32_t) * (size_t)p->local.nnz);
  p->local.compute_max_degree();

  // Halo: every remote v reached by an owned row.
  std::unique_ptr<uint8_t[]> mark(new uint8_t[n > 0 ? n : 1]);
  std::memset(mark.get(), 0, (size_t)n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < p->local.nnz; ++i) {
    int32_t v = p->local.colidx[i];
    if (v < b || v >= e) __atomic_store_n(&mark[v], (uint8_t)1, __ATOMIC_RELAXED);
  }
  p->n_halo = compact_marked(mark.get(), n, p->halo_ids);
  mark.reset();
  const int32_t nh = p->n_halo;
  std::unique_ptr<int64_t[]> hdeg(new int64_t[(size_t)nh + 1]);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nh; ++i) {
    int32_t v = p->halo_ids[i];
    hdeg[i] = dag->rowptr[v + 1] - dag->rowptr[v];
  }
  p->halo.allocate(nh, 0);
  exclusive_scan(hdeg.get(), nh, p->halo.rowptr.get());
  p->halo.nnz = p->halo.rowptr[nh];
  p->halo.colidx.reset(new int32_t[p->halo.nnz > 0 ? p->halo.nnz : 1]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < nh; ++i) {
    int32_t v = p->halo_ids[i];
    std::memcpy(p->halo.colidx.get() + p->halo.rowptr[i], dag->colidx + dag->rowptr[v],
                sizeof(int32_t) * (size_t)hdeg[i]);
  }
  p->halo.compute_max_degree();
  *out = p.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_partition_get(const tch_partition* p, tch_partition_view* o) try {
  if (!p || !o) return fail(TC_EINVAL, "null argument");
  o->n_global = p->n_global;
  o->row_begin = p->row_begin;
  o->row_end = p->row_end;
  o->local_rowptr = p->local.rowptr.get();
  o->local_colidx = p->local.colidx.get();
  o->local_nnz = p->local.nnz;
  o->n_halo = p->n_halo;
  o->halo_ids = p->halo_ids.get();
  o->halo_rowptr = p->halo.rowptr.get();
  o->halo_colidx = p->halo.colidx.get();
  o->halo_nnz = p->halo.nnz;
  return TC_OK;
} catch (...) {
  return tch::caught();
}

void tch_partition_free(tch_partition* p) { delete p; }

tc_status tch_partition_assemble(const tch_partition_view* p, tch_csr** out) try {
  if (!p || !out) return fail(TC_EINVAL, "null argument");
  const int32_t n = p->n_global, b = p->row_begin, e = p->row_end;
  if (n < 0 || b < 0 || e > n || b > e || p->n_halo < 0)
    return fail(TC_EINVAL, "partition view ranges are inconsistent");
  std::unique_ptr<int64_t[]> len(new int64_t[(size_t)n + 1]);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < n; ++v) len[v] = 0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < e - b; ++i) len[b + i] = p->local_rowptr[i + 1] - p->local_rowptr[i];
  int64_t bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (int64_t h = 0; h < p->n_halo; ++h) {
    const int32_t v = p->halo_ids[h];
    if (v < 0 || v >= n || (v >= b && v < e) || (h > 0 && p->halo_ids[h - 1] >= v)) { ++bad; continue; }
    len[v] = p->halo_rowptr[h + 1] - p->halo_rowptr[h];
  }
  if (bad) return fail(TC_EINVAL, "halo ids must be sorted, unique, in range and not owned");
  std::unique_ptr<tch_csr> g(new tch_csr);
  g->g.allocate(n, 0);
  exclusive_scan(len.get(), n, g->g.rowptr.get());
  g->g.nnz = g->g.rowptr[n];
  g->g.colidx.reset(new int32_t[g->g.nnz > 0 ? g->g.nnz : 1]);
  if (e > b)
    std::memcpy(g->g.colidx.get() + g->g.rowptr[b], p->local_colidx,
                sizeof(int32_t) * (size_t)(p->local_rowptr[e - b] - p->local_rowptr[0]));
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t h = 0; h < p->n_halo; ++h) {
    const int32_t v = p->halo_ids[h];
    std::memcpy(g->g.colidx.get() + g->g.rowptr[v], p->halo_colidx + p->halo_rowptr[h],
                sizeof(int32_t) * (size_t)len[v]);
  }
  g->g.compute_max_degree();
  *out = g.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

void tch_free_ids(int32_t* p) { delete[] p; }

tc_status tch_partition_induced(const tc_csr_view* g, int32_t k, int32_t part, tch_csr** sub_out,
                                int32_t* local_begin, int32_t* local_end, int32_t** idx_map_out) try {
  if (!g || !sub_out || !local_begin || !local_end || !idx_map_out)
    return fail(TC_EINVAL, "null argument");
  if (k < 1 || part < 0 || part >= k) return fail(TC_EINVAL, "need k >= 1 and 0 <= part < k");
  const int32_t n = g->n_rows;
  int32_t b, e;
  range_of(n, k, part, b, e);
  std::unique_ptr<uint8_t[]> mark(new uint8_t[n > 0 ? n : 1]);
  std::memset(mark.get(), 0, (size_t)n);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = b; v < e; ++v) {
    __atomic_store_n(&mark[v], (uint8_t)1, __ATOMIC_RELAXED);
    for (int64_t x = g->rowptr[v]; x < g->rowptr[v + 1]; ++x)
      __atomic_store_n(&mark[g->colidx[x]], (uint8_t)1, __ATOMIC_RELAXED);
  }
  std::unique_ptr<int32_t[]> ids;
  const int32_t nv = compact_marked(mark.get(), n, ids);
  std::unique_ptr<int32_t[]> new_id(new int32_t[n > 0 ? n : 1]);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nv; ++i) new_id[ids[i]] = (int32_t)i;
  std::unique_ptr<int64_t[]> deg(new int64_t[(size_t)nv + 1]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < nv; ++i) {
    int32_t v = ids[i];
    int64_t c = 0;
    for (int64_t x = g->rowptr[v]; x < g->rowptr[v + 1]; ++x) c += mark[g->colidx[x]] ? 1 : 0;
    deg[i] = c;
  }
  std::unique_ptr<tch_csr> sub(new tch_csr);
  sub->g.allocate(nv, 0);
  exclusive_scan(deg.get(), nv, sub->g.rowptr.get());
  sub->g.nnz = sub->g.rowptr[nv];
  sub->g.colidx.reset(new int32_t[sub->g.nnz > 0 ? sub->g.nnz : 1]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < nv; ++i) {
    int32_t v = ids[i];
    int64_t o = sub->g.rowptr[i];
    for (int64_t x = g->rowptr[v]; x < g->rowptr[v + 1]; ++x) {
      int32_t u = g->colidx[x];
      if (mark[u]) sub->g.colidx[o++] = new_id[u];
    }
  }
  sub->g.compute_max_degree();
  // Owned ids are contiguous globally, hence contiguous among the ascending local ids.
  *local_begin = (e > b) ? new_id[b] : 0;
  *local_end = (e > b) ? new_id[e - 1] + 1 : 0;
  int32_t* map = new int32_t[nv > 0 ? nv : 1];
  std::memcpy(map, ids.get(), sizeof(int32_t) * (size_t)nv);
  *idx_map_out = map;
  *sub_out = sub.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

}  // extern "C"

namespace {

// Owner counts of the partner scheme over the edges of `rows`; deg(v) and
// slot(v) give d+ and the owner-count slot of any vertex the rows reach.
template <class Deg, class Slot>
static void partner_units(int64_t nl, const int64_t* rowptr, const int32_t* colidx, int64_t first_row,
                          std::vector<int32_t>& cnt, Deg deg, Slot slot, int64_t* units, int64_t* unit_bytes) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < nl; ++i) {
    const int64_t du = rowptr[i + 1] - rowptr[i];
    int32_t own = 0;
    for (int64_t x = rowptr[i]; x < rowptr[i + 1]; ++x) {
      const int32_t v = colidx[x];
      if (deg(v) <= du) {
        ++own;
      } else {
#pragma omp atomic
        cnt[slot(v)] += 1;
      }
    }
#pragma omp atomic
    cnt[slot((int32_t)(first_row + i))] += own;
  }
  int64_t u = 0, b = 0;
  const int64_t ns = (int64_t)cnt.size();
#pragma omp parallel for reduction(+ : u, b)
  for (int64_t k = 0; k < ns; ++k) {
    const int64_t c = cnt[k];
    if (!c) continue;
    const int64_t K = (c + TCH_PARTNERS_PER_ROW - 1) / TCH_PARTNERS_PER_ROW;
    u += K;
    b += K * 32;
  }
  *units = u;
  *unit_bytes = b;
}

}  // namespace

extern "C" {

tc_status tch_work_of(const tc_csr_view* dag, int32_t row_begin, int32_t row_end, tch_work* out) try {
  if (!dag || !out) return fail(TC_EINVAL, "null argument");
  if (row_begin < 0 || row_end > dag->n_rows || row_begin > row_end)
    return fail(TC_EINVAL, "row range outside [0, n]");
  int64_t m = 0, wedge = 0, mins = 0;
  int32_t mo = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : m, wedge, mins) reduction(max : mo)
  for (int64_t u = row_begin; u < row_end; ++u) {
    const int64_t du = dag->rowptr[u + 1] - dag->rowptr[u];
    m += du;
    if (du > mo) mo = (int32_t)du;
    for (int64_t x = dag->rowptr[u]; x < dag->rowptr[u + 1]; ++x) {
      const int32_t v = dag->colidx[x];
      const int64_t dv = dag->rowptr[v + 1] - dag->rowptr[v];
      wedge += dv;
      mins += std::min(du, dv);
    }
  }
  const int64_t R = row_end - row_begin;
  out->rows = R;
  out->edges = m;
  out->wedge_elems = wedge;
  out->min_elems = mins;
  out->alg_bytes = 8 * (R + 1) + 20 * m + 4 * wedge;
  out->max_out = mo;
  std::vector<int32_t> cnt((size_t)dag->n_rows, 0);
  const int64_t* rp = dag->rowptr;
  int64_t units = 0, ub = 0;
  partner_units(R, rp + row_begin, dag->colidx, row_begin, cnt,
                [rp](int32_t v) { return rp[v + 1] - rp[v]; }, [](int32_t v) { return (int64_t)v; }, &units, &ub);
  // the owners' lists: 4 d+(x) per unit
  int64_t setb = 0;
  const int64_t n = dag->n_rows;
#pragma omp parallel for reduction(+ : setb)
  for (int64_t x = 0; x < n; ++x)
    if (cnt[x]) setb += 4 * ((cnt[x] + TCH_PARTNERS_PER_ROW - 1) / TCH_PARTNERS_PER_ROW) * (rp[x + 1] - rp[x]);
  out->partner_rows = units;
  out->partner_bytes = ub + setb + 8 * m + 4 * mins;
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_work_of_partition(const tch_partition_view* p, tch_work* out) try {
  if (!p || !out) return fail(TC_EINVAL, "null argument");
  const int32_t b = p->row_begin, nl = p->row_end - p->row_begin;
  int64_t m = 0, wedge = 0, mins = 0;
  int32_t mo = 0;
  int64_t missing = 0;
  auto halo_of = [p](int32_t v) -> int64_t {
    const int32_t* it = std::lower_bound(p->halo_ids, p->halo_ids + p->n_halo, v);
    return (it == p->halo_ids + p->n_halo || *it != v) ? -1 : (int64_t)(it - p->halo_ids);
  };
  auto deg = [p, b, halo_of](int32_t v) -> int64_t {
    if (v >= b && v < p->row_end) return p->local_rowptr[v - b + 1] - p->local_rowptr[v - b];
    const int64_t h = halo_of(v);
    return h < 0 ? -1 : p->halo_rowptr[h + 1] - p->halo_rowptr[h];
  };
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : m, wedge, mins, missing) reduction(max : mo)
  for (int64_t i = 0; i < nl; ++i) {
    const int64_t du = p->local_rowptr[i + 1] - p->local_rowptr[i];
    m += du;
    if (du > mo) mo = (int32_t)du;
    for (int64_t x = p->local_rowptr[i]; x < p->local_rowptr[i + 1]; ++x) {
      const int64_t dv = deg(p->local_colidx[x]);
      if (dv < 0) { ++missing; continue; }
      wedge += dv;
      mins += std::min(du, dv);
    }
  }
  if (missing) return fail(TC_EINVAL, "halo cache misses a remote list reached by an owned row");
  out->rows = nl;
  out->edges = m;
  out->wedge_elems = wedge;
  out->min_elems = mins;
  out->alg_bytes = 8 * ((int64_t)nl + 1) + 20 * m + 4 * wedge;
  out->max_out = mo;
  // owner slots: local rows [0, nl), then halo lists [nl, nl + n_halo)
  std::vector<int32_t> cnt((size_t)nl + (size_t)p->n_halo, 0);
  auto slot = [p, b, nl, halo_of](int32_t v) -> int64_t {
    if (v >= b && v < p->row_end) return v - b;
    return nl + halo_of(v);
  };
  int64_t units = 0, ub = 0;
  partner_units(nl, p->local_rowptr, p->local_colidx, b, cnt, deg, slot, &units, &ub);
  int64_t setb = 0;
  const int64_t ns = (int64_t)cnt.size();
#pragma omp parallel for reduction(+ : setb)
  for (int64_t k = 0; k < ns; ++k) {
    if (!cnt[k]) continue;
    const int64_t d = k < nl ? p->local_rowptr[k + 1] - p->local_rowptr[k]
                             : p->halo_rowptr[k - nl + 1] - p->halo_rowptr[k - nl];
    setb += 4 * ((cnt[k] + TCH_PARTNERS_PER_ROW - 1) / TCH_PARTNERS_PER_ROW) * d;
  }
  out->partner_rows = units;
  out->partner_bytes = ub + setb + 8 * m + 4 * mins;
  return TC_OK;
} catch (...) {
  return tch::caught();
}

}  // extern "C"
