// Seeded synthetic graphs in place of the GraphMiner input files the
// reference reads (successful_push_pull/main.cc:14).  Every pair is a pure
// function of (seed, pair index), so output does not depend on thread count.
#include <algorithm>
#include <cmath>

#include "host_common.hpp"

namespace tch {

int32_t ceil_log2(int64_t n) {
  int32_t s = 0;
  while (((int64_t)1 << s) < n) ++s;
  return s;
}

// Multiply-high maps a uniform 32-bit draw onto [0, n).
static inline int32_t scale32(uint32_t r, int32_t n) {
  return (int32_t)(((uint64_t)r * (uint64_t)n) >> 32);
}

void gen_er_pairs(int32_t n, int64_t m, uint64_t seed, int32_t* src, int32_t* dst, int threads) {
  const int T = threads <= 1 ? 1 : threads;
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    const uint64_t r = draw(seed, (uint64_t)i, 0);
    const int32_t u = scale32((uint32_t)(r >> 32), n);
    const int32_t v = scale32((uint32_t)r, n);
    if (u == v) { src[i] = -1; dst[i] = -1; continue; }
    src[i] = u;
    dst[i] = v;
  }
}

// RMAT: `scale` recursive quadrant choices per pair; level l uses the high
// (even l) or low (odd l) 32 bits of draw (seed, i, l/2), compared against
// floor(p * 2^32) cumulative thresholds.
void gen_rmat_pairs(int32_t n, int32_t scale, int64_t m, double a, double b, double c,
                    uint64_t seed, int32_t* src, int32_t* dst, int threads) {
  const int T = threads <= 1 ? 1 : threads;
  const double two32 = 4294967296.0;
  const uint64_t ta = (uint64_t)std::floor(a * two32);
  const uint64_t tb = (uint64_t)std::floor((a + b) * two32);
  const uint64_t tc = (uint64_t)std::floor((a + b + c) * two32);
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    uint32_t u = 0, v = 0;
    uint64_t r = 0;
    for (int32_t l = 0; l < scale; ++l) {
      if ((l & 1) == 0) r = draw(seed, (uint64_t)i, (uint64_t)(l >> 1));
      const uint64_t x = (l & 1) ? (r & 0xffffffffull) : (r >> 32);
      uint32_t bu, bv;
      if (x < ta) { bu = 0; bv = 0; }
      else if (x < tb) { bu = 0; bv = 1; }
      else if (x < tc) { bu = 1; bv = 0; }
      else { bu = 1; bv = 1; }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    if ((int64_t)u >= n || (int64_t)v >= n || u == v) { src[i] = -1; dst[i] = -1; continue; }
    src[i] = (int32_t)u;
    dst[i] = (int32_t)v;
  }
}

static tc_status gen_er(int32_t n, double avg_degree, uint64_t seed, int threads, Csr& out) {
  if (n < 1) return fail(TC_EINVAL, "ER needs n >= 1");
  if (!(avg_degree >= 0) || avg_degree > 1e6) return fail(TC_EINVAL, "ER avg_degree out of range");
  const int64_t m = llround((double)n * avg_degree / 2.0);
  std::unique_ptr<int32_t[]> s(new int32_t[m > 0 ? m : 1]), d(new int32_t[m > 0 ? m : 1]);
  gen_er_pairs(n, m, seed, s.get(), d.get(), threads);
  return build_symmetric(n, s.get(), d.get(), m, threads, out);
}

static tc_status gen_rmat(int32_t n, double ef, double a, double b, double c, uint64_t seed,
                          int threads, Csr& out) {
  if (n < 2) return fail(TC_EINVAL, "RMAT needs n >= 2");
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0))
    return fail(TC_EINVAL, "RMAT probabilities must be >= 0 with a+b+c <= 1");
  if (!(ef >= 0) || ef > 1e6) return fail(TC_EINVAL, "RMAT edge factor out of range");
  const int32_t scale = ceil_log2(n);
  const int64_t m = llround(ef * (double)n);
  std::unique_ptr<int32_t[]> s(new int32_t[m > 0 ? m : 1]), d(new int32_t[m > 0 ? m : 1]);
  gen_rmat_pairs(n, scale, m, a, b, c, seed, s.get(), d.get(), threads);
  return build_symmetric(n, s.get(), d.get(), m, threads, out);
}

static tc_status batch_spec(uint64_t batch_seed, int32_t i, tch_batch_spec& sp) {
  if (i < 0) return fail(TC_EINVAL, "negative batch index");
  const double u_n = unit(draw(batch_seed, (uint64_t)i, 1));
  const double u_d = unit(draw(batch_seed, (uint64_t)i, 2));
  int64_t n = (int64_t)std::floor(std::exp2(8.0 + 8.0 * u_n));
  if (n < 256) n = 256;
  if (n > 65536) n = 65536;
  sp.kind = (i & 1) ? 1 : 0;
  sp.n = (int32_t)n;
  sp.avg_degree = 4.0 + 60.0 * u_d;
  sp.seed = draw(batch_seed, (uint64_t)i, 3);
  return TC_OK;
}

static tc_status batch_graph(uint64_t batch_seed, int32_t i, int threads, Csr& dag) {
  tch_batch_spec sp;
  tc_status st = batch_spec(batch_seed, i, sp);
  if (st != TC_OK) return st;
  Csr sym;
  if (sp.kind == 0)
    st = gen_er(sp.n, sp.avg_degree, sp.seed, threads, sym);
  else
    st = gen_rmat(sp.n, sp.avg_degree / 2.0, 0.57, 0.19, 0.19, sp.seed, threads, sym);
  if (st != TC_OK) return st;
  return orient(sym.view(), threads, dag);
}

}  // namespace tch

using namespace tch;

extern "C" {

tc_status tch_gen_er(int32_t n, double avg_degree, uint64_t seed, tch_csr** out) try {
  if (!out) return fail(TC_EINVAL, "null out");
  std::unique_ptr<tch_csr> h(new tch_csr);
  tc_status st = gen_er(n, avg_degree, seed, omp_get_max_threads(), h->g);
  if (st != TC_OK) return st;
  *out = h.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_gen_rmat(int32_t n, double edge_factor, double a, double b, double c, uint64_t seed,
                       tch_csr** out) try {
  if (!out) return fail(TC_EINVAL, "null out");
  std::unique_ptr<tch_csr> h(new tch_csr);
  tc_status st = gen_rmat(n, edge_factor, a, b, c, seed, omp_get_max_threads(), h->g);
  if (st != TC_OK) return st;
  *out = h.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_batch_spec_of(uint64_t batch_seed, int32_t index, tch_batch_spec* out) try {
  if (!out) return fail(TC_EINVAL, "null out");
  return batch_spec(batch_seed, index, *out);
} catch (...) {
  return tch::caught();
}

tc_status tch_batch_graph(uint64_t batch_seed, int32_t index, tch_csr** dag_out) try {
  if (!dag_out) return fail(TC_EINVAL, "null out");
  std::unique_ptr<tch_csr> h(new tch_csr);
  tc_status st = batch_graph(batch_seed, index, omp_get_max_threads(), h->g);
  if (st != TC_OK) return st;
  *dag_out = h.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_batch_graphs(uint64_t batch_seed, int32_t first, int32_t count, tch_csr** dags_out) try {
  if (!dags_out || first < 0 || count < 0) return fail(TC_EINVAL, "bad batch range");
  std::vector<tc_status> st(count, TC_OK);
  std::vector<std::string> err(count);
  for (int32_t k = 0; k < count; ++k) dags_out[k] = nullptr;
  // Largest graphs first so the dynamic schedule ends evenly.
  std::vector<std::pair<int64_t, int32_t>> order(count);
  for (int32_t k = 0; k < count; ++k) {
    tch_batch_spec sp;
    batch_spec(batch_seed, first + k, sp);
    order[k] = {-(int64_t)(sp.n * sp.avg_degree), k};
  }
  std::sort(order.begin(), order.end());
#pragma omp parallel for schedule(dynamic, 1)
  for (int32_t j = 0; j < count; ++j) {
    const int32_t k = order[j].second;
    std::unique_ptr<tch_csr> h(new tch_csr);
    st[k] = batch_graph(batch_seed, first + k, 1, h->g);
    if (st[k] == TC_OK) dags_out[k] = h.release();
    else err[k] = tch_last_error();
  }
  for (int32_t k = 0; k < count; ++k) {
    if (st[k] != TC_OK) {
      for (int32_t j = 0; j < count; ++j) { delete dags_out[j]; dags_out[j] = nullptr; }
      return fail(st[k], err[k]);
    }
  }
  return TC_OK;
} catch (...) {
  return tch::caught();
}

}  // extern "C"
