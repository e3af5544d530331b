// Host CSR substrate: symmetric CSR construction, validation, degree-ordered
// DAG orientation, neighbour sorting and GraphMiner binary I/O.
//
// Reference behaviour followed here:
//   orientation       successful_push_pull/graph.cc:252-298
//   sort_neighbors    successful_push_pull/graph.cc:156-164
//   loader / format   successful_push_pull/graph.cc:31-55 (meta.txt, vertex.bin, edge.bin)
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <fstream>
#include <new>
#include <sstream>
#include <stdexcept>

#include <sys/stat.h>

#include "host_common.hpp"

namespace tch {

static thread_local std::string g_last_error;

tc_status fail(tc_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}
void clear_error() { g_last_error.clear(); }

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

int64_t file_bytes(const std::string& path) {
  struct stat sb;
  if (::stat(path.c_str(), &sb) != 0 || !S_ISREG(sb.st_mode)) return -1;
  return (int64_t)sb.st_size;
}

bool read_exact(const std::string& path, void* dst, size_t bytes) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) return false;
  size_t got = bytes ? std::fread(dst, 1, bytes, fp) : 0;
  int extra = std::fgetc(fp);
  std::fclose(fp);
  return got == bytes && extra == EOF;
}

bool write_exact(const std::string& path, const void* src, size_t bytes) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) return false;
  size_t put = bytes ? std::fwrite(src, 1, bytes, fp) : 0;
  return std::fclose(fp) == 0 && put == bytes;
}

void Csr::compute_max_degree() {
  int32_t md = 0;
#pragma omp parallel for reduction(max : md) schedule(static)
  for (int64_t v = 0; v < n; ++v) {
    int64_t d = rowptr[v + 1] - rowptr[v];
    if (d > md) md = (int32_t)d;
  }
  max_degree = md;
}

tc_status build_symmetric(int32_t n, const int32_t* src, const int32_t* dst, int64_t m,
                          int threads, Csr& out) {
  if (n < 0 || m < 0) return fail(TC_EINVAL, "negative vertex or pair count");
  if (m > 0 && (!src || !dst)) return fail(TC_EINVAL, "null pair arrays");
  int T = threads <= 1 ? 1 : threads;
  // Per-thread histograms cost T*n*4 bytes; shrink T for huge n.
  const size_t kHistCap = (size_t)6 << 30;
  while (T > 1 && (size_t)T * (size_t)(n > 0 ? n : 1) * 4 > kHistCap) T /= 2;

  // Range check of the endpoints.
  int64_t bad = 0;
#pragma omp parallel for num_threads(T) reduction(+ : bad) schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    int32_t a = src[i], b = dst[i];
    if (a >= n || b >= n) ++bad;
  }
  if (bad) return fail(TC_EINVAL, "pair endpoint out of range [0, n)");

  const size_t nn = (size_t)(n > 0 ? n : 1);
  std::unique_ptr<int32_t[]> hist(new int32_t[(size_t)T * nn]);
  auto chunk_lo = [&](int t) { return m * t / T; };

  // Pass 1: per-virtual-thread out-degree histograms (both directions).
#pragma omp parallel for num_threads(T) schedule(static, 1)
  for (int t = 0; t < T; ++t) {
    int32_t* h = hist.get() + (size_t)t * nn;
    std::memset(h, 0, nn * sizeof(int32_t));
    for (int64_t i = chunk_lo(t); i < chunk_lo(t + 1); ++i) {
      int32_t a = src[i], b = dst[i];
      if (a < 0 || b < 0 || a == b) continue;
      ++h[a];
      ++h[b];
    }
  }
  // Row sizes and per-thread relative write cursors.
  std::unique_ptr<int64_t[]> deg(new int64_t[nn]);
#pragma omp parallel for num_threads(T) schedule(static, 4096)
  for (int64_t v = 0; v < n; ++v) {
    int64_t s = 0;
    for (int t = 0; t < T; ++t) {
      int32_t c = hist[(size_t)t * nn + v];
      hist[(size_t)t * nn + v] = (int32_t)s;
      s += c;
    }
    deg[v] = s;
  }
  std::unique_ptr<int64_t[]> raw_ptr(new int64_t[nn + 1]);
  exclusive_scan(deg.get(), n, raw_ptr.get());
  const int64_t raw_nnz = raw_ptr[n];
  std::unique_ptr<int32_t[]> raw(new int32_t[raw_nnz > 0 ? raw_nnz : 1]);

  // Pass 2: scatter into rows at deterministic positions.
#pragma omp parallel for num_threads(T) schedule(static, 1)
  for (int t = 0; t < T; ++t) {
    int32_t* h = hist.get() + (size_t)t * nn;
    for (int64_t i = chunk_lo(t); i < chunk_lo(t + 1); ++i) {
      int32_t a = src[i], b = dst[i];
      if (a < 0 || b < 0 || a == b) continue;
      raw[raw_ptr[a] + h[a]++] = b;
      raw[raw_ptr[b] + h[b]++] = a;
    }
  }
  hist.reset();

  // Sort and deduplicate every row.
#pragma omp parallel for num_threads(T) schedule(dynamic, 256)
  for (int64_t v = 0; v < n; ++v) {
    int32_t* b = raw.get() + raw_ptr[v];
    int32_t* e = raw.get() + raw_ptr[v + 1];
    std::sort(b, e);
    deg[v] = std::unique(b, e) - b;
  }
  out.allocate(n, 0);
  exclusive_scan(deg.get(), n, out.rowptr.get());
  out.nnz = out.rowptr[n];
  out.colidx.reset(new int32_t[out.nnz > 0 ? out.nnz : 1]);
#pragma omp parallel for num_threads(T) schedule(dynamic, 1024)
  for (int64_t v = 0; v < n; ++v)
    std::memcpy(out.colidx.get() + out.rowptr[v], raw.get() + raw_ptr[v],
                sizeof(int32_t) * (size_t)deg[v]);
  out.compute_max_degree();
  return TC_OK;
}

// Orientation rule of graph.cc:265-266 / 284-285, with undirected degrees.
static inline bool keep_edge(int64_t dsrc, int64_t ddst, int32_t src, int32_t dst) {
  return ddst > dsrc || (ddst == dsrc && dst > src);
}

tc_status orient(const tc_csr_view& g, int threads, Csr& dag) {
  const int T = threads <= 1 ? 1 : threads;
  const int32_t n = g.n_rows;
  std::unique_ptr<int64_t[]> newdeg(new int64_t[(size_t)n + 1]);
#pragma omp parallel for num_threads(T) schedule(dynamic, 1024)
  for (int64_t s = 0; s < n; ++s) {
    const int64_t ds = g.rowptr[s + 1] - g.rowptr[s];
    int64_t c = 0;
    for (int64_t e = g.rowptr[s]; e < g.rowptr[s + 1]; ++e) {
      const int32_t d = g.colidx[e];
      if (keep_edge(ds, g.rowptr[d + 1] - g.rowptr[d], (int32_t)s, d)) ++c;
    }
    newdeg[s] = c;
  }
  dag.allocate(n, 0);
  exclusive_scan(newdeg.get(), n, dag.rowptr.get());
  dag.nnz = dag.rowptr[n];
  dag.colidx.reset(new int32_t[dag.nnz > 0 ? dag.nnz : 1]);
#pragma omp parallel for num_threads(T) schedule(dynamic, 1024)
  for (int64_t s = 0; s < n; ++s) {
    const int64_t ds = g.rowptr[s + 1] - g.rowptr[s];
    int64_t o = dag.rowptr[s];
    for (int64_t e = g.rowptr[s]; e < g.rowptr[s + 1]; ++e) {
      const int32_t d = g.colidx[e];
      if (keep_edge(ds, g.rowptr[d + 1] - g.rowptr[d], (int32_t)s, d)) dag.colidx[o++] = d;
    }
  }
  dag.compute_max_degree();
  return TC_OK;
}

static tc_status validate(const tc_csr_view& g, uint32_t check) {
  if (g.n_rows < 0 || g.nnz < 0) return fail(TC_EINVAL, "negative n_rows or nnz");
  if (!g.rowptr) return fail(TC_EINVAL, "null rowptr");
  if (g.nnz > 0 && !g.colidx) return fail(TC_EINVAL, "null colidx");
  const int32_t n = g.n_rows;
  if (check & TC_CHECK_SHAPE) {
    if (g.rowptr[0] != 0) return fail(TC_EINVAL, "rowptr[0] != 0");
    if (g.rowptr[n] != g.nnz) return fail(TC_EINVAL, "rowptr[n] != nnz");
    int64_t bad_mono = 0;
#pragma omp parallel for reduction(+ : bad_mono) schedule(static)
    for (int64_t v = 0; v < n; ++v)
      if (g.rowptr[v + 1] < g.rowptr[v]) ++bad_mono;
    if (bad_mono) return fail(TC_EINVAL, "rowptr is not monotone");
    int64_t bad_id = 0;
#pragma omp parallel for reduction(+ : bad_id) schedule(static)
    for (int64_t e = 0; e < g.nnz; ++e)
      if (g.colidx[e] < 0 || g.colidx[e] >= n) ++bad_id;
    if (bad_id) return fail(TC_EINVAL, "neighbour id out of range [0, n)");
  }
  if (check & (TC_CHECK_SORTED | TC_CHECK_NOSELF)) {
    int64_t unsorted = 0, self = 0;
#pragma omp parallel for reduction(+ : unsorted, self) schedule(dynamic, 1024)
    for (int64_t v = 0; v < n; ++v) {
      for (int64_t e = g.rowptr[v]; e < g.rowptr[v + 1]; ++e) {
        if (e > g.rowptr[v] && g.colidx[e] <= g.colidx[e - 1]) ++unsorted;
        if (g.colidx[e] == v) ++self;
      }
    }
    if ((check & TC_CHECK_SORTED) && unsorted)
      return fail(TC_EINVAL, "a row is not strictly increasing (unsorted or duplicate ids)");
    if ((check & TC_CHECK_NOSELF) && self) return fail(TC_EINVAL, "self-loop present");
  }
  return TC_OK;
}

}  // namespace tch

using namespace tch;

extern "C" {

const char* tch_last_error(void) { return tch::g_last_error.c_str(); }

void tch_csr_free(tch_csr* g) { delete g; }

tc_status tch_csr_view(const tch_csr* g, tc_csr_view* out) try {
  if (!g || !out) return fail(TC_EINVAL, "null argument");
  *out = g->g.view();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

int32_t tch_csr_max_degree(const tch_csr* g) { return g ? g->g.max_degree : -1; }

tc_status tch_validate(const tc_csr_view* g, uint32_t check) try {
  if (!g) return fail(TC_EINVAL, "null view");
  return tch::validate(*g, check);
} catch (...) {
  return tch::caught();
}

int64_t tch_intersect_num(const tc_csr_view* g, int32_t v, int32_t u, const uint8_t* vlabels, uint8_t label) {
  if (!g || !g->rowptr || v < 0 || u < 0 || v >= g->n_rows || u >= g->n_rows) {
    tch::fail(TC_EINVAL, "intersect_num: vertex out of range");
    return -1;
  }
  const int32_t* a = g->colidx + g->rowptr[v];
  const int32_t* b = g->colidx + g->rowptr[u];
  const int64_t na = g->rowptr[v + 1] - g->rowptr[v], nb = g->rowptr[u + 1] - g->rowptr[u];
  int64_t i = 0, j = 0, num = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) {
      ++i;
    } else if (b[j] < a[i]) {
      ++j;
    } else {
      if (!vlabels || vlabels[a[i]] == label) ++num;
      ++i;
      ++j;
    }
  }
  return num;
}

tc_status tch_csr_copy(const tc_csr_view* src, uint32_t check, tch_csr** out) try {
  if (!src || !out) return fail(TC_EINVAL, "null argument");
  tc_status st = tch::validate(*src, check | TC_CHECK_SHAPE);
  if (st != TC_OK) return st;
  std::unique_ptr<tch_csr> h(new tch_csr);
  h->g.allocate(src->n_rows, src->nnz);
  std::memcpy(h->g.rowptr.get(), src->rowptr, sizeof(int64_t) * ((size_t)src->n_rows + 1));
  if (src->nnz) std::memcpy(h->g.colidx.get(), src->colidx, sizeof(int32_t) * (size_t)src->nnz);
  h->g.compute_max_degree();
  *out = h.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_csr_from_pairs(int32_t n, const int32_t* src, const int32_t* dst, int64_t m,
                             tch_csr** out) try {
  if (!out) return fail(TC_EINVAL, "null out");
  std::unique_ptr<tch_csr> h(new tch_csr);
  tc_status st = build_symmetric(n, src, dst, m, omp_get_max_threads(), h->g);
  if (st != TC_OK) return st;
  *out = h.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_orient(const tc_csr_view* sym, tch_csr** dag_out) try {
  if (!sym || !dag_out) return fail(TC_EINVAL, "null argument");
  tc_status st = tch::validate(*sym, TC_CHECK_SHAPE);
  if (st != TC_OK) return st;
  std::unique_ptr<tch_csr> h(new tch_csr);
  st = orient(*sym, omp_get_max_threads(), h->g);
  if (st != TC_OK) return st;
  *dag_out = h.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_sort_neighbors(tch_csr* g) try {
  if (!g) return fail(TC_EINVAL, "null graph");
  Csr& c = g->g;
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < c.n; ++v)
    std::sort(c.colidx.get() + c.rowptr[v], c.colidx.get() + c.rowptr[v + 1]);
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_read_graphminer(const char* prefix, int use_dag, tch_csr** out) try {
  if (!prefix || !out) return fail(TC_EINVAL, "null argument");
  std::string p(prefix);
  std::ifstream meta(p + ".meta.txt");
  if (!meta) return fail(TC_EIO, "cannot open " + p + ".meta.txt");
  int64_t n = 0, m = 0;
  int vid_size = 0, eid_size = 0, vlabel_size = 0, elabel_size = 0;
  int64_t max_degree = 0, feat_len = 0, nvc = 0, nec = 0;
  meta >> n >> m >> vid_size >> eid_size >> vlabel_size >> elabel_size >> max_degree >> feat_len >>
      nvc >> nec;
  if (!meta) return fail(TC_EINVAL, "malformed meta file " + p + ".meta.txt");
  if (vid_size != 4) return fail(TC_EINVAL, "vid_size must be 4 (int32 ids)");
  if (eid_size != 8) return fail(TC_EINVAL, "eid_size must be 8 (int64 offsets)");
  if (n <= 0 || n > INT32_MAX || m < 0) return fail(TC_ERANGE, "vertex or edge count out of range");
  if (!(max_degree > 0 && max_degree < n))
    return fail(TC_EINVAL, "meta max_degree must satisfy 0 < max_degree < n");
  // the files must have the sizes the meta data gives before anything is
  // allocated from those counts
  if (file_bytes(p + ".vertex.bin") != 8 * (n + 1))
    return fail(TC_EIO, "cannot read " + p + ".vertex.bin (size must be 8*(n+1))");
  if (m > INT64_MAX / 4 || file_bytes(p + ".edge.bin") != 4 * m)
    return fail(TC_EIO, "cannot read " + p + ".edge.bin (size must be 4*m)");
  std::unique_ptr<tch_csr> h(new tch_csr);
  h->g.allocate((int32_t)n, m);
  if (!read_exact(p + ".vertex.bin", h->g.rowptr.get(), sizeof(int64_t) * ((size_t)n + 1)))
    return fail(TC_EIO, "cannot read " + p + ".vertex.bin (size must be 8*(n+1))");
  if (!read_exact(p + ".edge.bin", h->g.colidx.get(), sizeof(int32_t) * (size_t)m))
    return fail(TC_EIO, "cannot read " + p + ".edge.bin (size must be 4*m)");
  tc_status st = tch::validate(h->g.view(), TC_CHECK_SHAPE);
  if (st != TC_OK) return st;
  h->g.compute_max_degree();
  if (use_dag) {
    std::unique_ptr<tch_csr> d(new tch_csr);
    st = orient(h->g.view(), omp_get_max_threads(), d->g);
    if (st != TC_OK) return st;
    h.swap(d);
  }
  *out = h.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_write_graphminer(const char* prefix, const tc_csr_view* g) try {
  if (!prefix || !g) return fail(TC_EINVAL, "null argument");
  tc_status st = tch::validate(*g, TC_CHECK_SHAPE);
  if (st != TC_OK) return st;
  std::string p(prefix);
  int64_t md = 0;
  for (int64_t v = 0; v < g->n_rows; ++v) md = std::max<int64_t>(md, g->rowptr[v + 1] - g->rowptr[v]);
  std::ofstream meta(p + ".meta.txt");
  if (!meta) return fail(TC_EIO, "cannot write " + p + ".meta.txt");
  // n m vid_size eid_size vlabel_size elabel_size max_degree feat_len nvc nec
  meta << g->n_rows << "\n" << g->nnz << "\n4 8 1 2\n" << md << "\n0\n0\n0\n";
  meta.close();
  if (!meta) return fail(TC_EIO, "cannot write " + p + ".meta.txt");
  if (!write_exact(p + ".vertex.bin", g->rowptr, sizeof(int64_t) * ((size_t)g->n_rows + 1)))
    return fail(TC_EIO, "cannot write " + p + ".vertex.bin");
  if (!write_exact(p + ".edge.bin", g->colidx, sizeof(int32_t) * (size_t)g->nnz))
    return fail(TC_EIO, "cannot write " + p + ".edge.bin");
  return TC_OK;
} catch (...) {
  return tch::caught();
}

}  // extern "C"
