// The reference's per-PE METIS part files, read and written, and the CSR a
// PE counts from them.
//
// Reference behaviour followed here:
//   file set, one per PE i   successful_push_pull/graph_partition.cc:501-566
//     <prefix>{i}.meta.txt        n, n_real, E, max_subg_nv, max_subg_ne, max_degree
//     <prefix>{i}.vertex.bin      int64 rowptr[n+1] over all global ids; only the
//                                 PE's own rows are non-empty (:467-487)
//     <prefix>{i}.edge.bin        int32 colidx[E], global ids (:489-497)
//     <prefix>{i}.vertex_list.bin int32 owned ids [n_real], ascending (:461-464)
//     <prefix>{i}.part.bin        int64 (METIS idx_t) owner PE of every id [n]
//   reader                   successful_push_pull/graph.cc:16-23,33-34,56-59
//   per-PE pivots            successful_push_pull/bs_warp_vertex_nvshmem.cuh:25-28
//                            (u = vertex_list[i]), remote N+(v) pulled from
//                            PE part[v] (:110-136)
// The .txt dumps the reference writes beside the binaries are debug output
// that its loader never reads; they are not written.  METIS itself is absent:
// any part array (a range split, a balanced split, an external METIS run)
// produces the same file set.
#include <algorithm>
#include <fstream>

#include "host_common.hpp"

struct tch_metis {
  int64_t n = 0, n_real = 0, nnz = 0, max_subg_nv = 0, max_subg_ne = 0, max_degree = 0;
  std::unique_ptr<int64_t[]> rowptr;
  std::unique_ptr<int32_t[]> colidx;
  std::unique_ptr<int32_t[]> vertex_list;
  std::unique_ptr<int64_t[]> part;
};

namespace {

using tch::fail;

std::string file_of(const char* prefix, int32_t i, const char* ext) {
  return std::string(prefix) + std::to_string(i) + ext;
}

}  // namespace

extern "C" {

tc_status tch_metis_write(const char* prefix, const tc_csr_view* g, const int64_t* part, int32_t k) try {
  if (!prefix || !g || !part) return fail(TC_EINVAL, "null argument");
  if (k < 1) return fail(TC_EINVAL, "need k >= 1");
  tc_status st = tch_validate(g, TC_CHECK_SHAPE);
  if (st != TC_OK) return st;
  const int64_t n = g->n_rows;
  std::vector<int64_t> nreal((size_t)k, 0), ne((size_t)k, 0);
  int64_t md = 0;
  for (int64_t v = 0; v < n; ++v) {
    if (part[v] < 0 || part[v] >= k) return fail(TC_EINVAL, "part values must be in [0, k)");
    const int64_t d = g->rowptr[v + 1] - g->rowptr[v];
    nreal[(size_t)part[v]] += 1;
    ne[(size_t)part[v]] += d;
    md = std::max(md, d);
  }
  const int64_t max_nv = *std::max_element(nreal.begin(), nreal.end());
  const int64_t max_ne = *std::max_element(ne.begin(), ne.end());
  std::unique_ptr<int64_t[]> rp(new int64_t[(size_t)n + 1]);
  for (int32_t i = 0; i < k; ++i) {
    std::vector<int32_t> owned;
    owned.reserve((size_t)nreal[(size_t)i]);
    std::vector<int32_t> ci;
    ci.reserve((size_t)ne[(size_t)i]);
    rp[0] = 0;
    for (int64_t v = 0; v < n; ++v) {
      if (part[v] == i) {
        owned.push_back((int32_t)v);
        ci.insert(ci.end(), g->colidx + g->rowptr[v], g->colidx + g->rowptr[v + 1]);
      }
      rp[v + 1] = (int64_t)ci.size();
    }
    std::ofstream meta(file_of(prefix, i, ".meta.txt"));
    if (!meta) return fail(TC_EIO, "cannot write " + file_of(prefix, i, ".meta.txt"));
    meta << n << "\n" << nreal[(size_t)i] << "\n" << ne[(size_t)i] << "\n" << max_nv << "\n" << max_ne << "\n"
         << md << "\n";
    meta.close();
    if (!meta) return fail(TC_EIO, "cannot write " + file_of(prefix, i, ".meta.txt"));
    if (!tch::write_exact(file_of(prefix, i, ".vertex.bin"), rp.get(), sizeof(int64_t) * ((size_t)n + 1)) ||
        !tch::write_exact(file_of(prefix, i, ".edge.bin"), ci.data(), sizeof(int32_t) * ci.size()) ||
        !tch::write_exact(file_of(prefix, i, ".vertex_list.bin"), owned.data(), sizeof(int32_t) * owned.size()) ||
        !tch::write_exact(file_of(prefix, i, ".part.bin"), part, sizeof(int64_t) * (size_t)n))
      return fail(TC_EIO, "cannot write the binary files of " + file_of(prefix, i, ""));
  }
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_metis_read(const char* prefix, int32_t i, tch_metis** out) try {
  if (!prefix || !out) return fail(TC_EINVAL, "null argument");
  if (i < 0) return fail(TC_EINVAL, "PE index must be >= 0");
  std::unique_ptr<tch_metis> m(new tch_metis);
  {
    std::ifstream meta(file_of(prefix, i, ".meta.txt"));
    if (!meta) return fail(TC_EIO, "cannot open " + file_of(prefix, i, ".meta.txt"));
    meta >> m->n >> m->n_real >> m->nnz >> m->max_subg_nv >> m->max_subg_ne >> m->max_degree;
    if (!meta) return fail(TC_EINVAL, "malformed meta file " + file_of(prefix, i, ".meta.txt"));
  }
  const int64_t n = m->n;
  if (n <= 0 || n > INT32_MAX || m->n_real < 0 || m->n_real > n || m->nnz < 0 || m->nnz > INT64_MAX / 4)
    return fail(TC_ERANGE, "meta counts out of range in " + file_of(prefix, i, ".meta.txt"));
  // the files must have the sizes the meta data gives before anything is
  // allocated from those counts
  const struct {
    const char* ext;
    int64_t bytes;
    const char* rule;
  } expect[] = {{".vertex.bin", 8 * (n + 1), "8*(n+1)"},
                {".edge.bin", 4 * m->nnz, "4*E"},
                {".vertex_list.bin", 4 * m->n_real, "4*n_real"},
                {".part.bin", 8 * n, "8*n"}};
  for (const auto& f : expect)
    if (tch::file_bytes(file_of(prefix, i, f.ext)) != f.bytes)
      return fail(TC_EIO, "cannot read " + file_of(prefix, i, f.ext) + " (size must be " + f.rule + ")");
  m->rowptr.reset(new int64_t[(size_t)n + 1]);
  m->colidx.reset(new int32_t[m->nnz > 0 ? (size_t)m->nnz : 1]);
  m->vertex_list.reset(new int32_t[m->n_real > 0 ? (size_t)m->n_real : 1]);
  m->part.reset(new int64_t[(size_t)n]);
  if (!tch::read_exact(file_of(prefix, i, ".vertex.bin"), m->rowptr.get(), sizeof(int64_t) * ((size_t)n + 1)))
    return fail(TC_EIO, "cannot read " + file_of(prefix, i, ".vertex.bin") + " (size must be 8*(n+1))");
  if (!tch::read_exact(file_of(prefix, i, ".edge.bin"), m->colidx.get(), sizeof(int32_t) * (size_t)m->nnz))
    return fail(TC_EIO, "cannot read " + file_of(prefix, i, ".edge.bin") + " (size must be 4*E)");
  if (!tch::read_exact(file_of(prefix, i, ".vertex_list.bin"), m->vertex_list.get(),
                sizeof(int32_t) * (size_t)m->n_real))
    return fail(TC_EIO, "cannot read " + file_of(prefix, i, ".vertex_list.bin") + " (size must be 4*n_real)");
  if (!tch::read_exact(file_of(prefix, i, ".part.bin"), m->part.get(), sizeof(int64_t) * (size_t)n))
    return fail(TC_EIO, "cannot read " + file_of(prefix, i, ".part.bin") + " (size must be 8*n)");
  const tc_csr_view v{m->rowptr.get(), m->colidx.get(), (int32_t)n, m->nnz};
  tc_status st = tch_validate(&v, TC_CHECK_SHAPE);
  if (st != TC_OK) return st;
  // owned ids ascending and in range, owned by this PE; every other row empty
  int64_t owned_edges = 0;
  for (int64_t x = 0; x < m->n_real; ++x) {
    const int32_t u = m->vertex_list[x];
    if (u < 0 || u >= n || (x > 0 && m->vertex_list[x - 1] >= u) || m->part[u] != i)
      return fail(TC_EINVAL, "vertex_list must hold this PE's ids, ascending");
    owned_edges += m->rowptr[u + 1] - m->rowptr[u];
  }
  if (owned_edges != m->nnz) return fail(TC_EINVAL, "rows outside vertex_list must be empty");
  *out = m.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_metis_get(const tch_metis* m, tch_metis_view* o) try {
  if (!m || !o) return fail(TC_EINVAL, "null argument");
  o->n = (int32_t)m->n;
  o->n_real = (int32_t)m->n_real;
  o->nnz = m->nnz;
  o->max_subg_nv = m->max_subg_nv;
  o->max_subg_ne = m->max_subg_ne;
  o->max_degree = m->max_degree;
  o->rowptr = m->rowptr.get();
  o->colidx = m->colidx.get();
  o->vertex_list = m->vertex_list.get();
  o->part = m->part.get();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

void tch_metis_free(tch_metis* m) { delete m; }

tc_status tch_metis_assemble(const tch_metis* const* pes, int32_t k, int32_t i, tch_csr** out,
                             uint8_t* owned) try {
  if (!pes || !out) return fail(TC_EINVAL, "null argument");
  if (k < 1 || i < -1 || i >= k) return fail(TC_EINVAL, "need k >= 1 and -1 <= i < k");
  for (int32_t p = 0; p < k; ++p)
    if (!pes[p]) return fail(TC_EINVAL, "every PE's files are needed");
  const int64_t n = pes[0]->n;
  for (int32_t p = 1; p < k; ++p)
    if (pes[p]->n != n) return fail(TC_EINVAL, "PE files disagree on n");
  const int64_t* part = pes[i < 0 ? 0 : i]->part.get();
  for (int64_t v = 0; v < n; ++v)
    if (part[v] < 0 || part[v] >= k) return fail(TC_EINVAL, "part values must be in [0, k)");
  // rows: all of them (i = -1), or PE i's own rows plus every remote list
  // they reach, each from its owner's file
  std::vector<uint8_t> keep((size_t)n, i < 0 ? 1 : 0);
  if (i >= 0) {
    const tch_metis& me = *pes[i];
    for (int64_t x = 0; x < me.n_real; ++x) {
      const int32_t u = me.vertex_list[x];
      keep[(size_t)u] = 1;
      for (int64_t e = me.rowptr[u]; e < me.rowptr[u + 1]; ++e) {
        const int32_t v = me.colidx[e];
        if (v < 0 || v >= n) return fail(TC_EINVAL, "colidx out of range");
        keep[(size_t)v] = 1;
      }
    }
  }
  std::unique_ptr<tch_csr> c(new tch_csr);
  std::unique_ptr<int64_t[]> len(new int64_t[(size_t)n]);
  for (int64_t v = 0; v < n; ++v) {
    const tch_metis& src = *pes[part[v]];
    len[v] = keep[(size_t)v] ? src.rowptr[v + 1] - src.rowptr[v] : 0;
  }
  c->g.allocate((int32_t)n, 0);
  tch::exclusive_scan(len.get(), n, c->g.rowptr.get());
  c->g.nnz = c->g.rowptr[n];
  c->g.colidx.reset(new int32_t[c->g.nnz > 0 ? (size_t)c->g.nnz : 1]);
  for (int64_t v = 0; v < n; ++v)
    if (len[v])
      std::memcpy(c->g.colidx.get() + c->g.rowptr[v], pes[part[v]]->colidx.get() + pes[part[v]]->rowptr[v],
                  sizeof(int32_t) * (size_t)len[v]);
  c->g.compute_max_degree();
  if (owned)
    for (int64_t v = 0; v < n; ++v) owned[v] = (i < 0 || part[v] == i) ? 1 : 0;
  *out = c.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

}  // extern "C"
