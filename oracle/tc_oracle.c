/*
 * tc_oracle.c -- CPU ORACLE for the triangle-counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C with OpenMP, the
 * reference's CPU algorithm for the path so that tests, the smoke check and
 * the bench's cpu_baseline leg can check the HIP path against it.  No product
 * code links, loads or calls it.
 *
 * Reference behaviour restated (paths relative to the reference repo):
 *   oracle_binary_search      successful_push_pull/graph.cc:363-374
 *   oracle_intersect_merge    successful_push_pull/graph.cc:376-391 (label test dropped)
 *   oracle_intersect_bsearch  successful_push_pull/graph_gpu.h:228-252 (shorter list
 *                             searched, element by element, in the longer one)
 *   oracle_orient_*           successful_push_pull/graph.cc:252-298
 *   oracle_tc_count           bs_warp_vertex.cuh:13-22 loop nest, run as the GraphMiner
 *                             tc_omp_base OpenMP loop (successful_push_pull/Makefile:5-7,
 *                             absent from the tree): parallel for schedule(dynamic, chunk)
 *                             over pivots u with a u64 reduction
 *   oracle_tc_count_halo      successful_push_pull/bs_warp_vertex_nvshmem.cuh:27-40 (owned
 *                             pivots) with every remote list read from a halo cache, as the
 *                             pull branch does (:110-136)
 *   oracle_tc_pushpull_pe     the per-PE attribution of the push/pull kernel:
 *                             local edges (:36-38), pulled edges with d+(u) >= d+(v)
 *                             (:127-136), pushed-in edges with d+(u) < d+(v) (:228-232)
 *   oracle_tc_pushpull_bounds the same attribution under any split points
 *   oracle_tc_edge_counts     per DAG edge (u, v), the term bs_warp_vertex.cuh:20 adds:
 *                             |N+(u) ∩ N+(v)| by the graph.cc:376-391 merge
 *
 * Pinning: see the header of oracle.py and the tests under tests/ -- closed
 * forms, brute force, networkx/scipy golden fixtures and invariants.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* graph.cc:363-374: l/r inclusive bounds, mid = l + (r-l)/2. */
int oracle_binary_search(const int32_t* list, int64_t begin, int64_t end, int32_t key) {
  int64_t l = begin, r = end - 1;
  while (r >= l) {
    int64_t mid = l + (r - l) / 2;
    int32_t value = list[mid];
    if (value == key) return 1;
    if (value < key) l = mid + 1;
    else r = mid - 1;
  }
  return 0;
}

/* graph.cc:376-391: two-pointer merge count of |a ∩ b| for sorted lists. */
int64_t oracle_intersect_merge(const int32_t* a, int64_t na, const int32_t* b, int64_t nb) {
  int64_t num = 0, i = 0, j = 0;
  while (i < na && j < nb) {
    int32_t x = a[i], y = b[j];
    if (x <= y) i++;
    if (y <= x) j++;
    if (x == y) num++;
  }
  return num;
}

/* graph_gpu.h:228-252: lookup = shorter list, search = longer list. */
int64_t oracle_intersect_bsearch(const int32_t* a, int64_t na, const int32_t* b, int64_t nb) {
  if (na == 0 || nb == 0) return 0;
  const int32_t* lookup = a; int64_t nl = na;
  const int32_t* search = b; int64_t ns = nb;
  if (na > nb) { lookup = b; nl = nb; search = a; ns = na; }
  int64_t c = 0;
  for (int64_t i = 0; i < nl; i++) c += oracle_binary_search(search, 0, ns, lookup[i]);
  return c;
}

static int keep_edge(int64_t dsrc, int64_t ddst, int32_t src, int32_t dst) {
  return ddst > dsrc || (ddst == dsrc && dst > src);
}

/* graph.cc:252-298, pass 1: DAG out-degrees -> rowptr (exclusive scan). */
int64_t oracle_orient_rowptr(int32_t n, const int64_t* rowptr, const int32_t* colidx,
                             int64_t* out_rowptr) {
  int64_t* nd = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  if (!nd) return -1;
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t s = 0; s < n; s++) {
    int64_t ds = rowptr[s + 1] - rowptr[s], c = 0;
    for (int64_t e = rowptr[s]; e < rowptr[s + 1]; e++) {
      int32_t d = colidx[e];
      if (keep_edge(ds, rowptr[d + 1] - rowptr[d], (int32_t)s, d)) c++;
    }
    nd[s] = c;
  }
  int64_t acc = 0;
  for (int64_t s = 0; s < n; s++) { out_rowptr[s] = acc; acc += nd[s]; }
  out_rowptr[n] = acc;
  free(nd);
  return acc;
}

/* graph.cc:252-298, pass 2: fill the kept neighbours in original order. */
void oracle_orient_fill(int32_t n, const int64_t* rowptr, const int32_t* colidx,
                        const int64_t* out_rowptr, int32_t* out_colidx) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t s = 0; s < n; s++) {
    int64_t ds = rowptr[s + 1] - rowptr[s], o = out_rowptr[s];
    for (int64_t e = rowptr[s]; e < rowptr[s + 1]; e++) {
      int32_t d = colidx[e];
      if (keep_edge(ds, rowptr[d + 1] - rowptr[d], (int32_t)s, d)) out_colidx[o++] = d;
    }
  }
}

/* bs_warp_vertex.cuh:13-22: total += |N+(u) ∩ N+(v)| for v in N+(u), u in
 * [row_begin, row_end).  method 0 = merge (graph.cc:376), 1 = binary search
 * of the shorter list in the longer (graph_gpu.h:228). */
uint64_t oracle_tc_count(int32_t n, const int64_t* rowptr, const int32_t* colidx,
                         int32_t row_begin, int32_t row_end, int chunk, int threads, int method) {
  (void)n;
  if (chunk < 1) chunk = 1024;
  if (threads < 1) threads = omp_get_max_threads();
  uint64_t total = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, chunk) reduction(+ : total)
  for (int64_t u = row_begin; u < row_end; u++) {
    const int32_t* a = colidx + rowptr[u];
    int64_t da = rowptr[u + 1] - rowptr[u];
    for (int64_t j = 0; j < da; j++) {
      int32_t v = a[j];
      const int32_t* b = colidx + rowptr[v];
      int64_t db = rowptr[v + 1] - rowptr[v];
      total += (uint64_t)(method ? oracle_intersect_bsearch(a, da, b, db)
                                 : oracle_intersect_merge(a, da, b, db));
    }
  }
  return total;
}

/* Count over an explicit list of pivot rows (a bounded sample of a graph). */
uint64_t oracle_tc_count_list(const int64_t* rowptr, const int32_t* colidx, const int32_t* rows,
                              int64_t nrows, int chunk, int threads) {
  if (chunk < 1) chunk = 1024;
  if (threads < 1) threads = omp_get_max_threads();
  uint64_t total = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, chunk) reduction(+ : total)
  for (int64_t i = 0; i < nrows; i++) {
    int32_t u = rows[i];
    const int32_t* a = colidx + rowptr[u];
    int64_t da = rowptr[u + 1] - rowptr[u];
    for (int64_t j = 0; j < da; j++) {
      int32_t v = a[j];
      total += (uint64_t)oracle_intersect_merge(a, da, colidx + rowptr[v], rowptr[v + 1] - rowptr[v]);
    }
  }
  return total;
}

/* Partition with halo cache: pivots are the owned rows [row_begin, row_end)
 * of local_rowptr/local_colidx; a neighbour v inside the range reads its list
 * from the local rows, any other v from the halo (sorted halo_ids). */
static int64_t halo_find(const int32_t* ids, int64_t n, int32_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (ids[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && ids[lo] == v) ? lo : -1;
}

uint64_t oracle_tc_count_halo(int32_t row_begin, int32_t row_end, const int64_t* local_rowptr,
                              const int32_t* local_colidx, int32_t n_halo, const int32_t* halo_ids,
                              const int64_t* halo_rowptr, const int32_t* halo_colidx,
                              int threads, int64_t* missing_out) {
  if (threads < 1) threads = omp_get_max_threads();
  uint64_t total = 0;
  int64_t missing = 0;
  const int64_t nl = (int64_t)row_end - row_begin;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1024) reduction(+ : total, missing)
  for (int64_t i = 0; i < nl; i++) {
    const int32_t* a = local_colidx + local_rowptr[i];
    int64_t da = local_rowptr[i + 1] - local_rowptr[i];
    for (int64_t j = 0; j < da; j++) {
      int32_t v = a[j];
      const int32_t* b;
      int64_t db;
      if (v >= row_begin && v < row_end) {
        b = local_colidx + local_rowptr[v - row_begin];
        db = local_rowptr[v - row_begin + 1] - local_rowptr[v - row_begin];
      } else {
        int64_t h = halo_find(halo_ids, n_halo, v);
        if (h < 0) { missing++; continue; }
        b = halo_colidx + halo_rowptr[h];
        db = halo_rowptr[h + 1] - halo_rowptr[h];
      }
      total += (uint64_t)oracle_intersect_merge(a, da, b, db);
    }
  }
  if (missing_out) *missing_out = missing;
  return total;
}

/* Triangles per vertex (every triangle credited to its three corners, the
 * quantity networkx.triangles reports), from the DAG. */
void oracle_tc_per_vertex(int32_t n, const int64_t* rowptr, const int32_t* colidx, uint64_t* tri) {
  memset(tri, 0, sizeof(uint64_t) * (size_t)n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t u = 0; u < n; u++) {
    const int32_t* a = colidx + rowptr[u];
    int64_t da = rowptr[u + 1] - rowptr[u];
    for (int64_t j = 0; j < da; j++) {
      int32_t v = a[j];
      const int32_t* b = colidx + rowptr[v];
      int64_t db = rowptr[v + 1] - rowptr[v], x = 0, y = 0;
      while (x < da && y < db) {
        if (a[x] < b[y]) x++;
        else if (b[y] < a[x]) y++;
        else {
          int32_t w = a[x];
#pragma omp atomic
          tri[u]++;
#pragma omp atomic
          tri[v]++;
#pragma omp atomic
          tri[w]++;
          x++; y++;
        }
      }
    }
  }
}

/* Per-PE count of the reference push/pull kernel for a contiguous k-way
 * split, PE p: owned pivots count local edges and pulled edges
 * (d+(u) >= d+(v)); edges with d+(u) < d+(v) are credited to v's owner. */
uint64_t oracle_tc_pushpull_pe(int32_t n, const int64_t* rowptr, const int32_t* colidx,
                               int32_t k, int32_t pe) {
  int64_t chunk = n / k + (n % k != 0);
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : total)
  for (int64_t u = 0; u < n; u++) {
    const int32_t* a = colidx + rowptr[u];
    int64_t da = rowptr[u + 1] - rowptr[u];
    int32_t pu = (int32_t)(u / chunk);
    for (int64_t j = 0; j < da; j++) {
      int32_t v = a[j];
      int32_t pv = (int32_t)(v / chunk);
      int64_t db = rowptr[v + 1] - rowptr[v];
      int32_t owner;
      if (pv == pu) owner = pu;
      else owner = (da < db) ? pv : pu;
      if (owner != pe) continue;
      total += (uint64_t)oracle_intersect_merge(a, da, colidx + rowptr[v], db);
    }
  }
  return total;
}

/* Per DAG edge e = (u, v), in CSR order: |N+(u) ∩ N+(v)| (the term the
 * loop nest of bs_warp_vertex.cuh:13-22 adds for that edge). */
void oracle_tc_edge_counts(int32_t n, const int64_t* rowptr, const int32_t* colidx, uint32_t* out) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t u = 0; u < n; u++) {
    const int32_t* a = colidx + rowptr[u];
    int64_t da = rowptr[u + 1] - rowptr[u];
    for (int64_t j = 0; j < da; j++) {
      int32_t v = a[j];
      out[rowptr[u] + j] = (uint32_t)oracle_intersect_merge(a, da, colidx + rowptr[v],
                                                            rowptr[v + 1] - rowptr[v]);
    }
  }
}

/* Per-PE sums of the push/pull kernel (rule of oracle_tc_pushpull_pe) when
 * PE p owns ids [bounds[p], bounds[p+1]). */
void oracle_tc_pushpull_bounds(int32_t n, const int64_t* rowptr, const int32_t* colidx, int32_t k,
                               const int32_t* bounds, uint64_t* out) {
  int32_t* part = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  for (int32_t p = 0; p < k; p++)
    for (int32_t v = bounds[p]; v < bounds[p + 1]; v++) part[v] = p;
  memset(out, 0, sizeof(uint64_t) * (size_t)k);
  for (int64_t u = 0; u < n; u++) {
    const int32_t* a = colidx + rowptr[u];
    int64_t da = rowptr[u + 1] - rowptr[u];
    for (int64_t j = 0; j < da; j++) {
      int32_t v = a[j];
      int64_t db = rowptr[v + 1] - rowptr[v];
      int32_t owner = part[u];
      if (part[v] != part[u] && da < db) owner = part[v];
      out[owner] += (uint64_t)oracle_intersect_merge(a, da, colidx + rowptr[v], db);
    }
  }
  free(part);
}
