/*
 * tc_host.h -- C ABI of the host graph substrate (libtc_host.so).
 *
 * Replaces the host side of the reference's triangle-counting path:
 * the Graph loader and CSR container (successful_push_pull/graph.h:14-188,
 * graph.cc:4-142), degree-ordered DAG orientation (graph.cc:252-298),
 * neighbour sorting (graph.cc:156-164) and the contiguous-range partitioner
 * (graph_partition.cc:27-135).  Seeded ER / RMAT generators stand in for the
 * GraphMiner input files the reference reads.
 *
 * All functions are thread-safe for distinct objects, never exit, and report
 * failures through tc_status plus tch_last_error() (per calling thread).
 * Parallel work uses OpenMP and honours OMP_NUM_THREADS.
 */
#ifndef TC_HOST_H
#define TC_HOST_H

#include "tc_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Owned host CSR: int64 rowptr[n+1], int32 colidx[nnz]. */
typedef struct tch_csr tch_csr;

/* Message for the last failing call on this thread ("" if none). */
const char* tch_last_error(void);

void tch_csr_free(tch_csr* g);

/* Borrowed view of an owned CSR; valid until tch_csr_free. */
tc_status tch_csr_view(const tch_csr* g, tc_csr_view* out);

/* Max row length (the reference's Graph::get_max_degree, graph.h:84). */
int32_t tch_csr_max_degree(const tch_csr* g);

/* Deep copy of a borrowed CSR (checked with `check` flags, TC_CHECK_*). */
tc_status tch_csr_copy(const tc_csr_view* src, uint32_t check, tch_csr** out);

/* Build the symmetric, self-loop-free, deduplicated, row-sorted CSR of an
 * undirected graph from m endpoint pairs (src[i], dst[i]); pairs with an
 * endpoint < 0 are skipped, endpoints >= n are an error. */
tc_status tch_csr_from_pairs(int32_t n, const int32_t* src, const int32_t* dst,
                             int64_t m, tch_csr** out);

/* Structural checks on a CSR (TC_CHECK_* flags). */
tc_status tch_validate(const tc_csr_view* g, uint32_t check);

/* |N(v) ∩ N(u)| of two sorted rows by a two-pointer merge, counting only
 * common ids w with vlabels[w] == label when vlabels is not NULL (one byte
 * per row).  Graph::intersect_num (graph.cc:376-391) on the host; it is the
 * reference's Graph API, not part of the GPU count.  Returns -1 (and sets
 * the error) when v or u is out of range. */
int64_t tch_intersect_num(const tc_csr_view* g, int32_t v, int32_t u, const uint8_t* vlabels, uint8_t label);

/* Seeded generators (symmetric CSR out).
 *  ER: G(n, M) with M = llround(n*avg_degree/2) uniformly drawn pairs (with
 *      replacement), self-loops dropped, symmetrised and deduplicated.
 *  RMAT: M = llround(edge_factor*n) pairs drawn in [0, 2^scale)^2 with
 *      quadrant probabilities (a, b, c, 1-a-b-c), scale = ceil(log2 n), no
 *      noise and no relabelling; pairs with an endpoint >= n are rejected,
 *      self-loops dropped, symmetrised and deduplicated.
 * Both are counter-based: pair i depends only on (seed, i), so the output is
 * identical for any thread count. */
tc_status tch_gen_er(int32_t n, double avg_degree, uint64_t seed, tch_csr** out);
tc_status tch_gen_rmat(int32_t n, double edge_factor, double a, double b, double c,
                       uint64_t seed, tch_csr** out);

/* One graph of the batched workload (4096 graphs; n log-uniform in
 * [2^8, 2^16]; even i -> ER, odd i -> RMAT(.57,.19,.19,.05); average degree
 * uniform in [4, 64], RMAT edge factor = degree/2). */
typedef struct tch_batch_spec {
  int32_t kind;        /* 0 = ER, 1 = RMAT */
  int32_t n;
  double avg_degree;
  uint64_t seed;       /* generator seed of this graph */
} tch_batch_spec;
tc_status tch_batch_spec_of(uint64_t batch_seed, int32_t index, tch_batch_spec* out);
/* Generate graph `index` of the batch and orient it (DAG out). */
tc_status tch_batch_graph(uint64_t batch_seed, int32_t index, tch_csr** dag_out);
/* Generate and orient graphs [first, first+count) in parallel over graphs. */
tc_status tch_batch_graphs(uint64_t batch_seed, int32_t first, int32_t count,
                           tch_csr** dags_out);

/* Degree-ordered DAG orientation of a symmetric CSR (graph.cc:252-298):
 * keep src->dst iff deg(dst) > deg(src) or (deg(dst) == deg(src) and
 * dst > src), with undirected degrees; row order is preserved. */
tc_status tch_orient(const tc_csr_view* sym, tch_csr** dag_out);

/* Sort every row in place (graph.cc:156-164). */
tc_status tch_sort_neighbors(tch_csr* g);

/* ---- partitioning ----------------------------------------------------- */

/* One partition of a contiguous-range split of a DAG into k parts
 * (graph_partition.cc:85-107: chunk = ceil(n/k), part p owns ids
 * [p*chunk, min(n, (p+1)*chunk))), plus the HBM halo cache: every remote
 * adjacency list N+(v), v outside the owned range, that an owned row's edges
 * reach.  All ids stay global. */
typedef struct tch_partition tch_partition;

typedef struct tch_partition_view {
  int32_t n_global;          /* vertices of the whole graph                  */
  int32_t row_begin;         /* owned ids [row_begin, row_end)               */
  int32_t row_end;
  const int64_t* local_rowptr; /* (row_end-row_begin)+1 offsets              */
  const int32_t* local_colidx; /* owned rows' DAG lists, global ids          */
  int64_t local_nnz;
  int32_t n_halo;            /* remote vertices whose lists are cached       */
  const int32_t* halo_ids;   /* sorted ascending                             */
  const int64_t* halo_rowptr;/* n_halo+1 offsets                             */
  const int32_t* halo_colidx;/* their full DAG lists, global ids             */
  int64_t halo_nnz;
} tch_partition_view;

tc_status tch_partition_range(const tc_csr_view* dag, int32_t k, int32_t part,
                              tch_partition** out);

/* Balanced contiguous ranges: bounds[0..k] with bounds[0] = 0, bounds[k] = n,
 * part p owning [bounds[p], bounds[p+1]).  The criterion weighs each row u:
 *   VERTICES  1 per row, exactly the reference split (graph_partition.cc:88-107)
 *   EDGES     d+(u)
 *   WEDGES    d+(u) + sum over v in N+(u) of d+(v)   (the elements an
 *             edge-iterator count streams, bs_warp_vertex.cuh:13-22)
 *   PARTNER   sum over v in N+(u) of 65536 * (6 + min(d+(u), d+(v)))
 *             + floor(65536 * 1503 * ceil(c_o / TCH_PARTNERS_PER_ROW) / c_o),
 *             c_o = the edges the edge's owner o owns (o = v when
 *             d+(v) > d+(u), else u): the ids the partner-scheme count
 *             streams, plus per edge and per partner row the costs measured
 *             on MI355X, in streamed-id equivalents (fixed point)
 * and each split point is the row where the prefix of weights is nearest to
 * p/k of the total.  No counterpart in the reference, whose vertex split puts
 * about half of an RMAT graph's edges into part 0. */
enum {
  TCH_BALANCE_VERTICES = 0,
  TCH_BALANCE_EDGES = 1,
  TCH_BALANCE_WEDGES = 2,
  TCH_BALANCE_PARTNER = 3
};
tc_status tch_partition_bounds(const tc_csr_view* dag, int32_t k, int32_t balance,
                               int32_t* bounds);
/* The partition owning rows [row_begin, row_end) with its halo cache (the
 * same structure tch_partition_range builds for a reference range). */
tc_status tch_partition_span(const tc_csr_view* dag, int32_t row_begin, int32_t row_end,
                             tch_partition** out);
tc_status tch_partition_get(const tch_partition* p, tch_partition_view* out);
void tch_partition_free(tch_partition* p);

/* Lays a partition out as one CSR over all n_global ids: owned rows and
 * halo rows hold their DAG lists, every other row is empty (the layout of the
 * reference's per-PE METIS subgraph files, graph_partition.cc:467-497, with
 * the halo rows added).  Counting rows [row_begin, row_end) of it gives the
 * partition's count. */
tc_status tch_partition_assemble(const tch_partition_view* p, tch_csr** out);

/* The reference's vertex-induced 1-D split (graph_partition.cc:27-135):
 * subgraph p keeps the owned vertices plus their neighbours (in `g`), with
 * ids re-indexed in ascending global order; local_begin/local_end give the
 * local ids of the first and one-past-last owned vertex and idx_map (length
 * n of the subgraph) maps local ids back to global ids. */
tc_status tch_partition_induced(const tc_csr_view* g, int32_t k, int32_t part,
                                tch_csr** sub_out, int32_t* local_begin,
                                int32_t* local_end, int32_t** idx_map_out);
void tch_free_ids(int32_t* p);

/* ---- the reference's per-PE METIS part files -----------------------------
 * File set of PE i (graph_partition.cc:501-566, read by graph.cc:16-23,
 * 33-34,56-59), <prefix> typically ".../metis/pgraph":
 *   <prefix>i.meta.txt        n, n_real, E, max_subg_nv, max_subg_ne, max_degree
 *   <prefix>i.vertex.bin      int64 rowptr[n+1] over all global ids, only
 *                             PE i's rows non-empty
 *   <prefix>i.edge.bin        int32 colidx[E], global ids
 *   <prefix>i.vertex_list.bin int32 PE i's ids [n_real], ascending
 *   <prefix>i.part.bin        int64 owner PE of every id [n]
 * tch_metis_write lays out any DAG under any part array (values in [0, k))
 * as k such file sets (METIS itself is not linked: its part array, or a range
 * split, is an input).  tch_metis_read loads and checks one PE's set.
 * tch_metis_assemble builds, from all k PEs' files, the CSR PE i counts: its
 * own rows plus every remote list they reach (each from its owner's file,
 * the lists the reference pulls, bs_warp_vertex_nvshmem.cuh:110-136), other
 * rows empty; `owned` (n bytes, may be NULL) gets 1 for PE i's rows, the
 * pivots of its count (vertex_list, bs_warp_vertex_nvshmem.cuh:25-28).
 * i = -1 assembles the whole DAG (every row from its owner, owned all 1). */
typedef struct tch_metis tch_metis;
typedef struct tch_metis_view {
  int32_t n;
  int32_t n_real;
  int64_t nnz;
  int64_t max_subg_nv, max_subg_ne, max_degree;
  const int64_t* rowptr;
  const int32_t* colidx;
  const int32_t* vertex_list;
  const int64_t* part;
} tch_metis_view;
tc_status tch_metis_write(const char* prefix, const tc_csr_view* dag, const int64_t* part, int32_t k);
tc_status tch_metis_read(const char* prefix, int32_t i, tch_metis** out);
tc_status tch_metis_get(const tch_metis* m, tch_metis_view* out);
void tch_metis_free(tch_metis* m);
tc_status tch_metis_assemble(const tch_metis* const* pes, int32_t k, int32_t i, tch_csr** out,
                             uint8_t* owned);

/* ---- algorithmic work of a count ----------------------------------------- */

typedef struct tch_work {
  int64_t rows;          /* R: pivot rows counted                              */
  int64_t edges;         /* m_R: DAG edges of those rows                       */
  int64_t wedge_elems;   /* sum over (u,v), u in R, of d+(v)                    */
  int64_t min_elems;     /* sum over (u,v) of min(d+(u), d+(v))                 */
  int64_t alg_bytes;     /* edge-iterator bytes (every N+(v) streamed per edge):
                            8*(R+1) + 20*m_R + 4*wedge_elems                     */
  int32_t max_out;       /* max d+ over pivot rows                               */
  int64_t partner_rows;  /* work units of the partner scheme: per owner x,
                            ceil(c(x) / TCH_PARTNERS_PER_ROW), c(x) = edges it owns */
  int64_t partner_bytes; /* bytes the partner scheme's count reads with no reuse:
                            per unit its 16-byte row record, 16 bytes of row
                            offsets and the owner's list (4*d+(x)); per edge an
                            8-byte list reference and 4*min(d+(u), d+(v))        */
} tch_work;

/* Partner scheme: edge (u, v) is owned by u when d+(v) <= d+(u), else by v;
 * the owner's list is the set, the other endpoint's list is streamed.  An
 * owner's edges are cut into units of at most this many. */
#define TCH_PARTNERS_PER_ROW 1024

/* Work of counting rows [row_begin, row_end) of a DAG (rows outside use the
 * lists in `dag` directly). */
tc_status tch_work_of(const tc_csr_view* dag, int32_t row_begin, int32_t row_end,
                      tch_work* out);
/* Same for a partition with its halo cache. */
tc_status tch_work_of_partition(const tch_partition_view* p, tch_work* out);

/* ---- arena bookkeeping for device memory pools ---------------------------
 * First-fit placement with coalescing over [0, capacity); offsets are
 * multiples of `align`.  Pure host bookkeeping (no device calls): the
 * serving loop places device-resident CSRs in one HBM allocation with it.
 * No counterpart in the reference, which allocates per graph
 * (successful_push_pull/graph_nvshmem.h:38-43). */
typedef struct tch_arena tch_arena;
tc_status tch_arena_create(uint64_t capacity, uint64_t align, tch_arena** out);
void tch_arena_destroy(tch_arena* a);
/* TC_ENOMEM when no free extent is large enough. */
tc_status tch_arena_alloc(tch_arena* a, uint64_t bytes, uint64_t* offset);
/* TC_EINVAL when offset is not the start of a live extent. */
tc_status tch_arena_free(tch_arena* a, uint64_t offset);
typedef struct tch_arena_stats {
  uint64_t capacity, used, peak_used, largest_free;
  int64_t live_extents, free_extents, allocs, failed_allocs;
} tch_arena_stats;
tc_status tch_arena_stats_get(const tch_arena* a, tch_arena_stats* out);

/* ---- GraphMiner on-disk format (graph.cc:31-55) -------------------------- */

/* Reads prefix.meta.txt ("n m vid_size eid_size vlabel_size elabel_size
 * max_degree feat_len n_vlabel_classes n_elabel_classes"), prefix.vertex.bin
 * (int64[n+1]) and prefix.edge.bin (int32[m]).  If use_dag is non-zero the
 * graph is oriented after loading (graph.cc:135-138). */
tc_status tch_read_graphminer(const char* prefix, int use_dag, tch_csr** out);
tc_status tch_write_graphminer(const char* prefix, const tc_csr_view* g);

#ifdef __cplusplus
}
#endif
#endif /* TC_HOST_H */
