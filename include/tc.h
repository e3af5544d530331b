/*
 * tc.h -- C ABI of the MI355X triangle-counting library (libtc_hip.so).
 *
 * The hot path is the reference's bs_warp_vertex set-intersection count
 * (bs_warp_vertex.cuh:2-35; loop nest :13-22): for every pivot u and every
 * v in N+(u) of the degree-ordered DAG, add |N+(u) ∩ N+(v)|.  Here it runs as
 * hand-written HIP kernels for gfx950.  All calls are stream-ordered on the
 * context's HIP stream and return when their results are on the host.
 *
 * Ownership: host arrays passed in are only read during the call; device
 * memory belongs to the context until the matching *_free.  Counts are
 * written (never accumulated into) the caller's output.
 *
 * Thread-safety: one context may be shared by threads (calls serialise on
 * an internal lock); tc_last_error reports the calling thread's last failure
 * on that context.
 */
#ifndef TC_H
#define TC_H

#include "tc_types.h"
#include "tc_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tc_ctx tc_ctx;
typedef struct tc_graph tc_graph;
typedef struct tc_batch tc_batch;
typedef struct tc_server tc_server;

/* Launch parameters of the degree-binned count (defaults, or picked by
 * tc_autotune).  The thresholds choose each pivot's group width by its wedge
 * work T(u) = sum of d+(v) over v in N+(u): at most t_g8 -> 8 lanes, ...,
 * above t_wave -> a whole workgroup; 0 for a threshold empties its bin. */
typedef struct tc_tune {
  int32_t t_g8;    /* max wedge elements of a pivot handled by an 8-lane group      */
  int32_t t_g16;   /* ... by a 16-lane group                                        */
  int32_t t_g32;   /* ... by a 32-lane group                                        */
  int32_t t_wave;  /* ... by one wave64; heavier pivots get a whole workgroup       */
  int32_t rounds;  /* pivots per group in one light-bin tile                          */
  int32_t grid;    /* persistent grid in workgroups (0 = full residency)             */
  int32_t pivot_set;  /* LDS layout of N+(u): 0 = cuckoo hash where it fits, else the
                         binary-search tree; 1 = always the tree; 2 = the tree for
                         8/16/32-lane groups, the hash for waves and workgroups      */
  int32_t hash_kicks; /* evictions allowed per insert before a hash build gives up and
                         the pivot falls back to the tree (1..1024)                   */
} tc_tune;

#define TC_NUM_BINS 6

/* Per-call measurements. */
typedef struct tc_stats {
  double total_ms;            /* host wall time of the call                        */
  double plan_ms;             /* degree-binning kernels (HIP events)               */
  double count_ms;            /* counting kernel(s) (HIP events)                   */
  int64_t pivots;             /* pivot rows considered                             */
  int64_t bin_pivots[TC_NUM_BINS]; /* pivots per bin: global-list, workgroup, wave,
                                      32-lane, 16-lane, 8-lane                      */
  int64_t hash_fallbacks;     /* pivot sets whose LDS hash build did not converge and
                                 fell back to the search tree (binned counts)       */
  int64_t single_launch;      /* 1: counted by the single plain launch                */
  int64_t plan_kept;          /* 1: the binned count reused the graph's kept plan      */
} tc_stats;

/* ---- context ------------------------------------------------------------
 * Replaces the per-process device setup of the reference solver
 * (successful_push_pull/multigpu_nvshmem_pushpull.cu:20-51: device pick and
 * the maximum_residency grid).  A device index the runtime does not have
 * returns TC_EDEVICE instead of exit(1) (:33-36). */
tc_status tc_ctx_create(int32_t device, tc_ctx** out);
void tc_ctx_destroy(tc_ctx* ctx);
const char* tc_last_error(const tc_ctx* ctx);    /* ctx may be NULL */
tc_status tc_get_tune(tc_ctx* ctx, tc_tune* out);
tc_status tc_set_tune(tc_ctx* ctx, const tc_tune* tune);
tc_status tc_device_name(tc_ctx* ctx, char* buf, int32_t len);

/* ---- graphs on the device ----------------------------------------------
 * Replaces GraphGPU::init (successful_push_pull/graph_gpu.h:84-137: device
 * rowptr/colidx + H2D).  `dag` must be the oriented DAG (tch_orient);
 * `check` (TC_CHECK_*) validates it on the host first. */
tc_status tc_graph_upload(tc_ctx* ctx, const tc_csr_view* dag, uint32_t check, tc_graph** out);
/* Same, counting only the rows u with owned[u] != 0 (n_rows bytes, copied):
 * a PE's share when its rows are an arbitrary vertex set, as with METIS
 * parts (bs_warp_vertex_nvshmem.cuh:25-28 loops over vertex_list); the other
 * rows hold the remote lists they reach (tch_metis_assemble).  NULL = all. */
tc_status tc_graph_upload_rows(tc_ctx* ctx, const tc_csr_view* dag, const uint8_t* owned, uint32_t check,
                               tc_graph** out);

/* One GPU's share of the push/pull scheme (GraphNVSHMEM,
 * successful_push_pull/graph_nvshmem.h:26-50, plus the pull branch of
 * bs_warp_vertex_nvshmem.cuh:110-136): the owned rows and the halo cache of
 * every remote list they reach, laid out in one global-id CSR in HBM.
 * tc_count on it counts the owned pivots only. */
tc_status tc_partition_upload(tc_ctx* ctx, const tch_partition_view* part, uint32_t check,
                              tc_graph** out);
tc_status tc_graph_free(tc_ctx* ctx, tc_graph* g);
/* Device bytes held by a graph. */
int64_t tc_graph_bytes(const tc_graph* g);

/* What an upload built.  Besides the DAG, every upload builds on the device
 * the partner structure the count works from: each DAG edge (u, v) is owned
 * by u when d+(v) <= d+(u), else by v; the owner's list N+(x) becomes the
 * LDS set and the other endpoint's list is streamed through it, so an edge
 * streams min(d+(u), d+(v)) ids (the reference streams d+(v) for every edge,
 * bs_warp_vertex.cuh:13-22).  An owner's edges are cut into units of at most
 * TCH_PARTNERS_PER_ROW ("partner rows").  No counterpart in the reference. */
typedef struct tc_graph_info {
  int32_t graphs;         /* 1 for a graph, the graph count for a batch          */
  int64_t rows;           /* DAG rows                                           */
  int64_t nnz;            /* DAG ids                                            */
  int64_t partner_rows;   /* partner rows built (units of work of the count)    */
  double build_ms;        /* device time of the partner build                   */
  int64_t bytes;          /* device bytes held                                  */
} tc_graph_info;
tc_status tc_graph_get_info(const tc_graph* g, tc_graph_info* out);

/* ---- counting -----------------------------------------------------------
 * tc_count: degree-binned persistent count over the graph's partner rows
 * (8/16/32-lane groups for light rows, one wave per medium row, one
 * workgroup per heavy row; the owner's list held in LDS as a hash or a
 * search tree, the partners' lists streamed through it).  Same total as the
 * reference's loop nest.  Replaces TCSolver's launch + copy-back
 * (successful_push_pull/multigpu_nvshmem_pushpull.cu:69-83) around
 * warp_vertex (bs_warp_vertex.cuh:2-35). */
tc_status tc_count(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_stats* stats);

/* tc_count_csr: one count of a host DAG, nothing kept -- the whole of
 * TCSolver(Graph&, uint64_t& total, ...) (successful_push_pull/main.cc:5;
 * multigpu_nvshmem_pushpull.cu:29-85: device copy, launch, copy-back).
 * check as for tc_graph_upload.  A graph that tc_count would count by its
 * single plain launch (see below) is copied
 * into a device buffer the context keeps for this call (grown when needed)
 * and counted without a partner structure; any other graph is uploaded,
 * counted and freed.  stats->single_launch tells which. */
tc_status tc_count_csr(tc_ctx* ctx, const tc_csr_view* dag, uint32_t check, uint64_t* total, tc_stats* stats);
/* tc_count_binned: always the degree-binned count (tc_count takes it for all
 * but small graphs; see below).  The binned count's plan is built on a
 * graph's first binned count and kept in the graph (device bytes: 10 per
 * partner-row slot) for later counts under the same tc_tune. */
tc_status tc_count_binned(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_stats* stats);
/* tc_count on a graph whose every list is under 1024 ids and whose counted
 * rows hold at most 2^18 DAG ids -- or, for a flat graph, at most 2^21 ids
 * in at most 2^17 rows with lists of at most 64 -- runs tc_count_plain's
 * single launch instead (stats->single_launch = 1): there one launch is
 * faster than planning plus the binned kernels' launches.  Both are exact.
 * tc_count_plain: one launch of the plain kernel, one wave64 per pivot,
 * grid-stride over pivots at full residency -- the direct restatement of
 * warp_vertex (bs_warp_vertex.cuh:2-35, grid a17 of
 * multigpu_nvshmem_pushpull.cu:41-51). */
tc_status tc_count_plain(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_stats* stats);

/* tc_count_noalloc: tc_count's count (the single plain launch for a small
 * graph, else the binned count, planned on every call) that allocates
 * nothing: no hipMalloc, hipFree or allocating library call.  No counterpart
 * in the reference; it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it).  Its plan
 * lives in a static device array of the library, about 20 MB, held on every
 * GPU the library runs on until the process ends: at most 2^21 plan slots
 * (rows + ids/1024 + 2), larger graphs return TC_ERANGE.  The graph must be
 * resident already (tc_graph_upload allocates).  Calls on one device share
 * that array: calls that overlap on one GPU race (one context serialises
 * its own calls; two contexts on one GPU must not count at the same time). */
tc_status tc_count_noalloc(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_stats* stats);

/* tc_count_local: the same count, also attributing every triangle to its
 * edges, its corners and the parts of a split.  Outputs (host memory, each
 * optional):
 *   edge_counts[e]   for each DAG edge e = (u, v) of the counted rows, in
 *                    CSR order from the first counted row (nnz entries for a
 *                    whole graph or a tc_graph_upload_rows graph, whose
 *                    uncounted rows read 0; local_nnz for a partition):
 *                    |N+(u) ∩ N+(v)|, the per-edge term the reference sums
 *                    (bs_warp_vertex.cuh:20);
 *   vertex_counts[v] for each of the graph's n rows: the counted triangles
 *                    with a corner at v (for a whole graph the local
 *                    clustering numerator; for a partition, the share of the
 *                    triangles whose lowest vertex it owns);
 *   part_counts[p]   for p < parts (1..1024) under split points
 *                    bounds[0..parts] (0 = bounds[0] <= ... <= bounds[parts]
 *                    = n): the sum each PE of the reference's push/pull
 *                    kernel reports (successful_push_pull/
 *                    bs_warp_vertex_nvshmem.cuh:127-136): an edge's count
 *                    goes to u's part when v is in the same part or
 *                    d+(u) >= d+(v), else to v's part.
 * The reference emits none of these but the per-PE sums; they make a
 * split's attribution checkable.  A call without vertex_counts attributes
 * each triangle to its edge only (edge_counts and part_counts are the same;
 * RMAT-22: 11.8 ms against 19.5 ms with vertex_counts).  A call with
 * part_counts alone (edge_counts and vertex_counts null, parts > 0) adds
 * each partner row's triangles to its owner's part, which is the part that
 * computes the row's edges under the push/pull rule: no per-edge counts,
 * the plain count's kernels, and no local buffers beyond the split points
 * and sums (the partner rows' owners come with every upload).  One graph
 * only (not batches); the first call with edge or vertex outputs adds the
 * partner entries' edge positions, per-edge and per-vertex buffers and the
 * per-part pass's table (each edge's row and push bit) to the graph
 * (tc_graph_get_info reports the bytes). */
typedef struct tc_local {
  uint32_t* edge_counts;
  uint64_t* vertex_counts;
  int32_t parts;
  const int32_t* bounds;
  uint64_t* part_counts;
} tc_local;
tc_status tc_count_local(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_local* out, tc_stats* stats);

/* ---- batches of graphs --------------------------------------------------
 * Many graphs resident at once; tc_batch_count runs one hipGraph-captured
 * sequence (binning + one persistent kernel pulling (graph, degree-bin)
 * tiles from a device work queue) and writes counts[i] for graph i.
 * tc_batch_count_each is the baseline: per graph one plain launch, a copy
 * back and a stream sync (ms_each may be NULL). */
tc_status tc_batch_create(tc_ctx* ctx, const tc_csr_view* dags, int32_t n_graphs, uint32_t check,
                          tc_batch** out);
tc_status tc_batch_count(tc_ctx* ctx, tc_batch* b, uint64_t* counts, tc_stats* stats);
tc_status tc_batch_count_each(tc_ctx* ctx, tc_batch* b, uint64_t* counts, double* ms_each,
                              tc_stats* stats);
tc_status tc_batch_free(tc_ctx* ctx, tc_batch* b);
int64_t tc_batch_bytes(const tc_batch* b);
int32_t tc_batch_graphs(const tc_batch* b);
tc_status tc_batch_get_info(const tc_batch* b, tc_graph_info* out);

/* ---- autotuning ----------------------------------------------------------
 * No counterpart in the reference, whose launch shape is fixed (one warp per
 * vertex, maximum-residency grid: multigpu_nvshmem_pushpull.cu:41-51).
 * Times the context's count on exactly one workload -- a resident graph `g`
 * (tc_count) or a batch `b` (tc_batch_count, re-captured per setting) --
 * under candidate tc_tune settings by coordinate descent over each field
 * (minimum of `reps` timed calls after one untimed call per setting), stops
 * after `budget_ms` of wall time, installs the fastest setting on the context
 * and writes it to *best.  Every setting's counts must equal the first
 * setting's (TC_EDEVICE otherwise).  best and report may be NULL. */
typedef struct tc_autotune_report {
  int32_t trials;       /* settings timed                                  */
  double start_ms;      /* time of the starting setting                    */
  double best_ms;       /* time of the installed setting                   */
  double elapsed_ms;    /* wall time of the whole search                   */
} tc_autotune_report;
tc_status tc_autotune(tc_ctx* ctx, tc_graph* g, tc_batch* b, int32_t reps, double budget_ms,
                      tc_tune* best, tc_autotune_report* report);
/* Tune file: one line "<device>|<key> t_g8 t_g16 t_g32 t_wave rounds grid
 * pivot_set hash_kicks"
 * per (device, workload key); <device> is tc_device_name with blanks and '|'
 * replaced by '_'.  Save replaces the context's record for `key`
 * (keeps other records); load installs it (TC_EIO if absent or malformed). */
tc_status tc_tune_save(tc_ctx* ctx, const char* path, const char* key);
tc_status tc_tune_load(tc_ctx* ctx, const char* path, const char* key);

/* ---- serving loop --------------------------------------------------------
 * Continuous batching of TC requests.  No counterpart in the reference,
 * which counts one pre-loaded graph per process (successful_push_pull/main.cc:11-36).
 * A server thread admits submitted requests in arrival order when their CSR
 * footprint fits the HBM pool (a request that cannot fit yet waits; up to
 * `lookahead` later ones may pass it); a request of at least `heavy_ids`
 * DAG ids is uploaded by DMA on a copy stream at admission while earlier
 * launches run, smaller ones by one copy kernel at the start of their
 * launch.  Whenever one of `slots` launch slots is free it launches every
 * admitted light request (up to the batch limits) as one batched count on
 * that slot's stream -- when another launch is running, only once
 * `fill_graphs` light requests are ready or the oldest has waited
 * `fill_wait_us`, so launches under load are large; a heavy request goes
 * alone, on any slot but the last when there are several, so light requests
 * never wait for one and always keep a slot.  Results are
 * delivered by tc_server_poll. */
typedef struct tc_server_config {
  uint64_t hbm_budget;       /* bytes of HBM for resident request CSRs            */
  int32_t slots;             /* concurrent launches (streams), 1..8               */
  int32_t max_batch_graphs;  /* requests per launch                               */
  int64_t max_batch_rows;    /* pivot rows per launch (sizes per-slot scratch)    */
  int32_t max_queue;         /* unfinished requests before submit returns TC_EBUSY */
  int32_t lookahead;         /* requests that may be admitted past a blocked head */
  uint64_t host_staging;     /* bytes of pinned host memory submit copies requests
                                into (async DMA to HBM); 0 = pageable copies     */
  int64_t max_batch_ids;     /* DAG ids per launch (sizes per-slot build scratch,
                                24 bytes per id); 0 = 32 Mi ids                  */
  int64_t heavy_ids;         /* DAG ids from which a request is heavy: uploaded by
                                DMA at admission, launched alone; 0 = 4 Mi ids   */
  int32_t fill_graphs;       /* a light launch that would share the GPU with
                                another waits until this many light requests are
                                ready ... (0 = 256; 1 = launch at once)          */
  int32_t fill_wait_us;      /* ... or the oldest has waited this long (0 = 2000) */
} tc_server_config;

typedef struct tc_request_result {
  uint64_t tag;
  uint64_t count;            /* u64 triangle count of the request's DAG           */
  int32_t status;            /* TC_OK or the failure of this request              */
  int32_t batch_graphs;      /* requests in the launch that counted it            */
  double t_submit_ms, t_admit_ms, t_launch_ms, t_done_ms; /* tc_clock_ms() times  */
} tc_request_result;

typedef struct tc_server_stats {
  int64_t submitted, completed, rejected, launches, queued, in_flight;
  double mean_batch_graphs;
  uint64_t pool_capacity, pool_used, pool_peak;
  /* where the time went since the server started (ms): callers' time inside
   * tc_server_submit (checks + copy into staging, summed over threads); the
   * server thread's time retiring finished launches, admitting requests,
   * issuing launches, and sleeping with work in flight or waiting for work;
   * device time of the launches (first command to completion, summed over
   * launches, which overlap on the slots' streams) */
  double submit_ms, retire_ms, admit_ms, launch_ms, wait_busy_ms, wait_idle_ms, device_ms;
  uint64_t staging_used;     /* pinned staging bytes held by requests not yet uploaded */
} tc_server_stats;

/* Monotonic clock (ms) used for every timestamp of the serving loop. */
double tc_clock_ms(void);
/* HBM bytes a request of this shape holds while resident. */
uint64_t tc_request_footprint(int32_t n_rows, int64_t nnz);
tc_status tc_server_create(tc_ctx* ctx, const tc_server_config* cfg, tc_server** out);
/* Copies the DAG (caller keeps ownership).  TC_EBUSY: queue full;
 * TC_ENOMEM: footprint exceeds the whole pool; TC_ERANGE: more rows than
 * max_batch_rows or more ids than max_batch_ids. */
tc_status tc_server_submit(tc_server* s, const tc_csr_view* dag, uint64_t tag, uint32_t check);
/* tc_server_submit at a scheduled time: waits (sleeping, then spinning for
 * the last ~2 ms) until tc_clock_ms() >= t_ms, then submits.  For clients
 * and load generators replaying an arrival schedule from several threads. */
tc_status tc_server_submit_at(tc_server* s, const tc_csr_view* dag, uint64_t tag, uint32_t check,
                              double t_ms);
/* Up to `max` finished results; waits up to timeout_ms for the first one. */
tc_status tc_server_poll(tc_server* s, tc_request_result* out, int32_t max, int32_t* n_out,
                         int32_t timeout_ms);
/* Waits until every submitted request has finished (TC_ETIMEDOUT otherwise). */
tc_status tc_server_drain(tc_server* s, int32_t timeout_ms);
tc_status tc_server_get_stats(tc_server* s, tc_server_stats* out);
/* Finishes all submitted work, then frees the server. */
tc_status tc_server_destroy(tc_server* s);

/* ---- measurement utility ------------------------------------------------
 * Copies `bytes` (a multiple of 16) between two device buffers with a plain
 * streaming kernel `iters` times and reports the mean kernel time.  No
 * counterpart in the reference: it gives a kernel of known HBM traffic
 * (bytes read + bytes written) to calibrate profiler traffic counters. */
tc_status tc_probe_copy(tc_ctx* ctx, int64_t bytes, int32_t iters, double* ms_per_iter);
/* Reads `bytes` (a multiple of 16) of a device buffer with a plain streaming
 * kernel `iters` times and reports the mean kernel time: the practical HBM
 * read rate, against which a kernel that only reads is measured.  No
 * counterpart in the reference. */
tc_status tc_probe_read(tc_ctx* ctx, int64_t bytes, int32_t iters, double* ms_per_iter);

#ifdef __cplusplus
}
#endif
#endif /* TC_H */
