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

/* Launch parameters of the degree-binned count (picked by tc_autotune). */
typedef struct tc_tune {
  int32_t t_g8;    /* max wedge elements of a pivot handled by an 8-lane group      */
  int32_t t_g16;   /* ... by a 16-lane group                                        */
  int32_t t_g32;   /* ... by a 32-lane group                                        */
  int32_t t_wave;  /* ... by one wave64; heavier pivots get a whole workgroup       */
  int32_t rounds;  /* pivots per group in one light-bin tile                          */
  int32_t grid;    /* persistent grid in workgroups (0 = full residency)             */
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

/* ---- counting -----------------------------------------------------------
 * tc_count: degree-binned persistent count (8/16/32-lane groups for light
 * pivots, one wave per medium pivot, one workgroup per heavy pivot; pivot
 * list staged in LDS).  Replaces TCSolver's launch + copy-back
 * (successful_push_pull/multigpu_nvshmem_pushpull.cu:69-83) around
 * warp_vertex (bs_warp_vertex.cuh:2-35). */
tc_status tc_count(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_stats* stats);
/* tc_count_plain: one launch of the plain kernel, one wave64 per pivot,
 * grid-stride over pivots at full residency -- the direct restatement of
 * warp_vertex (bs_warp_vertex.cuh:2-35, grid a17 of
 * multigpu_nvshmem_pushpull.cu:41-51). */
tc_status tc_count_plain(tc_ctx* ctx, tc_graph* g, uint64_t* total, tc_stats* stats);

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

/* ---- measurement utility ------------------------------------------------
 * Copies `bytes` (a multiple of 16) between two device buffers with a plain
 * streaming kernel `iters` times and reports the mean kernel time.  No
 * counterpart in the reference: it gives a kernel of known HBM traffic
 * (bytes read + bytes written) to calibrate profiler traffic counters. */
tc_status tc_probe_copy(tc_ctx* ctx, int64_t bytes, int32_t iters, double* ms_per_iter);

#ifdef __cplusplus
}
#endif
#endif /* TC_H */
