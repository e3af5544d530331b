/*
 * tc_types.h -- types shared by the host graph library (tc_host.h) and the
 * device counting library (tc.h).
 *
 * Id and offset widths follow the reference's GraphMiner typedefs as used on
 * the triangle-counting path: vidType = int32 (nvshmemx_int32_get_warp on
 * colidx, successful_push_pull/bs_warp_vertex_nvshmem.cuh:134), eidType = int64
 * (nvshmem_int64_g on rowptr, bs_warp_vertex_nvshmem.cuh:114-115), AccType =
 * uint64 (uint64_t total, successful_push_pull/main.cc:32).
 */
#ifndef TC_TYPES_H
#define TC_TYPES_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Every entry point returns a status instead of calling exit() or assert()
 * (the reference exits on CUDA errors and asserts on bad meta data:
 * successful_push_pull/multigpu_nvshmem_pushpull.cu:33-36, graph.cc:47). */
typedef enum tc_status {
  TC_OK = 0,
  TC_EINVAL = 1,    /* malformed argument or malformed CSR                      */
  TC_ENOMEM = 2,    /* host/device allocation failed, or over the HBM budget     */
  TC_EDEVICE = 3,   /* HIP runtime error                                         */
  TC_EBUSY = 4,     /* queue full / object in use                                */
  TC_ERANGE = 5,    /* size beyond a supported maximum                           */
  TC_EIO = 6,       /* file could not be opened, read or written                 */
  TC_ETIMEDOUT = 7, /* a bounded wait expired                                    */
  TC_ESTATE = 8     /* call not valid in the object's current state              */
} tc_status;

/* A borrowed CSR: rowptr has n_rows+1 entries with rowptr[0] == 0 and
 * rowptr[n_rows] == nnz; colidx holds nnz neighbour ids in [0, n_rows), each
 * row sorted ascending.  The same layout as the reference's Graph
 * (vertices eid[n+1], edges vid[m]; successful_push_pull/graph.h:35-36). */
typedef struct tc_csr_view {
  const int64_t* rowptr;
  const int32_t* colidx;
  int32_t n_rows;
  int64_t nnz;
} tc_csr_view;

/* Validation flags for tch_validate / tc_graph_upload. */
#define TC_CHECK_SHAPE   1u  /* rowptr[0]==0, monotone, rowptr[n]==nnz, ids in range */
#define TC_CHECK_SORTED  2u  /* every row strictly increasing (sorted, no duplicates) */
#define TC_CHECK_NOSELF  4u  /* no self-loops                                         */
#define TC_CHECK_ALL     7u

#ifdef __cplusplus
}
#endif
#endif /* TC_TYPES_H */
