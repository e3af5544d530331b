"""ctypes bindings of the two native libraries (``lib/libtc_host.so`` and
``lib/libtc_hip.so``).

The libraries are built in-tree by ``build.py``.  Loading fails loudly when
one is missing: there is no Python or CPU fallback for any device entry
point.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
# HPP_TC_LIB_DIR selects another build of the two libraries (e.g. a tuning variant).
LIB_DIR = os.environ.get("HPP_TC_LIB_DIR") or os.path.join(_PKG, "lib")
HOST_LIB = os.path.join(LIB_DIR, "libtc_host.so")
HIP_LIB = os.path.join(LIB_DIR, "libtc_hip.so")

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_dbl = ctypes.c_double
i32p = ctypes.POINTER(c_i32)
i64p = ctypes.POINTER(c_i64)
u64p = ctypes.POINTER(c_u64)
vp = ctypes.c_void_p

TC_OK, TC_EINVAL, TC_ENOMEM, TC_EDEVICE, TC_EBUSY, TC_ERANGE, TC_EIO, TC_ETIMEDOUT, TC_ESTATE = range(9)
STATUS_NAMES = {0: "TC_OK", 1: "TC_EINVAL", 2: "TC_ENOMEM", 3: "TC_EDEVICE", 4: "TC_EBUSY",
                5: "TC_ERANGE", 6: "TC_EIO", 7: "TC_ETIMEDOUT", 8: "TC_ESTATE"}
TC_CHECK_SHAPE, TC_CHECK_SORTED, TC_CHECK_NOSELF, TC_CHECK_ALL = 1, 2, 4, 7


class TcError(RuntimeError):
    """A native call returned a non-OK tc_status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class CsrView(ctypes.Structure):
    _fields_ = [("rowptr", i64p), ("colidx", i32p), ("n_rows", c_i32), ("nnz", c_i64)]


class BatchSpec(ctypes.Structure):
    _fields_ = [("kind", c_i32), ("n", c_i32), ("avg_degree", c_dbl), ("seed", c_u64)]


class PartitionView(ctypes.Structure):
    _fields_ = [("n_global", c_i32), ("row_begin", c_i32), ("row_end", c_i32),
                ("local_rowptr", i64p), ("local_colidx", i32p), ("local_nnz", c_i64),
                ("n_halo", c_i32), ("halo_ids", i32p), ("halo_rowptr", i64p),
                ("halo_colidx", i32p), ("halo_nnz", c_i64)]


class MetisView(ctypes.Structure):
    _fields_ = [("n", c_i32), ("n_real", c_i32), ("nnz", c_i64), ("max_subg_nv", c_i64),
                ("max_subg_ne", c_i64), ("max_degree", c_i64), ("rowptr", i64p), ("colidx", i32p),
                ("vertex_list", i32p), ("part", i64p)]


class Work(ctypes.Structure):
    _fields_ = [("rows", c_i64), ("edges", c_i64), ("wedge_elems", c_i64), ("min_elems", c_i64),
                ("alg_bytes", c_i64), ("max_out", c_i32), ("partner_rows", c_i64),
                ("partner_bytes", c_i64)]


PARTNERS_PER_ROW = 1024  # TCH_PARTNERS_PER_ROW
# tch_partition_bounds criteria
BALANCE = {"vertices": 0, "edges": 1, "wedges": 2, "partner": 3}


class ArenaStats(ctypes.Structure):
    _fields_ = [("capacity", c_u64), ("used", c_u64), ("peak_used", c_u64), ("largest_free", c_u64),
                ("live_extents", c_i64), ("free_extents", c_i64), ("allocs", c_i64),
                ("failed_allocs", c_i64)]


_host = None


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def host():
    """The host graph library; raises if it has not been built."""
    global _host
    if _host is not None:
        return _host
    if not os.path.exists(HOST_LIB):
        raise RuntimeError(f"{HOST_LIB} is missing: run `python -m "
                           "hetegeneous_parallel_programming_amd.build`")
    L = ctypes.CDLL(HOST_LIB, mode=ctypes.RTLD_GLOBAL)
    P = vp
    _sig(L, "tch_last_error", ctypes.c_char_p, [])
    _sig(L, "tch_csr_free", None, [P])
    _sig(L, "tch_csr_view", c_i32, [P, ctypes.POINTER(CsrView)])
    _sig(L, "tch_csr_max_degree", c_i32, [P])
    _sig(L, "tch_csr_copy", c_i32, [ctypes.POINTER(CsrView), c_u32, ctypes.POINTER(P)])
    _sig(L, "tch_csr_from_pairs", c_i32, [c_i32, i32p, i32p, c_i64, ctypes.POINTER(P)])
    _sig(L, "tch_validate", c_i32, [ctypes.POINTER(CsrView), c_u32])
    _sig(L, "tch_intersect_num", c_i64, [ctypes.POINTER(CsrView), c_i32, c_i32,
                                        ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint8])
    _sig(L, "tch_gen_er", c_i32, [c_i32, c_dbl, c_u64, ctypes.POINTER(P)])
    _sig(L, "tch_gen_rmat", c_i32, [c_i32, c_dbl, c_dbl, c_dbl, c_dbl, c_u64, ctypes.POINTER(P)])
    _sig(L, "tch_batch_spec_of", c_i32, [c_u64, c_i32, ctypes.POINTER(BatchSpec)])
    _sig(L, "tch_batch_graph", c_i32, [c_u64, c_i32, ctypes.POINTER(P)])
    _sig(L, "tch_batch_graphs", c_i32, [c_u64, c_i32, c_i32, ctypes.POINTER(P)])
    _sig(L, "tch_orient", c_i32, [ctypes.POINTER(CsrView), ctypes.POINTER(P)])
    _sig(L, "tch_sort_neighbors", c_i32, [P])
    _sig(L, "tch_partition_range", c_i32, [ctypes.POINTER(CsrView), c_i32, c_i32, ctypes.POINTER(P)])
    _sig(L, "tch_partition_bounds", c_i32, [ctypes.POINTER(CsrView), c_i32, c_i32, i32p])
    _sig(L, "tch_partition_span", c_i32, [ctypes.POINTER(CsrView), c_i32, c_i32, ctypes.POINTER(P)])
    _sig(L, "tch_partition_get", c_i32, [P, ctypes.POINTER(PartitionView)])
    _sig(L, "tch_partition_free", None, [P])
    _sig(L, "tch_partition_induced", c_i32, [ctypes.POINTER(CsrView), c_i32, c_i32,
                                             ctypes.POINTER(P), i32p, i32p,
                                             ctypes.POINTER(i32p)])
    _sig(L, "tch_free_ids", None, [i32p])
    _sig(L, "tch_work_of", c_i32, [ctypes.POINTER(CsrView), c_i32, c_i32, ctypes.POINTER(Work)])
    _sig(L, "tch_work_of_partition", c_i32, [ctypes.POINTER(PartitionView), ctypes.POINTER(Work)])
    _sig(L, "tch_read_graphminer", c_i32, [ctypes.c_char_p, c_i32, ctypes.POINTER(P)])
    _sig(L, "tch_write_graphminer", c_i32, [ctypes.c_char_p, ctypes.POINTER(CsrView)])
    _sig(L, "tch_partition_assemble", c_i32, [ctypes.POINTER(PartitionView), ctypes.POINTER(P)])
    _sig(L, "tch_metis_write", c_i32, [ctypes.c_char_p, ctypes.POINTER(CsrView), i64p, c_i32])
    _sig(L, "tch_metis_read", c_i32, [ctypes.c_char_p, c_i32, ctypes.POINTER(P)])
    _sig(L, "tch_metis_get", c_i32, [P, ctypes.POINTER(MetisView)])
    _sig(L, "tch_metis_free", None, [P])
    _sig(L, "tch_metis_assemble", c_i32, [ctypes.POINTER(P), c_i32, c_i32, ctypes.POINTER(P),
                                          ctypes.POINTER(ctypes.c_uint8)])
    _sig(L, "tch_arena_create", c_i32, [c_u64, c_u64, ctypes.POINTER(P)])
    _sig(L, "tch_arena_destroy", None, [P])
    _sig(L, "tch_arena_alloc", c_i32, [P, c_u64, ctypes.POINTER(c_u64)])
    _sig(L, "tch_arena_free", c_i32, [P, c_u64])
    _sig(L, "tch_arena_stats_get", c_i32, [P, ctypes.POINTER(ArenaStats)])
    _host = L
    return L


def host_check(status: int) -> None:
    if status != TC_OK:
        raise TcError(status, host().tch_last_error().decode())
