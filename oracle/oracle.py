"""CPU ORACLE for the triangle-counting path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker.  Nothing in the
product package imports, links or calls it.

Two halves:

* ``liboracle.so`` (``tc_oracle.c``, built by ``oracle/Makefile``): the
  reference's CPU algorithm in plain C + OpenMP -- merge intersection
  (successful_push_pull/graph.cc:376-391), binary search (graph.cc:363-374,
  graph_gpu.h:228-252), orientation (graph.cc:252-298), the count loop nest
  (bs_warp_vertex.cuh:13-22) run as an OpenMP dynamic-schedule u64 reduction,
  the halo-cache partition count (bs_warp_vertex_nvshmem.cuh:27-40,110-136)
  and the push/pull per-PE attribution (:37-38,127-136,228-232).
* numpy restatements: the seeded ER/RMAT/batch generators (bit-exact with the
  product's C++ generators), symmetrisation, orientation, and brute-force
  triangle counts for tiny graphs.

How each function is pinned (tests/test_oracle.py):

* counts: closed forms (K_n -> C(n,3), wheels, bipartite/trees/even cycles ->
  0, disjoint unions add), O(n^3) brute force on n <= 64, and golden counts
  written by networkx.triangles and scipy trace(A^3)/6 in
  tests/golden/make_golden.py; merge and binary-search variants must agree.
* orientation: closed-form rule, acyclicity, nnz(DAG) == nnz(sym)/2,
  regular graphs give exactly the upper triangle.
* intersections: set-intersection size on random sorted lists.
* halo partition count: sum over all k parts equals the whole-graph count.
* push/pull attribution: sum over PEs equals the whole-graph count.
* generators: invariants (symmetric, sorted, deduplicated, loop-free) and
  golden SHA-256 digests of small generated CSRs.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")
_lib = None

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def lib():
    """Load liboracle.so (build it with ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH}; run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_binary_search.restype = ctypes.c_int
        L.oracle_binary_search.argtypes = [_i32p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32]
        for f in (L.oracle_intersect_merge, L.oracle_intersect_bsearch):
            f.restype = ctypes.c_int64
            f.argtypes = [_i32p, ctypes.c_int64, _i32p, ctypes.c_int64]
        L.oracle_orient_rowptr.restype = ctypes.c_int64
        L.oracle_orient_rowptr.argtypes = [ctypes.c_int32, _i64p, _i32p, _i64p]
        L.oracle_orient_fill.restype = None
        L.oracle_orient_fill.argtypes = [ctypes.c_int32, _i64p, _i32p, _i64p, _i32p]
        L.oracle_tc_count.restype = ctypes.c_uint64
        L.oracle_tc_count.argtypes = [ctypes.c_int32, _i64p, _i32p, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_tc_count_list.restype = ctypes.c_uint64
        L.oracle_tc_count_list.argtypes = [_i64p, _i32p, _i32p, ctypes.c_int64, ctypes.c_int,
                                           ctypes.c_int]
        L.oracle_tc_count_halo.restype = ctypes.c_uint64
        L.oracle_tc_count_halo.argtypes = [ctypes.c_int32, ctypes.c_int32, _i64p, _i32p,
                                           ctypes.c_int32, _i32p, _i64p, _i32p, ctypes.c_int,
                                           _i64p]
        L.oracle_tc_per_vertex.restype = None
        L.oracle_tc_per_vertex.argtypes = [ctypes.c_int32, _i64p, _i32p, _u64p]
        L.oracle_tc_edge_counts.restype = None
        L.oracle_tc_edge_counts.argtypes = [ctypes.c_int32, _i64p, _i32p,
                                            ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_tc_pushpull_bounds.restype = None
        L.oracle_tc_pushpull_bounds.argtypes = [ctypes.c_int32, _i64p, _i32p, ctypes.c_int32,
                                                _i32p, _u64p]
        L.oracle_tc_pushpull_pe.restype = ctypes.c_uint64
        L.oracle_tc_pushpull_pe.argtypes = [ctypes.c_int32, _i64p, _i32p, ctypes.c_int32,
                                            ctypes.c_int32]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# --------------------------------------------------------------------------
# C oracle wrappers
# --------------------------------------------------------------------------

def intersect_merge(a, b) -> int:
    a, b = _c32(a), _c32(b)
    return int(lib().oracle_intersect_merge(_p(a, _i32p), a.size, _p(b, _i32p), b.size))


def intersect_bsearch(a, b) -> int:
    a, b = _c32(a), _c32(b)
    return int(lib().oracle_intersect_bsearch(_p(a, _i32p), a.size, _p(b, _i32p), b.size))


def binary_search(lst, key) -> bool:
    lst = _c32(lst)
    return bool(lib().oracle_binary_search(_p(lst, _i32p), 0, lst.size, int(key)))


def orient(rowptr, colidx):
    """Degree-ordered DAG of a symmetric CSR (graph.cc:252-298)."""
    rowptr, colidx = _c64(rowptr), _c32(colidx)
    n = rowptr.size - 1
    out_rp = np.empty(n + 1, dtype=np.int64)
    nnz = lib().oracle_orient_rowptr(n, _p(rowptr, _i64p), _p(colidx, _i32p), _p(out_rp, _i64p))
    out_ci = np.empty(max(nnz, 1), dtype=np.int32)
    lib().oracle_orient_fill(n, _p(rowptr, _i64p), _p(colidx, _i32p), _p(out_rp, _i64p),
                             _p(out_ci, _i32p))
    return out_rp, out_ci[:nnz]


def tc_count(rowptr, colidx, row_begin=0, row_end=None, chunk=1024, threads=0, method=0) -> int:
    """u64 count over pivot rows [row_begin, row_end) of a DAG."""
    rowptr, colidx = _c64(rowptr), _c32(colidx)
    n = rowptr.size - 1
    if row_end is None:
        row_end = n
    return int(lib().oracle_tc_count(n, _p(rowptr, _i64p), _p(colidx, _i32p), row_begin, row_end,
                                     chunk, threads, method))


def tc_count_rows(rowptr, colidx, rows, chunk=64, threads=0) -> int:
    """u64 count over an explicit list of pivot rows of a DAG."""
    rowptr, colidx, rows = _c64(rowptr), _c32(colidx), _c32(rows)
    return int(lib().oracle_tc_count_list(_p(rowptr, _i64p), _p(colidx, _i32p), _p(rows, _i32p),
                                          rows.size, chunk, threads))


def tc_count_halo(row_begin, row_end, local_rowptr, local_colidx, halo_ids, halo_rowptr,
                  halo_colidx, threads=0) -> int:
    lr, lc = _c64(local_rowptr), _c32(local_colidx)
    hi, hr, hc = _c32(halo_ids), _c64(halo_rowptr), _c32(halo_colidx)
    missing = ctypes.c_int64(0)
    total = lib().oracle_tc_count_halo(row_begin, row_end, _p(lr, _i64p), _p(lc, _i32p), hi.size,
                                       _p(hi, _i32p), _p(hr, _i64p), _p(hc, _i32p), threads,
                                       ctypes.byref(missing))
    if missing.value:
        raise ValueError(f"halo misses {missing.value} remote lists")
    return int(total)


def tc_per_vertex(rowptr, colidx):
    rowptr, colidx = _c64(rowptr), _c32(colidx)
    n = rowptr.size - 1
    out = np.zeros(n, dtype=np.uint64)
    lib().oracle_tc_per_vertex(n, _p(rowptr, _i64p), _p(colidx, _i32p), _p(out, _u64p))
    return out


def tc_edge_counts(rowptr, colidx):
    """|N+(u) ∩ N+(v)| for every DAG edge (u, v), in CSR order."""
    rowptr, colidx = _c64(rowptr), _c32(colidx)
    out = np.zeros(colidx.size, dtype=np.uint32)
    lib().oracle_tc_edge_counts(rowptr.size - 1, _p(rowptr, _i64p), _p(colidx, _i32p),
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def tc_pushpull_bounds(rowptr, colidx, bounds):
    """Per-PE push/pull sums when PE p owns ids [bounds[p], bounds[p+1])."""
    rowptr, colidx = _c64(rowptr), _c32(colidx)
    b = _c32(bounds)
    out = np.zeros(b.size - 1, dtype=np.uint64)
    lib().oracle_tc_pushpull_bounds(rowptr.size - 1, _p(rowptr, _i64p), _p(colidx, _i32p),
                                    b.size - 1, _p(b, _i32p), _p(out, _u64p))
    return out


def tc_pushpull_pe(rowptr, colidx, k, pe) -> int:
    rowptr, colidx = _c64(rowptr), _c32(colidx)
    return int(lib().oracle_tc_pushpull_pe(rowptr.size - 1, _p(rowptr, _i64p), _p(colidx, _i32p),
                                           k, pe))


# --------------------------------------------------------------------------
# numpy restatement of the seeded generators
# --------------------------------------------------------------------------

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(x):
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def draw(seed, a, b):
    """Counter-based draw b of stream a (vectorised over a and/or b)."""
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    with np.errstate(over="ignore"):
        inner = np.uint64(seed) + np.uint64(0x632BE59BD9B4E019) * (a + np.uint64(1))
        return _mix64(_mix64(inner) + b)


def _unit(r):
    return (np.asarray(r, dtype=np.uint64) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def ceil_log2(n):
    s = 0
    while (1 << s) < n:
        s += 1
    return s


def rmat_pairs(n, edge_factor, a, b, c, seed):
    scale = ceil_log2(n)
    m = _llround(edge_factor * n)
    two32 = 4294967296.0
    ta, tb, tc = (int(math.floor(p * two32)) for p in (a, a + b, a + b + c))
    i = np.arange(m, dtype=np.uint64)
    u = np.zeros(m, dtype=np.uint64)
    v = np.zeros(m, dtype=np.uint64)
    r = None
    for lvl in range(scale):
        if lvl % 2 == 0:
            r = draw(seed, i, lvl >> 1)
        x = (r & np.uint64(0xFFFFFFFF)) if (lvl & 1) else (r >> np.uint64(32))
        x = x.astype(np.int64)
        bu = (x >= tb).astype(np.uint64)            # quadrants c, d -> row bit 1
        bv = (((x >= ta) & (x < tb)) | (x >= tc)).astype(np.uint64)  # b, d -> col bit 1
        u = (u << np.uint64(1)) | bu
        v = (v << np.uint64(1)) | bv
    u = u.astype(np.int64)
    v = v.astype(np.int64)
    keep = (u < n) & (v < n) & (u != v)
    return u[keep], v[keep]


def _llround(x):
    # C llround: half away from zero (inputs here are non-negative).
    return int(math.floor(x + 0.5))


def er_pairs(n, avg_degree, seed):
    m = _llround(n * avg_degree / 2.0)
    i = np.arange(m, dtype=np.uint64)
    r = draw(seed, i, 0)
    u = (((r >> np.uint64(32)) * np.uint64(n)) >> np.uint64(32)).astype(np.int64)
    v = (((r & np.uint64(0xFFFFFFFF)) * np.uint64(n)) >> np.uint64(32)).astype(np.int64)
    keep = u != v
    return u[keep], v[keep]


def symmetric_csr(n, src, dst):
    """Symmetrised, deduplicated, row-sorted CSR (int64 rowptr, int32 colidx)."""
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    keep = src != dst
    s = np.concatenate([src[keep], dst[keep]])
    d = np.concatenate([dst[keep], src[keep]])
    keys = np.unique(s * np.int64(n) + d)
    rows = keys // n
    cols = (keys % n).astype(np.int32)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=rowptr[1:])
    return rowptr, cols


def gen_er(n, avg_degree, seed):
    return symmetric_csr(n, *er_pairs(n, avg_degree, seed))


def gen_rmat(n, edge_factor, seed, a=0.57, b=0.19, c=0.19):
    return symmetric_csr(n, *rmat_pairs(n, edge_factor, a, b, c, seed))


def batch_spec(batch_seed, index):
    """(kind, n, avg_degree, seed) of graph `index` of the batched workload."""
    u_n = float(_unit(draw(batch_seed, index, 1)))
    u_d = float(_unit(draw(batch_seed, index, 2)))
    n = int(math.floor(2.0 ** (8.0 + 8.0 * u_n)))
    n = min(max(n, 256), 65536)
    kind = index & 1
    deg = 4.0 + 60.0 * u_d
    seed = int(draw(batch_seed, index, 3))
    return kind, n, deg, seed


def batch_graph(batch_seed, index):
    kind, n, deg, seed = batch_spec(batch_seed, index)
    rp, ci = gen_er(n, deg, seed) if kind == 0 else gen_rmat(n, deg / 2.0, seed)
    return orient(rp, ci)


def orient_numpy(rowptr, colidx):
    """Pure-numpy orientation (cross-checks the C restatement)."""
    rowptr = np.asarray(rowptr, dtype=np.int64)
    colidx = np.asarray(colidx, dtype=np.int64)
    n = rowptr.size - 1
    deg = np.diff(rowptr)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    keep = (deg[colidx] > deg[src]) | ((deg[colidx] == deg[src]) & (colidx > src))
    out_rp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(src[keep], minlength=n), out=out_rp[1:])
    return out_rp, colidx[keep].astype(np.int32)


def brute_force_triangles(n, src, dst) -> int:
    """O(n^3) count on an adjacency matrix (n <= 64)."""
    assert n <= 64
    A = np.zeros((n, n), dtype=np.int64)
    for a, b in zip(np.asarray(src).tolist(), np.asarray(dst).tolist()):
        if a != b:
            A[a, b] = A[b, a] = 1
    t = 0
    for i in range(n):
        for j in range(i + 1, n):
            if not A[i, j]:
                continue
            for k in range(j + 1, n):
                t += A[i, k] & A[j, k]
    return int(t)


def csr_digest(rowptr, colidx) -> str:
    import hashlib
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(rowptr, dtype="<i8").tobytes())
    h.update(np.ascontiguousarray(colidx, dtype="<i4").tobytes())
    return h.hexdigest()
