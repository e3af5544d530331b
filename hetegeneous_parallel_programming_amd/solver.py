"""Device-side API: the Python face of ``libtc_hip.so`` (C ABI in
``include/tc.h``).

:func:`TCSolver` mirrors the reference's boundary
``void TCSolver(Graph &g, uint64_t &total, int n_gpus, int chunk_size)``
(successful_push_pull/main.cc:5, multigpu_nvshmem_pushpull.cu:29-85) and
returns the u64 total.  Every count runs the hand-written HIP kernels; if
the library is missing the import of this module's entry points raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _native as nat
from ._native import CsrView, PartitionView, TcError, host
from .graph import Graph, HaloPartition

NUM_BINS = 6
BIN_NAMES = ("global", "workgroup", "wave", "g32", "g16", "g8")


class Stats(ctypes.Structure):
    _fields_ = [("total_ms", nat.c_dbl), ("plan_ms", nat.c_dbl), ("count_ms", nat.c_dbl),
                ("pivots", nat.c_i64), ("bin_pivots", nat.c_i64 * NUM_BINS),
                ("hash_fallbacks", nat.c_i64), ("single_launch", nat.c_i64), ("plan_kept", nat.c_i64)]

    def as_dict(self) -> dict:
        return {"total_ms": self.total_ms, "plan_ms": self.plan_ms, "count_ms": self.count_ms,
                "pivots": self.pivots,
                "bin_pivots": dict(zip(BIN_NAMES, list(self.bin_pivots))),
                "hash_fallbacks": self.hash_fallbacks, "single_launch": bool(self.single_launch),
                "plan_kept": bool(self.plan_kept)}


class Tune(ctypes.Structure):
    _fields_ = [("t_g8", nat.c_i32), ("t_g16", nat.c_i32), ("t_g32", nat.c_i32),
                ("t_wave", nat.c_i32), ("rounds", nat.c_i32), ("grid", nat.c_i32),
                ("pivot_set", nat.c_i32), ("hash_kicks", nat.c_i32)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}

    @classmethod
    def from_dict(cls, d: dict) -> "Tune":
        t = cls()
        for f, _ in cls._fields_:
            setattr(t, f, int(d[f]))
        return t


class Local(ctypes.Structure):
    _fields_ = [("edge_counts", ctypes.POINTER(ctypes.c_uint32)), ("vertex_counts", nat.u64p),
                ("parts", nat.c_i32), ("bounds", nat.i32p), ("part_counts", nat.u64p)]


class GraphInfo(ctypes.Structure):
    _fields_ = [("graphs", nat.c_i32), ("rows", nat.c_i64), ("nnz", nat.c_i64),
                ("partner_rows", nat.c_i64), ("build_ms", nat.c_dbl), ("bytes", nat.c_i64)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class AutotuneReport(ctypes.Structure):
    _fields_ = [("trials", nat.c_i32), ("start_ms", nat.c_dbl), ("best_ms", nat.c_dbl),
                ("elapsed_ms", nat.c_dbl)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", nat.c_u64), ("slots", nat.c_i32), ("max_batch_graphs", nat.c_i32),
                ("max_batch_rows", nat.c_i64), ("max_queue", nat.c_i32), ("lookahead", nat.c_i32),
                ("host_staging", nat.c_u64), ("max_batch_ids", nat.c_i64), ("heavy_ids", nat.c_i64),
                ("fill_graphs", nat.c_i32), ("fill_wait_us", nat.c_i32)]


class RequestResult(ctypes.Structure):
    _fields_ = [("tag", nat.c_u64), ("count", nat.c_u64), ("status", nat.c_i32),
                ("batch_graphs", nat.c_i32), ("t_submit_ms", nat.c_dbl), ("t_admit_ms", nat.c_dbl),
                ("t_launch_ms", nat.c_dbl), ("t_done_ms", nat.c_dbl)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", nat.c_i64), ("completed", nat.c_i64), ("rejected", nat.c_i64),
                ("launches", nat.c_i64), ("queued", nat.c_i64), ("in_flight", nat.c_i64),
                ("mean_batch_graphs", nat.c_dbl), ("pool_capacity", nat.c_u64),
                ("pool_used", nat.c_u64), ("pool_peak", nat.c_u64),
                ("submit_ms", nat.c_dbl), ("retire_ms", nat.c_dbl), ("admit_ms", nat.c_dbl),
                ("launch_ms", nat.c_dbl), ("wait_busy_ms", nat.c_dbl), ("wait_idle_ms", nat.c_dbl),
                ("device_ms", nat.c_dbl), ("staging_used", nat.c_u64)]


_dev = None


def device_lib():
    """Load libtc_hip.so; raises RuntimeError when it has not been built."""
    global _dev
    if _dev is not None:
        return _dev
    host()  # dependency, loaded RTLD_GLOBAL first
    if not os.path.exists(nat.HIP_LIB):
        raise RuntimeError(f"{nat.HIP_LIB} is missing: the HIP extension must be built "
                           "(python -m hetegeneous_parallel_programming_amd.build)")
    L = ctypes.CDLL(nat.HIP_LIB)
    P, PP = ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)
    S = nat._sig
    S(L, "tc_ctx_create", nat.c_i32, [nat.c_i32, PP])
    S(L, "tc_ctx_destroy", None, [P])
    S(L, "tc_last_error", ctypes.c_char_p, [P])
    S(L, "tc_get_tune", nat.c_i32, [P, ctypes.POINTER(Tune)])
    S(L, "tc_set_tune", nat.c_i32, [P, ctypes.POINTER(Tune)])
    S(L, "tc_device_name", nat.c_i32, [P, ctypes.c_char_p, nat.c_i32])
    S(L, "tc_graph_upload", nat.c_i32, [P, ctypes.POINTER(CsrView), nat.c_u32, PP])
    S(L, "tc_graph_upload_rows", nat.c_i32, [P, ctypes.POINTER(CsrView), ctypes.POINTER(ctypes.c_uint8),
                                             nat.c_u32, PP])
    S(L, "tc_partition_upload", nat.c_i32, [P, ctypes.POINTER(PartitionView), nat.c_u32, PP])
    S(L, "tc_graph_free", nat.c_i32, [P, P])
    S(L, "tc_graph_bytes", nat.c_i64, [P])
    S(L, "tc_count", nat.c_i32, [P, P, nat.u64p, ctypes.POINTER(Stats)])
    S(L, "tc_count_csr", nat.c_i32, [P, ctypes.POINTER(CsrView), nat.c_u32, nat.u64p, ctypes.POINTER(Stats)])
    S(L, "tc_count_plain", nat.c_i32, [P, P, nat.u64p, ctypes.POINTER(Stats)])
    S(L, "tc_count_binned", nat.c_i32, [P, P, nat.u64p, ctypes.POINTER(Stats)])
    S(L, "tc_count_noalloc", nat.c_i32, [P, P, nat.u64p, ctypes.POINTER(Stats)])
    S(L, "tc_count_local", nat.c_i32, [P, P, nat.u64p, ctypes.POINTER(Local), ctypes.POINTER(Stats)])
    S(L, "tc_batch_create", nat.c_i32, [P, ctypes.POINTER(CsrView), nat.c_i32, nat.c_u32, PP])
    S(L, "tc_batch_count", nat.c_i32, [P, P, nat.u64p, ctypes.POINTER(Stats)])
    S(L, "tc_batch_count_each", nat.c_i32, [P, P, nat.u64p, ctypes.POINTER(nat.c_dbl),
                                            ctypes.POINTER(Stats)])
    S(L, "tc_batch_free", nat.c_i32, [P, P])
    S(L, "tc_batch_bytes", nat.c_i64, [P])
    S(L, "tc_batch_graphs", nat.c_i32, [P])
    S(L, "tc_graph_get_info", nat.c_i32, [P, ctypes.POINTER(GraphInfo)])
    S(L, "tc_batch_get_info", nat.c_i32, [P, ctypes.POINTER(GraphInfo)])
    S(L, "tc_autotune", nat.c_i32, [P, P, P, nat.c_i32, nat.c_dbl, ctypes.POINTER(Tune),
                                    ctypes.POINTER(AutotuneReport)])
    S(L, "tc_tune_save", nat.c_i32, [P, ctypes.c_char_p, ctypes.c_char_p])
    S(L, "tc_tune_load", nat.c_i32, [P, ctypes.c_char_p, ctypes.c_char_p])
    S(L, "tc_probe_copy", nat.c_i32, [P, nat.c_i64, nat.c_i32, ctypes.POINTER(nat.c_dbl)])
    S(L, "tc_probe_read", nat.c_i32, [P, nat.c_i64, nat.c_i32, ctypes.POINTER(nat.c_dbl)])
    S(L, "tc_clock_ms", nat.c_dbl, [])
    S(L, "tc_request_footprint", nat.c_u64, [nat.c_i32, nat.c_i64])
    S(L, "tc_server_create", nat.c_i32, [P, ctypes.POINTER(ServerConfig), PP])
    S(L, "tc_server_submit", nat.c_i32, [P, ctypes.POINTER(CsrView), nat.c_u64, nat.c_u32])
    S(L, "tc_server_submit_at", nat.c_i32, [P, ctypes.POINTER(CsrView), nat.c_u64, nat.c_u32, nat.c_dbl])
    S(L, "tc_server_poll", nat.c_i32, [P, ctypes.POINTER(RequestResult), nat.c_i32,
                                       ctypes.POINTER(nat.c_i32), nat.c_i32])
    S(L, "tc_server_drain", nat.c_i32, [P, nat.c_i32])
    S(L, "tc_server_get_stats", nat.c_i32, [P, ctypes.POINTER(ServerStats)])
    S(L, "tc_server_destroy", nat.c_i32, [P])
    _dev = L
    return L


def _check(status: int) -> None:
    if status != nat.TC_OK:
        raise TcError(status, device_lib().tc_last_error(None).decode())


class DeviceGraph:
    """A DAG (or a partition with its halo cache) resident in HBM."""

    def __init__(self, ctx: "Context", handle: int, rows: int, edges: int):
        self.ctx = ctx
        self._h = ctypes.c_void_p(handle)
        self.rows = rows
        self.edges = edges

    @property
    def bytes(self) -> int:
        return int(device_lib().tc_graph_bytes(self._h))

    def info(self) -> dict:
        """What the upload built (partner rows, build time, device bytes)."""
        gi = GraphInfo()
        _check(device_lib().tc_graph_get_info(self._h, ctypes.byref(gi)))
        return gi.as_dict()

    def free(self) -> None:
        if self._h is not None and self._h.value and self.ctx._h is not None:
            _check(device_lib().tc_graph_free(self.ctx._h, self._h))
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Many DAGs resident at once; counted by one captured persistent launch."""

    def __init__(self, ctx: "Context", handle: int, n: int):
        self.ctx = ctx
        self._h = ctypes.c_void_p(handle)
        self.n = n

    def count(self):
        out = np.zeros(self.n, dtype=np.uint64)
        st = Stats()
        _check(device_lib().tc_batch_count(self.ctx._h, self._h,
                                           out.ctypes.data_as(nat.u64p), ctypes.byref(st)))
        return out, st.as_dict()

    def count_each(self):
        """Baseline: one plain launch + copy back + sync per graph."""
        out = np.zeros(self.n, dtype=np.uint64)
        ms = np.zeros(self.n, dtype=np.float64)
        st = Stats()
        _check(device_lib().tc_batch_count_each(self.ctx._h, self._h,
                                                out.ctypes.data_as(nat.u64p),
                                                ms.ctypes.data_as(ctypes.POINTER(nat.c_dbl)),
                                                ctypes.byref(st)))
        return out, ms, st.as_dict()

    @property
    def bytes(self) -> int:
        return int(device_lib().tc_batch_bytes(self._h))

    def info(self) -> dict:
        gi = GraphInfo()
        _check(device_lib().tc_batch_get_info(self._h, ctypes.byref(gi)))
        return gi.as_dict()

    def free(self) -> None:
        if self._h is not None and self._h.value and self.ctx._h is not None:
            _check(device_lib().tc_batch_free(self.ctx._h, self._h))
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One device, one HIP stream, the tuning parameters and plan scratch."""

    def __init__(self, device: int = 0):
        L = device_lib()
        h = ctypes.c_void_p()
        _check(L.tc_ctx_create(device, ctypes.byref(h)))
        self._h = h
        self.device = device

    @property
    def name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        _check(device_lib().tc_device_name(self._h, buf, 256))
        return buf.value.decode()

    @property
    def tune(self) -> dict:
        t = Tune()
        _check(device_lib().tc_get_tune(self._h, ctypes.byref(t)))
        return t.as_dict()

    @tune.setter
    def tune(self, d: dict) -> None:
        cur = self.tune
        cur.update(d)
        t = Tune.from_dict(cur)
        _check(device_lib().tc_set_tune(self._h, ctypes.byref(t)))

    def upload(self, g: Graph, check: int = nat.TC_CHECK_SHAPE, owned=None) -> DeviceGraph:
        """Make ``g`` resident; ``owned`` (one byte per row) restricts the
        counted rows to those marked 1 (``tc_graph_upload_rows``)."""
        h = ctypes.c_void_p()
        if owned is None:
            _check(device_lib().tc_graph_upload(self._h, ctypes.byref(g.view()), check,
                                                ctypes.byref(h)))
        else:
            own = np.ascontiguousarray(owned, dtype=np.uint8)
            if own.size != g.V():
                raise ValueError("owned needs one entry per row")
            _check(device_lib().tc_graph_upload_rows(
                self._h, ctypes.byref(g.view()), own.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                check, ctypes.byref(h)))
        return DeviceGraph(self, h.value, g.V(), g.E())

    def upload_arrays(self, rowptr, colidx, check: int = nat.TC_CHECK_SHAPE) -> DeviceGraph:
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        ci = np.ascontiguousarray(colidx, dtype=np.int32)
        v = CsrView(rp.ctypes.data_as(nat.i64p), ci.ctypes.data_as(nat.i32p), rp.size - 1, ci.size)
        h = ctypes.c_void_p()
        _check(device_lib().tc_graph_upload(self._h, ctypes.byref(v), check, ctypes.byref(h)))
        return DeviceGraph(self, h.value, rp.size - 1, ci.size)

    def upload_partition(self, p: HaloPartition, check: int = nat.TC_CHECK_SHAPE) -> DeviceGraph:
        h = ctypes.c_void_p()
        _check(device_lib().tc_partition_upload(self._h, ctypes.byref(p.view()), check,
                                                ctypes.byref(h)))
        return DeviceGraph(self, h.value, p.row_end - p.row_begin, int(p.local_colidx.size))

    def count(self, dg: DeviceGraph):
        """tc_count: the binned count, or the single plain launch for small graphs."""
        total = nat.c_u64(0)
        st = Stats()
        _check(device_lib().tc_count(self._h, dg._h, ctypes.byref(total), ctypes.byref(st)))
        return int(total.value), st.as_dict()

    def count_csr(self, g: Graph, check: int = nat.TC_CHECK_SHAPE):
        """tc_count_csr: one count of the host DAG g, nothing kept on the
        device (a small graph is counted without a partner structure)."""
        total = nat.c_u64(0)
        st = Stats()
        _check(device_lib().tc_count_csr(self._h, ctypes.byref(g.view()), check, ctypes.byref(total),
                                         ctypes.byref(st)))
        return int(total.value), st.as_dict()

    def count_binned(self, dg: DeviceGraph):
        """The degree-binned count whatever the graph's size (tc_count_binned)."""
        total = nat.c_u64(0)
        st = Stats()
        _check(device_lib().tc_count_binned(self._h, dg._h, ctypes.byref(total), ctypes.byref(st)))
        return int(total.value), st.as_dict()

    def count_noalloc(self, dg: DeviceGraph):
        """tc_count's count without any device allocation (tc_count_noalloc):
        its plan lives in the library's static scratch, at most 2^21 plan
        slots; larger graphs raise TC_ERANGE."""
        total = nat.c_u64(0)
        st = Stats()
        _check(device_lib().tc_count_noalloc(self._h, dg._h, ctypes.byref(total), ctypes.byref(st)))
        return int(total.value), st.as_dict()

    def count_plain(self, dg: DeviceGraph):
        total = nat.c_u64(0)
        st = Stats()
        _check(device_lib().tc_count_plain(self._h, dg._h, ctypes.byref(total), ctypes.byref(st)))
        return int(total.value), st.as_dict()

    def count_local(self, dg: DeviceGraph, edges: bool = True, vertices: bool = True,
                    bounds=None):
        """Count with per-edge, per-vertex and per-part attribution (``tc_count_local``).

        Returns (total, dict) with ``edge_counts`` (uint32 per DAG edge of the
        counted rows, CSR order), ``vertex_counts`` (uint64 per row of the
        uploaded graph), ``part_counts`` (uint64 per part of the split
        ``bounds``, the reference push/pull kernel's per-PE sums) and
        ``stats``; outputs not asked for are None.  Without ``vertices`` the
        kernels attribute each triangle to its edge only (same edge and part
        counts, less work); with ``bounds`` alone (no edges, no vertices)
        each partner row's triangles go to its owner's part (same part
        counts, the plain count's kernels)."""
        n = int(dg.info()["rows"])
        lo = Local()
        ec = vc = pc = None
        if edges:
            ec = np.zeros(dg.edges, np.uint32)
            lo.edge_counts = ec.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        if vertices:
            vc = np.zeros(n, np.uint64)
            lo.vertex_counts = vc.ctypes.data_as(nat.u64p)
        if bounds is not None:
            b = np.ascontiguousarray(bounds, dtype=np.int32)
            pc = np.zeros(b.size - 1, np.uint64)
            lo.parts = b.size - 1
            lo.bounds = b.ctypes.data_as(nat.i32p)
            lo.part_counts = pc.ctypes.data_as(nat.u64p)
        total = nat.c_u64(0)
        st = Stats()
        _check(device_lib().tc_count_local(self._h, dg._h, ctypes.byref(total), ctypes.byref(lo),
                                           ctypes.byref(st)))
        return int(total.value), {"edge_counts": ec, "vertex_counts": vc, "part_counts": pc,
                                  "stats": st.as_dict()}

    def batch(self, graphs: list, check: int = nat.TC_CHECK_SHAPE) -> Batch:
        views = (CsrView * max(len(graphs), 1))()
        for i, g in enumerate(graphs):
            views[i] = g.view()
        h = ctypes.c_void_p()
        _check(device_lib().tc_batch_create(self._h, views, len(graphs), check, ctypes.byref(h)))
        return Batch(self, h.value, len(graphs))

    def autotune(self, graph: DeviceGraph | None = None, batch: Batch | None = None,
                 reps: int = 3, budget_ms: float = 10000.0):
        """Pick the launch setting (bin thresholds = group widths, tile
        pivots, grid) by timing exactly one workload; installs it on this
        context and returns (tune dict, report dict)."""
        if (graph is None) == (batch is None):
            raise ValueError("give exactly one of graph= or batch=")
        t, rep = Tune(), AutotuneReport()
        _check(device_lib().tc_autotune(self._h, graph._h if graph is not None else None,
                                        batch._h if batch is not None else None, reps,
                                        float(budget_ms), ctypes.byref(t), ctypes.byref(rep)))
        return t.as_dict(), rep.as_dict()

    def save_tune(self, path: str, key: str) -> None:
        """Write this context's tune under (device name, key) into a tune file."""
        _check(device_lib().tc_tune_save(self._h, os.fsencode(path), key.encode()))

    def load_tune(self, path: str, key: str) -> bool:
        """Install the tune saved under (device name, key); False if none."""
        st = device_lib().tc_tune_load(self._h, os.fsencode(path), key.encode())
        if st == nat.TC_EIO:
            return False
        _check(st)
        return True

    def probe_copy(self, nbytes: int, iters: int = 10) -> float:
        """Mean ms of a plain device copy of nbytes (counter calibration)."""
        ms = nat.c_dbl(0)
        _check(device_lib().tc_probe_copy(self._h, nbytes, iters, ctypes.byref(ms)))
        return ms.value

    def probe_read(self, nbytes: int, iters: int = 10) -> float:
        """Mean ms of a plain streaming read of nbytes (practical read rate)."""
        ms = nat.c_dbl(0)
        _check(device_lib().tc_probe_read(self._h, nbytes, iters, ctypes.byref(ms)))
        return ms.value

    def close(self) -> None:
        if self._h is not None and self._h.value:
            device_lib().tc_ctx_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def clock_ms() -> float:
    """The serving loop's monotonic clock (ms)."""
    return float(device_lib().tc_clock_ms())


def request_footprint(n_rows: int, nnz: int) -> int:
    return int(device_lib().tc_request_footprint(n_rows, nnz))


class Server:
    """Continuous-batching TC server on one device (``tc_server_*``)."""

    def __init__(self, ctx: Context, hbm_budget: int = 8 << 30, slots: int = 3,
                 max_batch_graphs: int = 1024, max_batch_rows: int = 8 << 20,
                 max_queue: int = 1 << 16, lookahead: int = 64, host_staging: int = 1 << 30,
                 max_batch_ids: int = 0, heavy_ids: int = 0, fill_graphs: int = 0, fill_wait_us: int = 0):
        cfg = ServerConfig(hbm_budget, slots, max_batch_graphs, max_batch_rows, max_queue, lookahead,
                           host_staging, max_batch_ids, heavy_ids, fill_graphs, fill_wait_us)
        h = ctypes.c_void_p()
        _check(device_lib().tc_server_create(ctx._h, ctypes.byref(cfg), ctypes.byref(h)))
        self.ctx = ctx
        self._h = h
        self._buf = (RequestResult * 4096)()

    def submit(self, g: Graph, tag: int, check: int = nat.TC_CHECK_SHAPE) -> None:
        _check(device_lib().tc_server_submit(self._h, ctypes.byref(g.view()), tag, check))

    def submit_arrays(self, rowptr, colidx, tag: int, check: int = nat.TC_CHECK_SHAPE) -> None:
        """Submit a DAG given as CSR arrays (copied before this returns)."""
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        ci = np.ascontiguousarray(colidx, dtype=np.int32)
        v = CsrView(rp.ctypes.data_as(nat.i64p), ci.ctypes.data_as(nat.i32p), rp.size - 1, ci.size)
        _check(device_lib().tc_server_submit(self._h, ctypes.byref(v), tag, check))

    def try_submit(self, g: Graph, tag: int, check: int = nat.TC_CHECK_SHAPE) -> int:
        """Status code instead of an exception (TC_EBUSY when the queue is full)."""
        return int(device_lib().tc_server_submit(self._h, ctypes.byref(g.view()), tag, check))

    def submit_at(self, g: Graph, tag: int, t_ms: float, check: int = nat.TC_CHECK_SHAPE) -> None:
        """Submit at a scheduled time of clock_ms() (waits without holding the GIL)."""
        _check(device_lib().tc_server_submit_at(self._h, ctypes.byref(g.view()), tag, check, float(t_ms)))

    def poll(self, timeout_ms: int = 0, max_results: int = 4096) -> list[dict]:
        n = nat.c_i32(0)
        _check(device_lib().tc_server_poll(self._h, self._buf, min(max_results, 4096),
                                           ctypes.byref(n), timeout_ms))
        out = []
        for i in range(n.value):
            r = self._buf[i]
            out.append({"tag": r.tag, "count": r.count, "status": r.status,
                        "batch_graphs": r.batch_graphs, "t_submit_ms": r.t_submit_ms,
                        "t_admit_ms": r.t_admit_ms, "t_launch_ms": r.t_launch_ms,
                        "t_done_ms": r.t_done_ms})
        return out

    def drain(self, timeout_ms: int = 600_000) -> None:
        _check(device_lib().tc_server_drain(self._h, timeout_ms))

    def stats(self) -> dict:
        st = ServerStats()
        _check(device_lib().tc_server_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self) -> None:
        if self._h is not None and self._h.value:
            _check(device_lib().tc_server_destroy(self._h))
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def TCSolver(g: Graph, n_gpus: int = 1, chunk_size: int = 1024, ctx: Context | None = None,
             stats: dict | None = None) -> int:
    """Count the triangles of DAG `g` on the GPU; returns the u64 total.

    Mirrors ``TCSolver(Graph&, uint64_t& total, int n_gpus, int chunk_size)``
    (successful_push_pull/main.cc:5).  Only one GPU is driven here: n_gpus
    other than 1 raises.  chunk_size is accepted for signature parity; the
    reference's GPU solvers ignore it too (main.cc:25,28).  `stats`, if
    given, receives the count's statistics (launch form, plan and count ms).
    """
    if n_gpus != 1:
        raise TcError(nat.TC_EINVAL, "this library drives exactly one GPU (n_gpus must be 1)")
    own = ctx is None
    c = Context(0) if own else ctx
    try:
        total, st = c.count_csr(g)
        if stats is not None:
            stats.update(st)
        return total
    finally:
        if own:
            c.close()
