"""Host-side graph objects: the Python face of ``libtc_host.so``.

Mirrors the reference's host classes on the triangle-counting path:

* :class:`Graph` -- ``class Graph`` (successful_push_pull/graph.h:14-188):
  GraphMiner loader (graph.cc:4-142), ``orientation()`` (graph.cc:252-298),
  ``sort_neighbors()`` (graph.cc:156-164), ``V()/E()/N(v)/out_degree(v)/
  out_rowptr()/out_colidx()/get_max_degree()``.
* :class:`PartitionedGraph` -- ``class PartitionedGraph``
  (graph_partition.h:11-68): ``edgecut_induced_partition1D()``
  (graph_partition.cc:85-135) plus :meth:`halo_partition`, the one-GPU share
  of the push/pull scheme with its HBM halo cache.

Arrays returned by :meth:`Graph.out_rowptr` / :meth:`Graph.out_colidx` are
zero-copy numpy views of native memory owned by the Graph.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from ._native import CsrView, MetisView, PartitionView, Work, host, host_check


def _as_view(rowptr: np.ndarray, colidx: np.ndarray) -> CsrView:
    v = CsrView()
    v.rowptr = rowptr.ctypes.data_as(nat.i64p)
    v.colidx = colidx.ctypes.data_as(nat.i32p)
    v.n_rows = rowptr.size - 1
    v.nnz = int(colidx.size)
    return v


class Graph:
    """An owned host CSR (int64 offsets, int32 ids)."""

    def __init__(self, handle: int, name: str = ""):
        if not handle:
            raise ValueError("null graph handle")
        self._h = ctypes.c_void_p(handle)
        self.name = name
        self._refresh()

    # ---- construction ----------------------------------------------------
    @classmethod
    def _wrap(cls, fn, *args, name=""):
        out = ctypes.c_void_p()
        host_check(fn(*args, ctypes.byref(out)))
        return cls(out.value, name)

    @classmethod
    def from_file(cls, prefix: str, use_dag: bool = False) -> "Graph":
        """Load ``prefix.{meta.txt,vertex.bin,edge.bin}`` (graph.cc:31-55)."""
        return cls._wrap(host().tch_read_graphminer, prefix.encode(), int(use_dag), name=prefix)

    @classmethod
    def er(cls, n: int, avg_degree: float, seed: int) -> "Graph":
        return cls._wrap(host().tch_gen_er, n, float(avg_degree), seed,
                         name=f"er-n{n}-d{avg_degree:g}-s{seed}")

    @classmethod
    def rmat(cls, n: int, edge_factor: float, seed: int, a: float = 0.57, b: float = 0.19,
             c: float = 0.19) -> "Graph":
        return cls._wrap(host().tch_gen_rmat, n, float(edge_factor), a, b, c, seed,
                         name=f"rmat-n{n}-ef{edge_factor:g}-s{seed}")

    @classmethod
    def rmat_scale(cls, scale: int, edge_factor: float, seed: int) -> "Graph":
        return cls.rmat(1 << scale, edge_factor, seed)

    @classmethod
    def from_pairs(cls, n: int, src, dst) -> "Graph":
        s = np.ascontiguousarray(src, dtype=np.int32)
        d = np.ascontiguousarray(dst, dtype=np.int32)
        if s.size != d.size:
            raise ValueError("src and dst differ in length")
        return cls._wrap(host().tch_csr_from_pairs, n, s.ctypes.data_as(nat.i32p),
                         d.ctypes.data_as(nat.i32p), s.size)

    @classmethod
    def from_csr(cls, rowptr, colidx, check: int = nat.TC_CHECK_SHAPE) -> "Graph":
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        ci = np.ascontiguousarray(colidx, dtype=np.int32)
        v = _as_view(rp, ci)
        return cls._wrap(host().tch_csr_copy, ctypes.byref(v), check)

    @classmethod
    def batch_member(cls, batch_seed: int, index: int) -> "Graph":
        """DAG of graph `index` of the seeded batched workload."""
        return cls._wrap(host().tch_batch_graph, batch_seed, index, name=f"batch{batch_seed}-{index}")

    @staticmethod
    def batch(batch_seed: int, first: int, count: int) -> list["Graph"]:
        """DAGs of graphs [first, first+count), generated in parallel."""
        arr = (ctypes.c_void_p * max(count, 1))()
        host_check(host().tch_batch_graphs(batch_seed, first, count, arr))
        return [Graph(arr[i], f"batch{batch_seed}-{first + i}") for i in range(count)]

    # ---- lifetime --------------------------------------------------------
    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                host().tch_csr_free(h)
            except Exception:
                pass
            self._h = None

    def _refresh(self) -> None:
        v = CsrView()
        host_check(host().tch_csr_view(self._h, ctypes.byref(v)))
        self._view = v
        n, nnz = v.n_rows, v.nnz
        self._rowptr = np.ctypeslib.as_array(v.rowptr, shape=(n + 1,))
        self._colidx = (np.ctypeslib.as_array(v.colidx, shape=(nnz,)) if nnz > 0
                        else np.zeros(0, dtype=np.int32))

    def view(self) -> CsrView:
        return self._view

    # ---- reference accessors (graph.h:75-93) -----------------------------
    def V(self) -> int:
        return int(self._view.n_rows)

    def E(self) -> int:
        return int(self._view.nnz)

    num_vertices = V
    num_edges = E

    def get_max_degree(self) -> int:
        return int(host().tch_csr_max_degree(self._h))

    def out_degree(self, v: int) -> int:
        return int(self._rowptr[v + 1] - self._rowptr[v])

    get_degree = out_degree

    def N(self, v: int) -> np.ndarray:
        return self._colidx[self._rowptr[v]:self._rowptr[v + 1]]

    def out_rowptr(self) -> np.ndarray:
        return self._rowptr

    def out_colidx(self) -> np.ndarray:
        return self._colidx

    def print_meta_data(self) -> None:
        print(f"|V|: {self.V()}, |E|: {self.E()}, Max Degree: {self.get_max_degree()}")

    # ---- transformations --------------------------------------------------
    def orientation(self) -> "Graph":
        """Orient in place into the degree-ordered DAG (graph.cc:252-298)."""
        out = ctypes.c_void_p()
        host_check(host().tch_orient(ctypes.byref(self._view), ctypes.byref(out)))
        old = self._h
        self._h = out
        self._refresh()
        host().tch_csr_free(old)
        return self

    def oriented(self) -> "Graph":
        """A new DAG; this graph is left unchanged."""
        return Graph._wrap(host().tch_orient, ctypes.byref(self._view), name=self.name + "-dag")

    def sort_neighbors(self) -> None:
        host_check(host().tch_sort_neighbors(self._h))

    def validate(self, check: int = nat.TC_CHECK_ALL) -> None:
        host_check(host().tch_validate(ctypes.byref(self._view), check))

    def work(self, row_begin: int = 0, row_end: int | None = None) -> dict:
        """Algorithmic work of counting rows [row_begin, row_end) of this DAG."""
        w = Work()
        host_check(host().tch_work_of(ctypes.byref(self._view), row_begin,
                                      self.V() if row_end is None else row_end, ctypes.byref(w)))
        return {f: getattr(w, f) for f, _ in Work._fields_}

    def write(self, prefix: str) -> None:
        host_check(host().tch_write_graphminer(prefix.encode(), ctypes.byref(self._view)))

    def write_metis(self, prefix: str, part) -> None:
        """The reference's per-PE part files ``<prefix>{i}.*`` for every PE i
        of the part array (graph_partition.cc:501-566; ``tch_metis_write``)."""
        p = np.ascontiguousarray(part, dtype=np.int64)
        if p.size != self.V():
            raise ValueError("part must give every vertex a PE")
        k = int(p.max()) + 1 if p.size else 1
        host_check(host().tch_metis_write(prefix.encode(), ctypes.byref(self._view),
                                          p.ctypes.data_as(nat.i64p), k))


class HaloPartition:
    """Owned rows of one contiguous-range part of a DAG plus its halo cache."""

    def __init__(self, handle: int):
        self._h = ctypes.c_void_p(handle)
        v = PartitionView()
        host_check(host().tch_partition_get(self._h, ctypes.byref(v)))
        self._view = v
        nl = v.row_end - v.row_begin
        self.n_global = v.n_global
        self.row_begin, self.row_end = v.row_begin, v.row_end
        self.local_rowptr = np.ctypeslib.as_array(v.local_rowptr, shape=(nl + 1,))
        self.local_colidx = (np.ctypeslib.as_array(v.local_colidx, shape=(v.local_nnz,))
                             if v.local_nnz else np.zeros(0, np.int32))
        self.halo_ids = (np.ctypeslib.as_array(v.halo_ids, shape=(v.n_halo,))
                         if v.n_halo else np.zeros(0, np.int32))
        self.halo_rowptr = np.ctypeslib.as_array(v.halo_rowptr, shape=(v.n_halo + 1,))
        self.halo_colidx = (np.ctypeslib.as_array(v.halo_colidx, shape=(v.halo_nnz,))
                            if v.halo_nnz else np.zeros(0, np.int32))

    def view(self) -> PartitionView:
        return self._view

    def work(self) -> dict:
        w = Work()
        host_check(host().tch_work_of_partition(ctypes.byref(self._view), ctypes.byref(w)))
        return {f: getattr(w, f) for f, _ in Work._fields_}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                host().tch_partition_free(h)
            except Exception:
                pass
            self._h = None


class MetisPart:
    """One PE's METIS part files (graph.cc:16-23,33-34,56-59; ``tch_metis_read``)."""

    def __init__(self, prefix: str, i: int):
        h = ctypes.c_void_p()
        host_check(host().tch_metis_read(prefix.encode(), i, ctypes.byref(h)))
        self._h = h
        v = MetisView()
        host_check(host().tch_metis_get(h, ctypes.byref(v)))
        self.pe = i
        self.n, self.n_real, self.nnz = v.n, v.n_real, v.nnz
        self.max_subg_nv, self.max_subg_ne, self.max_degree = v.max_subg_nv, v.max_subg_ne, v.max_degree
        self.rowptr = np.ctypeslib.as_array(v.rowptr, shape=(v.n + 1,))
        self.colidx = (np.ctypeslib.as_array(v.colidx, shape=(v.nnz,)) if v.nnz
                       else np.zeros(0, np.int32))
        self.vertex_list = (np.ctypeslib.as_array(v.vertex_list, shape=(v.n_real,)) if v.n_real
                            else np.zeros(0, np.int32))
        self.part = np.ctypeslib.as_array(v.part, shape=(v.n,))

    @staticmethod
    def read_all(prefix: str, k: int) -> list:
        return [MetisPart(prefix, i) for i in range(k)]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                host().tch_metis_free(h)
            except Exception:
                pass
            self._h = None


def assemble_metis(parts: list, i: int):
    """(Graph, owned) for PE ``i`` of the METIS part files ``parts``: its own
    rows plus the remote lists they reach, with the owned-row mask of its
    pivots; ``i = -1`` gives the whole DAG (``tch_metis_assemble``)."""
    k = len(parts)
    arr = (ctypes.c_void_p * k)(*[p._h.value for p in parts])
    out = ctypes.c_void_p()
    owned = np.zeros(parts[0].n, np.uint8)
    host_check(host().tch_metis_assemble(ctypes.cast(arr, ctypes.POINTER(ctypes.c_void_p)), k, i,
                                         ctypes.byref(out),
                                         owned.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
    return Graph(out.value, f"metis-pe{i}"), owned


class PartitionedGraph:
    """Contiguous-range partitioner (graph_partition.h:11-68).

    ``balance`` picks the split points of :meth:`halo_partition`: "vertices"
    is the reference's ceil(n/k) split (graph_partition.cc:88-107); "edges",
    "wedges" and "partner" balance d+(u), the edge-iterator elements or the
    partner-scheme cost of the owned rows (streamed ids plus per-edge and
    per-partner-row costs measured on MI355X; ``tch_partition_bounds``).  The
    vertex-induced subgraphs of :meth:`edgecut_induced_partition1D` always
    use the reference split."""

    def __init__(self, g: Graph, nc: int, balance: str = "vertices"):
        if nc < 1:
            raise ValueError("need at least one vertex chunk")
        if balance not in nat.BALANCE:
            raise ValueError(f"balance must be one of {sorted(nat.BALANCE)}")
        self.g = g
        self.num_vertex_chunks = nc
        self.balance = balance
        self._bounds: np.ndarray | None = None
        self._subgraphs: list[Graph] = []
        self._local_begin: list[int] = []
        self._local_end: list[int] = []
        self.idx_map: list[np.ndarray] = []

    def edgecut_induced_partition1D(self) -> None:
        """Vertex-induced subgraph per range (graph_partition.cc:85-135)."""
        self._subgraphs, self._local_begin, self._local_end, self.idx_map = [], [], [], []
        for i in range(self.num_vertex_chunks):
            sub = ctypes.c_void_p()
            lb, le = nat.c_i32(0), nat.c_i32(0)
            ids = nat.i32p()
            host_check(host().tch_partition_induced(ctypes.byref(self.g.view()),
                                                    self.num_vertex_chunks, i, ctypes.byref(sub),
                                                    ctypes.byref(lb), ctypes.byref(le),
                                                    ctypes.byref(ids)))
            g = Graph(sub.value, f"{self.g.name}-part{i}")
            m = np.ctypeslib.as_array(ids, shape=(max(g.V(), 1),))[:g.V()].copy()
            host().tch_free_ids(ids)
            self._subgraphs.append(g)
            self._local_begin.append(lb.value)
            self._local_end.append(le.value)
            self.idx_map.append(m)

    def get_num_subgraphs(self) -> int:
        return len(self._subgraphs)

    def get_subgraph(self, i: int) -> Graph:
        return self._subgraphs[i]

    def get_local_begin(self, i: int) -> int:
        return self._local_begin[i]

    def get_local_end(self, i: int) -> int:
        return self._local_end[i]

    def bounds(self) -> np.ndarray:
        """Split points b[0..k]: part i owns global ids [b[i], b[i+1])."""
        if self._bounds is None:
            b = np.zeros(self.num_vertex_chunks + 1, np.int32)
            host_check(host().tch_partition_bounds(
                ctypes.byref(self.g.view()), self.num_vertex_chunks, nat.BALANCE[self.balance],
                b.ctypes.data_as(nat.i32p)))
            self._bounds = b
        return self._bounds

    def halo_partition(self, i: int) -> HaloPartition:
        """Owned rows of part ``i`` plus the halo cache of the remote lists they reach."""
        if not 0 <= i < self.num_vertex_chunks:
            raise IndexError(f"part {i} outside [0, {self.num_vertex_chunks})")
        out = ctypes.c_void_p()
        if self.balance == "vertices":
            host_check(host().tch_partition_range(ctypes.byref(self.g.view()),
                                                  self.num_vertex_chunks, i, ctypes.byref(out)))
        else:
            b = self.bounds()
            host_check(host().tch_partition_span(ctypes.byref(self.g.view()), int(b[i]),
                                                 int(b[i + 1]), ctypes.byref(out)))
        return HaloPartition(out.value)


class Arena:
    """First-fit extent bookkeeping (``tch_arena_*``) used for the server's HBM pool."""

    def __init__(self, capacity: int, align: int = 256):
        h = ctypes.c_void_p()
        host_check(host().tch_arena_create(capacity, align, ctypes.byref(h)))
        self._h = h

    def alloc(self, nbytes: int) -> int:
        off = nat.c_u64(0)
        host_check(host().tch_arena_alloc(self._h, nbytes, ctypes.byref(off)))
        return int(off.value)

    def free(self, offset: int) -> None:
        host_check(host().tch_arena_free(self._h, offset))

    def stats(self) -> dict:
        st = nat.ArenaStats()
        host_check(host().tch_arena_stats_get(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in nat.ArenaStats._fields_}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                host().tch_arena_destroy(h)
            except Exception:
                pass
            self._h = None


def batch_spec(batch_seed: int, index: int) -> dict:
    s = nat.BatchSpec()
    host_check(host().tch_batch_spec_of(batch_seed, index, ctypes.byref(s)))
    return {"kind": "er" if s.kind == 0 else "rmat", "n": s.n, "avg_degree": s.avg_degree,
            "seed": s.seed}
