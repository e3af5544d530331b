"""CPU tests of the product's host substrate (libtc_host.so) against the
oracle: generators, orientation, partitioner + halo cache, GraphMiner I/O,
validation and the work model."""
import json
import os

import numpy as np
import pytest

import oracle
from hetegeneous_parallel_programming_amd import Graph, PartitionedGraph, batch_spec
from hetegeneous_parallel_programming_amd._native import TC_EINVAL, TC_EIO, TC_ERANGE, TcError

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("n,d,seed", [(300, 6.0, 11), (1000, 8, 1), (5000, 12, 3), (16384, 16, 2),
                                      (777, 33.3, 9)])
def test_er_generator_bit_exact(n, d, seed):
    g = Graph.er(n, d, seed)
    rp, ci = oracle.gen_er(n, d, seed)
    assert np.array_equal(g.out_rowptr(), rp) and np.array_equal(g.out_colidx(), ci)


@pytest.mark.parametrize("n,ef,seed", [(256, 8, 5), (1000, 16, 3), (4096, 16, 2), (12000, 10, 6),
                                       (65536, 4, 1)])
def test_rmat_generator_bit_exact(n, ef, seed):
    g = Graph.rmat(n, ef, seed)
    rp, ci = oracle.gen_rmat(n, ef, seed)
    assert np.array_equal(g.out_rowptr(), rp) and np.array_equal(g.out_colidx(), ci)


def test_generators_match_golden_digests():
    gold = json.load(open(os.path.join(GOLDEN, "small.json")))
    for c in gold["cases"]:
        g = (Graph.er if c["kind"] == "er" else Graph.rmat)(c["n"], c["param"], c["seed"])
        assert oracle.csr_digest(g.out_rowptr(), g.out_colidx()) == c["sha256_sym"], c


def test_generated_graphs_are_simple_and_symmetric():
    for g in (Graph.er(3000, 20, 4), Graph.rmat(4096, 16, 8)):
        g.validate()  # shape, strictly increasing rows, no self-loops
        rp, ci = g.out_rowptr(), g.out_colidx()
        src = np.repeat(np.arange(g.V()), np.diff(rp))
        fwd = set(zip(src.tolist(), ci.tolist()))
        assert all((b, a) in fwd for a, b in list(fwd)[:20000])


def test_orientation_matches_oracle():
    for g in (Graph.er(16384, 16, 1), Graph.rmat(1 << 15, 16, 3)):
        d = g.oriented()
        rp, ci = oracle.orient(g.out_rowptr(), g.out_colidx())
        assert np.array_equal(d.out_rowptr(), rp) and np.array_equal(d.out_colidx(), ci)
        assert d.get_max_degree() == int(np.diff(rp).max())
        g2 = Graph.from_csr(g.out_rowptr(), g.out_colidx())
        g2.orientation()  # in place, as the reference's Graph::orientation
        assert np.array_equal(g2.out_colidx(), ci)


def test_batch_spec_and_graphs_match_oracle():
    gold = json.load(open(os.path.join(GOLDEN, "small.json")))
    for spec in gold["batch_specs"]:
        s = batch_spec(1, spec["index"])
        assert (s["n"], s["seed"]) == (spec["n"], spec["seed"])
        assert s["kind"] == ("er" if spec["kind"] == 0 else "rmat")
        assert s["avg_degree"] == spec["avg_degree"]
    many = Graph.batch(1, 0, 12)
    for i in (0, 1, 2, 5, 11):
        one = Graph.batch_member(1, i)
        rp, ci = oracle.batch_graph(1, i)
        assert np.array_equal(one.out_rowptr(), rp) and np.array_equal(one.out_colidx(), ci)
        assert np.array_equal(many[i].out_colidx(), ci)


def test_from_pairs_and_sort():
    g = Graph.from_pairs(5, [0, 1, 2, 2, 4, 3], [1, 0, 3, 3, 4, 0])  # dup, self-loop, reverse
    assert g.out_rowptr().tolist() == [0, 2, 3, 4, 6, 6]
    assert g.out_colidx().tolist() == [1, 3, 0, 3, 0, 2]
    with pytest.raises(TcError) as e:
        Graph.from_pairs(3, [0], [3])
    assert e.value.status == TC_EINVAL
    rp = np.array([0, 3, 3], np.int64)
    h = Graph.from_csr(rp, np.array([1, 0, 1], np.int32))
    h.sort_neighbors()
    assert h.out_colidx().tolist() == [0, 1, 1]


@pytest.mark.parametrize("rp,ci,check", [
    ([0, 2, 3], [1, 5, 0], 1),       # id out of range
    ([1, 2, 3], [1, 0, 0], 1),       # rowptr[0] != 0
    ([0, 3, 2], [1, 1, 0], 1),       # non-monotone
    ([0, 2, 4], [1, 0, 0, 1], 2),    # unsorted row (1, 0)
    ([0, 2, 3], [1, 1, 0], 2),       # duplicate neighbour
    ([0, 1, 2], [0, 0], 4),          # self-loop
])
def test_validation_rejects(rp, ci, check):
    with pytest.raises(TcError) as e:
        Graph.from_csr(np.array(rp, np.int64), np.array(ci, np.int32), check=check | 1)
    assert e.value.status == TC_EINVAL


def test_work_model_matches_numpy():
    d = Graph.rmat(1 << 14, 16, 5).oriented()
    rp, ci = d.out_rowptr(), d.out_colidx()
    deg = np.diff(rp)
    src = np.repeat(np.arange(d.V()), deg)
    w = d.work()
    assert w["edges"] == ci.size
    assert w["wedge_elems"] == int(deg[ci].sum())
    assert w["min_elems"] == int(np.minimum(deg[src], deg[ci]).sum())
    assert w["alg_bytes"] == 8 * (d.V() + 1) + 20 * ci.size + 4 * int(deg[ci].sum())
    lo, hi = 100, 5000
    sel = (src >= lo) & (src < hi)
    assert d.work(lo, hi)["wedge_elems"] == int(deg[ci[sel]].sum())


def _partner_model(rp, ci, lo, hi, per_row=1024):
    """Partner scheme restated in numpy: edge (u, v) is owned by u when
    d+(v) <= d+(u), else by v; an owner's edges form ceil(c / per_row) units."""
    deg = np.diff(rp).astype(np.int64)
    src = np.repeat(np.arange(deg.size), deg)
    sel = (src >= lo) & (src < hi)
    s, t = src[sel], ci[sel].astype(np.int64)
    du, dv = deg[s], deg[t]
    owner = np.where(dv <= du, s, t)
    cnt = np.bincount(owner, minlength=deg.size)
    units = -(-cnt // per_row)
    byts = 32 * units.sum() + 4 * (units * deg).sum() + 8 * s.size + 4 * np.minimum(du, dv).sum()
    return int(units.sum()), int(byts), int(cnt.max(initial=0))


@pytest.mark.parametrize("lo,hi", [(0, None), (1000, 30000), (0, 1)])
def test_partner_work_model_matches_numpy(lo, hi):
    d = Graph.rmat(1 << 16, 16, 5).oriented()
    rp, ci = d.out_rowptr(), d.out_colidx()
    hi = d.V() if hi is None else hi
    units, byts, cmax = _partner_model(rp, ci, lo, hi)
    w = d.work(lo, hi)
    assert (w["partner_rows"], w["partner_bytes"]) == (units, byts)
    if (lo, hi) == (0, d.V()):
        assert cmax > 1024  # some owners are cut into several units
        # every edge streams the shorter list: never more than the reference's
        assert w["min_elems"] <= w["wedge_elems"]


def test_partner_work_model_of_partition_matches_rows_of_dag():
    d = Graph.rmat(1 << 15, 16, 3).oriented()
    pg = PartitionedGraph(d, 3)
    for p in range(3):
        hp = pg.halo_partition(p)
        a, b = hp.work(), d.work(hp.row_begin, hp.row_end)
        assert (a["partner_rows"], a["partner_bytes"]) == (b["partner_rows"], b["partner_bytes"])


@pytest.mark.parametrize("k", [1, 2, 3, 4, 7])
def test_halo_partition_counts_sum_to_total(k):
    d = Graph.rmat(1 << 14, 16, 7).oriented()
    rp, ci = d.out_rowptr(), d.out_colidx()
    total = oracle.tc_count(rp, ci)
    pg = PartitionedGraph(d, k)
    chunk = -(-d.V() // k)
    acc = 0
    for p in range(k):
        hp = pg.halo_partition(p)
        b, e = hp.row_begin, hp.row_end
        assert (b, e) == (min(p * chunk, d.V()), min((p + 1) * chunk, d.V()))
        assert np.array_equal(hp.local_colidx, ci[rp[b]:rp[e]])
        assert np.array_equal(hp.local_rowptr, rp[b:e + 1] - rp[b])
        remote = np.unique(hp.local_colidx[(hp.local_colidx < b) | (hp.local_colidx >= e)])
        assert np.array_equal(hp.halo_ids, remote)  # exactly the lists the owned rows reach
        for i in range(0, remote.size, max(1, remote.size // 50)):
            v = remote[i]
            assert np.array_equal(hp.halo_colidx[hp.halo_rowptr[i]:hp.halo_rowptr[i + 1]],
                                  ci[rp[v]:rp[v + 1]])
        c = oracle.tc_count_halo(b, e, hp.local_rowptr, hp.local_colidx, hp.halo_ids,
                                 hp.halo_rowptr, hp.halo_colidx)
        assert c == oracle.tc_count(rp, ci, b, e)
        w = hp.work()
        assert w["wedge_elems"] == d.work(b, e)["wedge_elems"]
        acc += c
    assert acc == total


def _row_weights(rp, ci, balance):
    """Per-row weights of each split criterion, restated from their definitions."""
    deg = np.diff(rp).astype(np.int64)
    u = np.repeat(np.arange(rp.size - 1), deg)
    dv = deg[ci]
    if balance == "vertices":
        return np.ones(deg.size, np.int64)
    if balance == "edges":
        return deg
    if balance == "wedges":
        return deg + np.bincount(u, weights=dv, minlength=deg.size).astype(np.int64)
    # partner: streamed ids + 6 per edge + 1503 per partner row, shared by
    # the owner's edges; fixed point x 2^16, summed per row exactly
    owner = np.where(dv > deg[u], ci, u)
    oc = np.bincount(owner, minlength=deg.size).astype(np.int64)[owner]
    per_edge = ((np.minimum(deg[u], dv) + 6) << 16) + (1503 << 16) * ((oc + 1023) // 1024) // oc
    pre = np.concatenate([[0], np.cumsum(per_edge, dtype=np.int64)])
    return pre[rp[1:]] - pre[rp[:-1]]


def test_partner_weights_share_each_owners_rows():
    """Each owner's partner rows, shared over the edges it owns, add up to
    the partner rows of the library's work model (Graph.work, the same count
    the device build makes), up to the fixed point's floor per edge."""
    d = Graph.rmat(1 << 14, 16, 5).oriented()
    rp, ci = d.out_rowptr(), d.out_colidx()
    deg = np.diff(rp).astype(np.int64)
    u = np.repeat(np.arange(rp.size - 1), deg)
    dv = deg[ci]
    base = ((np.minimum(deg[u], dv) + 6) << 16).sum()
    rows_q = int(_row_weights(rp, ci, "partner").sum() - base)
    want = d.work()["partner_rows"] * (1503 << 16)
    assert want - ci.size <= rows_q <= want


@pytest.mark.parametrize("balance", ["vertices", "edges", "wedges", "partner"])
@pytest.mark.parametrize("k", [1, 2, 4, 5])
def test_balanced_bounds_split_the_weight_evenly(balance, k):
    d = Graph.rmat(1 << 14, 16, 5).oriented()
    rp, ci = d.out_rowptr(), d.out_colidx()
    b = PartitionedGraph(d, k, balance).bounds()
    n = d.V()
    assert b[0] == 0 and b[-1] == n and (np.diff(b) >= 0).all()
    if balance == "vertices":
        chunk = -(-n // k)
        assert list(b) == [min(p * chunk, n) for p in range(k + 1)]
        return
    w = _row_weights(rp, ci, balance)
    pre = np.concatenate([[0], np.cumsum(w)])
    total = pre[-1]
    for p in range(1, k):
        # split at the row whose prefix is nearest to p/k of the total
        t = total * p // k
        best = np.abs(pre - t).min()
        assert abs(int(pre[b[p]]) - int(t)) == best
    parts = np.diff(pre[b])
    assert parts.max() <= total / k + w.max()


@pytest.mark.parametrize("balance", ["edges", "wedges", "partner"])
def test_balanced_halo_partitions_sum_to_total(balance):
    d = Graph.rmat(1 << 14, 16, 7).oriented()
    rp, ci = d.out_rowptr(), d.out_colidx()
    pg = PartitionedGraph(d, 4, balance)
    b = pg.bounds()
    acc = 0
    for p in range(4):
        hp = pg.halo_partition(p)
        assert (hp.row_begin, hp.row_end) == (b[p], b[p + 1])
        assert np.array_equal(hp.local_colidx, ci[rp[b[p]]:rp[b[p + 1]]])
        c = oracle.tc_count_halo(hp.row_begin, hp.row_end, hp.local_rowptr, hp.local_colidx,
                                 hp.halo_ids, hp.halo_rowptr, hp.halo_colidx)
        assert c == oracle.tc_count(rp, ci, b[p], b[p + 1])
        acc += c
    assert acc == oracle.tc_count(rp, ci)
    # the contiguous vertex split leaves part 0 with far more than a quarter of the edges
    w0 = d.work(0, -(-d.V() // 4))["edges"]
    wb = max(d.work(int(b[p]), int(b[p + 1]))["edges"] for p in range(4))
    if balance == "edges":
        assert wb < w0


def test_partition_span_rejects_bad_ranges():
    from hetegeneous_parallel_programming_amd._native import host
    import ctypes
    d = Graph.er(1000, 8, 1).oriented()
    out = ctypes.c_void_p()
    for b, e in [(-1, 10), (10, 5), (0, 1001)]:
        assert host().tch_partition_span(ctypes.byref(d.view()), b, e, ctypes.byref(out)) == TC_EINVAL
    with pytest.raises(ValueError):
        PartitionedGraph(d, 2, "nodes")


def test_induced_partition_matches_reference_semantics():
    d = Graph.rmat(1 << 13, 16, 2).oriented()
    rp, ci = d.out_rowptr(), d.out_colidx()
    pg = PartitionedGraph(d, 3)
    pg.edgecut_induced_partition1D()
    assert pg.get_num_subgraphs() == 3
    acc = 0
    for p in range(3):
        sub = pg.get_subgraph(p)
        m = pg.idx_map[p]
        assert (np.diff(m) > 0).all()
        lb, le = pg.get_local_begin(p), pg.get_local_end(p)
        b, e = m[lb], m[le - 1] + 1
        c = oracle.tc_count(sub.out_rowptr(), sub.out_colidx(), lb, le)
        assert c == oracle.tc_count(rp, ci, b, e)
        acc += c
    assert acc == oracle.tc_count(rp, ci)


def _reference_metis_files(tmp_path, rp, ci, part, k):
    """The file set graph_partition.cc:501-566 writes for each PE, built here
    with numpy: meta lines n, n_real, E, max_subg_nv, max_subg_ne,
    max_degree; int64 rowptr over all ids with only the PE's rows filled;
    int32 colidx; int32 owned ids ascending; int64 part array."""
    n = rp.size - 1
    deg = np.diff(rp)
    nreal = [int((part == i).sum()) for i in range(k)]
    ne = [int(deg[part == i].sum()) for i in range(k)]
    files = {}
    for i in range(k):
        mine = part == i
        sub_rp = np.concatenate([[0], np.cumsum(np.where(mine, deg, 0))]).astype(np.int64)
        sub_ci = np.concatenate([ci[rp[v]:rp[v + 1]] for v in np.flatnonzero(mine)] or
                                [np.zeros(0, np.int32)]).astype(np.int32)
        meta = f"{n}\n{nreal[i]}\n{ne[i]}\n{max(nreal)}\n{max(ne)}\n{int(deg.max())}\n"
        files[i] = {".meta.txt": meta.encode(), ".vertex.bin": sub_rp.tobytes(),
                    ".edge.bin": sub_ci.tobytes(),
                    ".vertex_list.bin": np.flatnonzero(mine).astype(np.int32).tobytes(),
                    ".part.bin": part.astype(np.int64).tobytes()}
    return files


def test_metis_files_match_the_reference_layout(tmp_path):
    d = Graph.rmat(1 << 12, 8, 3).oriented()
    rp, ci = d.out_rowptr(), d.out_colidx()
    k = 3
    part = (np.arange(d.V(), dtype=np.int64) * 7919) % k   # scattered, as a METIS part is
    prefix = str(tmp_path / "pgraph")
    d.write_metis(prefix, part)
    want = _reference_metis_files(tmp_path, rp, ci, part, k)
    for i in range(k):
        for ext, data in want[i].items():
            assert open(f"{prefix}{i}{ext}", "rb").read() == data, (i, ext)


def test_metis_files_read_assemble_and_count(tmp_path):
    from hetegeneous_parallel_programming_amd import MetisPart, assemble_metis

    d = Graph.rmat(1 << 13, 16, 4).oriented()
    rp, ci = d.out_rowptr(), d.out_colidx()
    k = 4
    part = (np.arange(d.V(), dtype=np.int64) * 104729 + 17) % k
    prefix = str(tmp_path / "pgraph")
    # files written by the reference's layout, read by the product
    for i, files in _reference_metis_files(tmp_path, rp, ci, part, k).items():
        for ext, data in files.items():
            open(f"{prefix}{i}{ext}", "wb").write(data)
    pes = MetisPart.read_all(prefix, k)
    total = oracle.tc_count(rp, ci)
    acc = 0
    for i, pe in enumerate(pes):
        assert (pe.n, pe.n_real) == (d.V(), int((part == i).sum()))
        assert np.array_equal(pe.vertex_list, np.flatnonzero(part == i))
        assert np.array_equal(pe.part, part)
        g, owned = assemble_metis(pes, i)
        assert np.array_equal(owned, (part == i).astype(np.uint8))
        grp, gci = g.out_rowptr(), g.out_colidx()
        # own rows as written, remote lists they reach as in the whole DAG, the rest empty
        reach = np.zeros(d.V(), bool)
        reach[part == i] = True
        for u in np.flatnonzero(part == i):
            reach[ci[rp[u]:rp[u + 1]]] = True
        assert np.array_equal(np.diff(grp), np.where(reach, np.diff(rp), 0))
        assert np.array_equal(gci, np.concatenate([ci[rp[v]:rp[v + 1]] for v in np.flatnonzero(reach)]))
        c = oracle.tc_count_rows(grp, gci, pe.vertex_list)
        assert c == oracle.tc_count_rows(rp, ci, pe.vertex_list)
        acc += c
    assert acc == total
    whole, owned = assemble_metis(pes, -1)
    assert owned.all()
    assert np.array_equal(whole.out_rowptr(), rp) and np.array_equal(whole.out_colidx(), ci)


def test_metis_read_rejects_bad_files(tmp_path):
    from hetegeneous_parallel_programming_amd import MetisPart

    d = Graph.er(500, 6, 1).oriented()
    prefix = str(tmp_path / "pgraph")
    part = np.arange(d.V(), dtype=np.int64) % 2
    d.write_metis(prefix, part)
    with pytest.raises(TcError) as e:
        MetisPart(prefix, 2)          # no such PE
    assert e.value.status == TC_EIO
    vl = np.fromfile(f"{prefix}0.vertex_list.bin", np.int32)
    vl[0] = 1                         # an id PE 1 owns
    vl.tofile(f"{prefix}0.vertex_list.bin")
    with pytest.raises(TcError) as e:
        MetisPart(prefix, 0)
    assert e.value.status == TC_EINVAL
    with pytest.raises(TcError):
        d.write_metis(prefix, np.full(d.V(), -1, np.int64))


def test_graphminer_roundtrip(tmp_path):
    g = Graph.rmat(4096, 8, 3)
    prefix = str(tmp_path / "g")
    g.write(prefix)
    meta = open(prefix + ".meta.txt").read().split()
    assert meta[:4] == [str(g.V()), str(g.E()), "4", "8"]
    h = Graph.from_file(prefix)
    assert np.array_equal(h.out_rowptr(), g.out_rowptr())
    assert np.array_equal(h.out_colidx(), g.out_colidx())
    dag = Graph.from_file(prefix, use_dag=True)  # graph.cc:135-138
    ref = g.oriented()  # keep the graph alive: out_colidx() views its memory
    assert np.array_equal(dag.out_colidx(), ref.out_colidx())
    with pytest.raises(TcError) as e:
        Graph.from_file(str(tmp_path / "missing"))
    assert e.value.status == TC_EIO
    open(prefix + ".edge.bin", "ab").write(b"\0\0\0\0")  # wrong size
    with pytest.raises(TcError) as e:
        Graph.from_file(prefix)
    assert e.value.status == TC_EIO


def test_readers_check_file_sizes_before_allocating(tmp_path):
    # counts in the meta data far beyond the files (and beyond any memory) are
    # rejected from the files' sizes; nothing is allocated from them
    from hetegeneous_parallel_programming_amd import MetisPart

    g = Graph.er(64, 4.0, 1)
    prefix = str(tmp_path / "g")
    g.write(prefix)
    meta = open(prefix + ".meta.txt").read().split()
    meta[1] = str(1 << 60)
    open(prefix + ".meta.txt", "w").write("\n".join(meta) + "\n")
    with pytest.raises(TcError) as e:
        Graph.from_file(prefix)
    assert e.value.status == TC_EIO
    d = g.oriented()
    pre = str(tmp_path / "pg")
    d.write_metis(pre, np.zeros(d.V(), np.int64))
    assert MetisPart(pre, 0).n_real == d.V()
    for field, value, status in [(2, 1 << 60, TC_EIO), (1, 1 << 40, TC_ERANGE), (0, 1 << 40, TC_ERANGE)]:
        meta = open(pre + "0.meta.txt").read().split()
        good = list(meta)
        meta[field] = str(value)
        open(pre + "0.meta.txt", "w").write("\n".join(meta) + "\n")
        with pytest.raises(TcError) as e:
            MetisPart(pre, 0)
        assert e.value.status == status
        open(pre + "0.meta.txt", "w").write("\n".join(good) + "\n")


@pytest.mark.parametrize("balance", ["vertices", "edges", "wedges", "partner"])
def test_balanced_bounds_with_more_parts_than_rows(balance):
    d = Graph.er(5, 2.0, 3).oriented()
    b = PartitionedGraph(d, 9, balance).bounds()
    assert b[0] == 0 and b[-1] == d.V() and (np.diff(b) >= 0).all() and b.size == 10
    empty = Graph.from_pairs(4, np.zeros(0, np.int32), np.zeros(0, np.int32)).oriented()
    b = PartitionedGraph(empty, 3, balance).bounds()
    assert b[0] == 0 and b[-1] == 4 and (np.diff(b) >= 0).all()


def test_metis_files_with_a_pe_that_owns_nothing(tmp_path):
    from hetegeneous_parallel_programming_amd import MetisPart, assemble_metis

    d = Graph.rmat(1 << 10, 8, 9).oriented()
    rp, ci = d.out_rowptr(), d.out_colidx()
    part = np.where(np.arange(d.V()) % 2 == 0, 0, 2).astype(np.int64)   # PE 1 owns no vertex
    prefix = str(tmp_path / "pgraph")
    d.write_metis(prefix, part)
    pes = MetisPart.read_all(prefix, 3)
    assert pes[1].n_real == 0 and pes[1].nnz == 0
    acc = 0
    for i in range(3):
        g, owned = assemble_metis(pes, i)
        acc += oracle.tc_count_rows(g.out_rowptr(), g.out_colidx(), np.flatnonzero(owned))
    assert acc == oracle.tc_count(rp, ci)
