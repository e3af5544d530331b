"""GPU parity of tc_count_local: per-edge counts |N+(u) ∩ N+(v)|, per-vertex
triangle counts and the reference push/pull per-PE attribution, against
the oracle (bit-exact), on whole graphs and on halo partitions; and the
edges-only kernels (no vertex_counts asked for) against the same."""
import itertools
import math

import numpy as np
import pytest

import oracle
from hetegeneous_parallel_programming_amd import Graph, PartitionedGraph
from hetegeneous_parallel_programming_amd._native import TcError

pytestmark = pytest.mark.gpu

# the six-vertex graph of tests/test_oracle.py whose split {0} | {1..5} pushes an edge
PUSH_EDGES = [(0, 5), (3, 4), (0, 2), (0, 1), (2, 3), (1, 5), (2, 4), (0, 4), (2, 5), (4, 5)]


def _cliques(sizes):
    src, dst, base = [], [], 0
    for k in sizes:
        s, d = zip(*itertools.combinations(range(base, base + k), 2))
        src += s
        dst += d
        base += k
    return Graph.from_pairs(base, np.array(src, np.int32), np.array(dst, np.int32)).oriented()


def _check_whole(ctx, g, bounds=None):
    rp, ci = g.out_rowptr(), g.out_colidx()
    dg = ctx.upload(g)
    total, out = ctx.count_local(dg, bounds=bounds)
    assert total == oracle.tc_count(rp, ci)
    assert np.array_equal(out["edge_counts"], oracle.tc_edge_counts(rp, ci))
    assert np.array_equal(out["vertex_counts"], oracle.tc_per_vertex(rp, ci))
    if bounds is not None:
        assert np.array_equal(out["part_counts"], oracle.tc_pushpull_bounds(rp, ci, bounds))
    # without vertex_counts the kernels attribute to edges only: same edge and part counts
    total_e, out_e = ctx.count_local(dg, vertices=False, bounds=bounds)
    assert total_e == total and out_e["vertex_counts"] is None
    assert np.array_equal(out_e["edge_counts"], out["edge_counts"])
    if bounds is not None:
        assert np.array_equal(out_e["part_counts"], out["part_counts"])
        # part sums alone: each partner row's hits go to its owner's part, with
        # no per-edge counts; the same sums and their total
        total_p, out_p = ctx.count_local(dg, edges=False, vertices=False, bounds=bounds)
        assert total_p == total and out_p["edge_counts"] is None and out_p["vertex_counts"] is None
        assert np.array_equal(out_p["part_counts"], out["part_counts"])
    # the plain count still agrees after the local call rebuilt the partner entries
    assert ctx.count_binned(dg)[0] == total
    dg.free()
    return out


def test_hand_example_with_a_pushed_edge(ctx):
    s, d = zip(*PUSH_EDGES)
    g = Graph.from_pairs(6, np.array(s, np.int32), np.array(d, np.int32)).oriented()
    out = _check_whole(ctx, g, bounds=[0, 1, 6])
    assert out["edge_counts"].tolist() == [2, 1, 0, 1, 0, 1, 0, 1, 0, 0]
    assert out["vertex_counts"].tolist() == [4, 1, 4, 1, 4, 4]
    assert out["part_counts"].tolist() == [4, 2]


@pytest.mark.parametrize("n,ef,seed", [(1 << 12, 8, 1), (1 << 15, 16, 2), (1 << 16, 16, 3)])
def test_rmat_local_counts(ctx, n, ef, seed):
    g = Graph.rmat(n, ef, seed).oriented()
    k = 4
    chunk = -(-g.V() // k)
    _check_whole(ctx, g, bounds=[min(p * chunk, g.V()) for p in range(k + 1)])


def test_er_local_counts_balanced_split(ctx):
    g = Graph.er(16384, 16, 1).oriented()
    _check_whole(ctx, g, bounds=PartitionedGraph(g, 5, "partner").bounds())


def test_cliques_in_every_bin(ctx):
    # 4200 puts a pivot list past the workgroup's LDS (HBM set, global bin)
    sizes = [3, 5, 9, 17, 33, 65, 129, 260, 700, 1500, 4200]
    g = _cliques(sizes)
    bounds = np.concatenate([[0], np.cumsum(sizes)])
    dg = ctx.upload(g)
    total, out = ctx.count_local(dg, bounds=bounds)
    st = out["stats"]
    assert st["bin_pivots"]["global"] > 0 and st["bin_pivots"]["g8"] > 0
    assert total == sum(math.comb(k, 3) for k in sizes)
    # each vertex of K_k sits in C(k-1, 2) triangles; parts are the cliques
    want_v = np.concatenate([np.full(k, math.comb(k - 1, 2), np.uint64) for k in sizes])
    assert np.array_equal(out["vertex_counts"], want_v)
    assert out["part_counts"].tolist() == [math.comb(k, 3) for k in sizes]
    rp, ci = g.out_rowptr(), g.out_colidx()
    assert np.array_equal(out["edge_counts"], oracle.tc_edge_counts(rp, ci))
    dg.free()


@pytest.mark.parametrize("layout", [{"pivot_set": 1}, {"pivot_set": 0, "hash_kicks": 1},
                                    {"t_g8": 0, "t_g16": 0, "t_g32": 0, "t_wave": 0},
                                    {"t_wave": 1 << 20, "hash_kicks": 128}])
def test_layouts_same_local_counts(ctx, layout):
    g = Graph.rmat(1 << 15, 16, 4).oriented()
    saved = ctx.tune
    try:
        ctx.tune = layout
        _check_whole(ctx, g)
    finally:
        ctx.tune = saved


def test_partitions_add_up_to_the_whole(ctx):
    g = Graph.rmat(1 << 16, 16, 5).oriented()
    rp, ci = g.out_rowptr(), g.out_colidx()
    want_e = oracle.tc_edge_counts(rp, ci)
    want_v = oracle.tc_per_vertex(rp, ci)
    acc = np.zeros(g.V(), np.uint64)
    for balance in ("vertices", "partner"):
        acc[:] = 0
        pg = PartitionedGraph(g, 3, balance)
        bounds = pg.bounds()
        parts_acc = np.zeros(3, np.uint64)
        for p in range(3):
            part = pg.halo_partition(p)
            dg = ctx.upload_partition(part)
            # part sums alone first (no local rebuild yet): an edge's owner may
            # be a halo vertex, whose degree comes from the halo cache
            total_p, out_p = ctx.count_local(dg, edges=False, vertices=False, bounds=bounds)
            total, out = ctx.count_local(dg, bounds=bounds)
            b, e = part.row_begin, part.row_end
            assert np.array_equal(out["edge_counts"], want_e[rp[b]:rp[e]])
            assert total == int(out["edge_counts"].sum())
            assert total_p == total and np.array_equal(out_p["part_counts"], out["part_counts"])
            parts_acc += out_p["part_counts"]
            acc += out["vertex_counts"]
            dg.free()
        assert np.array_equal(acc, want_v)
        # over the partitions, each PE's sum is the whole graph's push/pull sum
        assert np.array_equal(parts_acc, oracle.tc_pushpull_bounds(rp, ci, bounds))


def test_repeated_calls_and_partial_outputs(ctx):
    g = Graph.rmat(1 << 14, 16, 6).oriented()
    rp, ci = g.out_rowptr(), g.out_colidx()
    dg = ctx.upload(g)
    first = ctx.count_local(dg)[1]
    total, out = ctx.count_local(dg, edges=False, vertices=False)
    assert out["edge_counts"] is None and out["vertex_counts"] is None
    assert total == oracle.tc_count(rp, ci)
    again = ctx.count_local(dg)[1]
    assert np.array_equal(first["edge_counts"], again["edge_counts"])
    assert np.array_equal(first["vertex_counts"], again["vertex_counts"])
    assert dg.info()["bytes"] > 12 * g.E()  # local buffers are reported
    dg.free()


def test_rejects_bad_split(ctx):
    g = Graph.er(1000, 8, 1).oriented()
    dg = ctx.upload(g)
    for bad in ([0, 500], [1, 500, 1000], [0, 600, 500, 1000]):
        with pytest.raises(TcError):
            ctx.count_local(dg, bounds=bad)
    with pytest.raises(TcError):
        ctx.count_local(dg, bounds=np.arange(0, 1026, dtype=np.int32))  # 1025 parts
    dg.free()


def test_parts_with_saturated_degrees(ctx):
    """A pushed edge between two rows whose out-degrees both pass the 16-bit
    saturation: the per-part pass (and, for part sums alone, the owner rule
    of the partner build) compares their exact degrees.  Vertex 0
    points at vertex 1 and 70,000 leaves, vertex 1 at 70,010 leaves, 50,000 of
    them shared; with 0 and 1 in different parts, d+(0) < d+(1) pushes the
    edge (0, 1) and its 50,000 triangles to vertex 1's part."""
    a = np.arange(2, 70_002, dtype=np.int32)
    b = np.concatenate([a[:50_000], np.arange(70_002, 90_012, dtype=np.int32)])
    n = 90_012
    rp = np.zeros(n + 1, np.int64)
    rp[1] = 1 + a.size
    rp[2:] = rp[1] + b.size
    ci = np.concatenate([np.array([1], np.int32), a, b])
    bounds = [0, 1, n]
    want = oracle.tc_pushpull_bounds(rp, ci, bounds)
    assert want.tolist() == [0, 50_000]
    dg = ctx.upload_arrays(rp, ci)
    for vertices in (True, False):
        total, out = ctx.count_local(dg, vertices=vertices, bounds=bounds)
        assert total == 50_000
        assert np.array_equal(out["part_counts"], want)
        assert np.array_equal(out["edge_counts"], oracle.tc_edge_counts(rp, ci))
    # part sums alone: the edge's owner (vertex 1, by exact degrees) takes it
    total, out = ctx.count_local(dg, edges=False, vertices=False, bounds=bounds)
    assert total == 50_000 and np.array_equal(out["part_counts"], want)
    dg.free()


def test_part_sums_alone_on_a_fresh_upload(ctx):
    """Part sums alone need only the partner rows' owners, which every upload
    keeps: the first such call adds just the split points and sums (no
    per-edge buffers, no rebuild), matches the oracle's push/pull sums, and
    agrees with a full call made afterwards; an owned-row mask too."""
    g = Graph.rmat(1 << 15, 16, 8).oriented()
    rp, ci = g.out_rowptr(), g.out_colidx()
    dg = ctx.upload(g)
    before = dg.info()["bytes"]
    for balance in ("vertices", "partner"):
        bounds = PartitionedGraph(g, 4, balance).bounds()
        total, out = ctx.count_local(dg, edges=False, vertices=False, bounds=bounds)
        assert total == oracle.tc_count(rp, ci)
        assert np.array_equal(out["part_counts"], oracle.tc_pushpull_bounds(rp, ci, bounds))
    assert dg.info()["bytes"] - before < 16 * 1024
    total_f, out_f = ctx.count_local(dg, bounds=bounds)  # rebuilds the partner entries
    total_p, out_p = ctx.count_local(dg, edges=False, vertices=False, bounds=bounds)
    assert total_p == total_f and np.array_equal(out_p["part_counts"], out_f["part_counts"])
    dg.free()
    odd = (np.arange(g.V()) % 2).astype(np.uint8)
    dg = ctx.upload(g, owned=odd)
    total_p, out_p = ctx.count_local(dg, edges=False, vertices=False, bounds=bounds)
    total_f, out_f = ctx.count_local(dg, bounds=bounds)
    assert total_p == total_f == oracle.tc_count_rows(rp, ci, np.flatnonzero(odd))
    assert np.array_equal(out_p["part_counts"], out_f["part_counts"])
    dg.free()
