"""GPU parity of the batched persistent kernel and the partition (halo) count."""
import numpy as np
import pytest

import oracle
from hetegeneous_parallel_programming_amd import Graph, PartitionedGraph

pytestmark = pytest.mark.gpu


def test_batch_matches_oracle(ctx):
    graphs = Graph.batch(1, 0, 96)
    want = np.array([oracle.tc_count(g.out_rowptr(), g.out_colidx()) for g in graphs], np.uint64)
    b = ctx.batch(graphs)
    for rep in range(3):  # replays of the captured graph
        got, st = b.count()
        assert np.array_equal(got, want), (rep, np.nonzero(got != want))
    each, ms, _ = b.count_each()
    assert np.array_equal(each, want)
    assert (ms > 0).all()
    b.free()


def test_batch_partner_rows_match_host_model(ctx):
    graphs = Graph.batch(2, 0, 40)
    b = ctx.batch(graphs)
    info = b.info()
    assert info["graphs"] == 40
    assert info["partner_rows"] == sum(g.work()["partner_rows"] for g in graphs)
    assert info["nnz"] == sum(g.E() for g in graphs)
    b.free()


def test_batch_with_empty_and_tiny_graphs(ctx):
    gs = [Graph.from_pairs(1, np.zeros(0, np.int32), np.zeros(0, np.int32)).oriented(),
          Graph.from_pairs(3, np.array([0, 1, 2], np.int32), np.array([1, 2, 0], np.int32)).oriented(),
          Graph.rmat(4096, 16, 5).oriented(),
          Graph.from_pairs(10, np.zeros(0, np.int32), np.zeros(0, np.int32)).oriented(),
          Graph.er(3000, 30, 8).oriented()]
    want = np.array([oracle.tc_count(g.out_rowptr(), g.out_colidx()) for g in gs], np.uint64)
    b = ctx.batch(gs)
    got, _ = b.count()
    assert np.array_equal(got, want)
    assert np.array_equal(b.count_each()[0], want)
    b.free()


@pytest.mark.parametrize("seg_ids", ["1", "40000"])
def test_batch_built_in_segments(ctx, monkeypatch, seg_ids):
    """A batch's partner build goes by segments of whole graphs (each with
    descriptors rebased to its first graph); a small segment size makes many
    segments, down to one graph each, and the counts and partner rows must
    equal the one-segment build's and the oracle's."""
    graphs = Graph.batch(5, 0, 64) + [Graph.from_pairs(7, np.zeros(0, np.int32), np.zeros(0, np.int32)).oriented()]
    want = np.array([oracle.tc_count(g.out_rowptr(), g.out_colidx()) for g in graphs], np.uint64)
    whole = ctx.batch(graphs)
    monkeypatch.setenv("TC_BATCH_SEGMENT_IDS", seg_ids)
    seg = ctx.batch(graphs)
    monkeypatch.delenv("TC_BATCH_SEGMENT_IDS")
    assert seg.info()["partner_rows"] == whole.info()["partner_rows"]
    for b in (whole, seg):
        got, _ = b.count()
        assert np.array_equal(got, want), np.nonzero(got != want)
    seg.free()
    whole.free()


def test_batch_interleaved_with_single_counts(ctx):
    graphs = Graph.batch(7, 100, 32)
    want = np.array([oracle.tc_count(g.out_rowptr(), g.out_colidx()) for g in graphs], np.uint64)
    b = ctx.batch(graphs)
    single = ctx.upload(graphs[3])
    for _ in range(3):
        assert np.array_equal(b.count()[0], want)
        assert ctx.count_binned(single)[0] == want[3]
    single.free()
    b.free()


@pytest.mark.parametrize("k,balance", [(2, "vertices"), (4, "vertices"), (5, "vertices"),
                                       (4, "edges"), (4, "wedges"), (4, "partner"),
                                       (7, "partner")])
def test_partition_halo_counts(ctx, k, balance):
    g = Graph.rmat(1 << 16, 16, 3).oriented()
    total = oracle.tc_count(g.out_rowptr(), g.out_colidx())
    pg = PartitionedGraph(g, k, balance)
    got_sum = 0
    for p in range(k):
        part = pg.halo_partition(p)
        want = oracle.tc_count_halo(part.row_begin, part.row_end, part.local_rowptr,
                                    part.local_colidx, part.halo_ids, part.halo_rowptr,
                                    part.halo_colidx)
        dg = ctx.upload_partition(part)
        assert dg.info()["partner_rows"] == part.work()["partner_rows"]
        got, st = ctx.count_binned(dg)
        assert got == want, (p, got, want)
        assert ctx.count_plain(dg)[0] == want
        dg.free()
        got_sum += got
    assert got_sum == total


def test_metis_parts_count_their_own_rows(ctx, tmp_path):
    from hetegeneous_parallel_programming_amd import MetisPart, assemble_metis

    g = Graph.rmat(1 << 15, 16, 8).oriented()
    rp, ci = g.out_rowptr(), g.out_colidx()
    k = 4
    part = (np.arange(g.V(), dtype=np.int64) * 104729 + 5) % k   # scattered ownership
    prefix = str(tmp_path / "pgraph")
    g.write_metis(prefix, part)
    pes = MetisPart.read_all(prefix, k)
    want_v = oracle.tc_per_vertex(rp, ci)
    acc, acc_v = 0, np.zeros(g.V(), np.uint64)
    for i, pe in enumerate(pes):
        sub, owned = assemble_metis(pes, i)
        want = oracle.tc_count_rows(rp, ci, pe.vertex_list)
        dg = ctx.upload(sub, owned=owned)
        assert ctx.count_binned(dg)[0] == want
        assert ctx.count_plain(dg)[0] == want
        total, out = ctx.count_local(dg, edges=False)
        assert total == want
        acc_v += out["vertex_counts"]
        dg.free()
        acc += want
    assert acc == oracle.tc_count(rp, ci)
    assert np.array_equal(acc_v, want_v)


def test_owned_mask_edge_cases(ctx):
    g = Graph.rmat(1 << 12, 8, 3).oriented()
    rp, ci = g.out_rowptr(), g.out_colidx()
    none = np.zeros(g.V(), np.uint8)
    dg = ctx.upload(g, owned=none)
    assert ctx.count_binned(dg)[0] == 0 and ctx.count_plain(dg)[0] == 0
    total, out = ctx.count_local(dg)
    assert total == 0 and not out["vertex_counts"].any() and not out["edge_counts"].any()
    dg.free()
    every = np.ones(g.V(), np.uint8)
    dg = ctx.upload(g, owned=every)
    assert ctx.count_binned(dg)[0] == oracle.tc_count(rp, ci)
    dg.free()
    odd = (np.arange(g.V()) % 2).astype(np.uint8)
    dg = ctx.upload(g, owned=odd)
    assert ctx.count_binned(dg)[0] == oracle.tc_count_rows(rp, ci, np.flatnonzero(odd))
    dg.free()
