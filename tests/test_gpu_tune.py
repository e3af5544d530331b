"""GPU tests of the autotuner and the per-device tune file.

The tuner may only change how the count is launched, never what it counts:
every setting it tries must give the oracle's counts, and the setting it
installs must too.
"""
import numpy as np
import pytest

import oracle
from hetegeneous_parallel_programming_amd import Graph
from hetegeneous_parallel_programming_amd._native import TcError

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_tune(ctx):
    saved = ctx.tune
    yield
    ctx.tune = saved


def ordered(t):
    return 0 <= t["t_g8"] <= t["t_g16"] <= t["t_g32"] <= t["t_wave"] and 1 <= t["rounds"]


def test_autotune_batch_keeps_counts(ctx, restore_tune):
    graphs = Graph.batch(3, 0, 64)
    want = np.array([oracle.tc_count(g.out_rowptr(), g.out_colidx()) for g in graphs], np.uint64)
    b = ctx.batch(graphs)
    t, rep = ctx.autotune(batch=b, reps=1, budget_ms=4000)
    assert ordered(t)
    assert ctx.tune == t
    assert rep["trials"] >= 1 and rep["best_ms"] <= rep["start_ms"]
    for _ in range(2):
        assert np.array_equal(b.count()[0], want)
    b.free()


def test_autotune_graph_keeps_count(ctx, restore_tune):
    g = Graph.rmat(1 << 15, 16, 21).oriented()
    want = oracle.tc_count(g.out_rowptr(), g.out_colidx())
    dg = ctx.upload(g)
    t, rep = ctx.autotune(graph=dg, reps=1, budget_ms=3000)
    assert ordered(t)
    assert ctx.count_binned(dg)[0] == want
    dg.free()


def test_autotune_budget_bounds_the_search(ctx, restore_tune):
    g = Graph.er(4096, 12, 2).oriented()
    dg = ctx.upload(g)
    _, rep = ctx.autotune(graph=dg, reps=1, budget_ms=0.0)
    assert rep["trials"] == 1  # only the starting setting
    dg.free()


def test_autotune_arguments(ctx):
    g = Graph.er(512, 8, 1).oriented()
    dg = ctx.upload(g)
    with pytest.raises(ValueError):
        ctx.autotune()
    b = ctx.batch([g])
    with pytest.raises(ValueError):
        ctx.autotune(graph=dg, batch=b)
    with pytest.raises(TcError):
        ctx.autotune(graph=dg, reps=0)
    b.free()
    dg.free()


def test_tune_file_round_trip(ctx, restore_tune, tmp_path):
    path = str(tmp_path / "tune.txt")
    assert not ctx.load_tune(path, "mix")  # no file yet
    a = {"t_g8": 128, "t_g16": 512, "t_g32": 2048, "t_wave": 8192, "rounds": 2, "grid": 0,
         "pivot_set": 0, "hash_kicks": 64}
    b = {"t_g8": 0, "t_g16": 256, "t_g32": 1024, "t_wave": 32768, "rounds": 8, "grid": 512,
         "pivot_set": 1, "hash_kicks": 128}
    ctx.tune = a
    ctx.save_tune(path, "mix")
    ctx.tune = b
    ctx.save_tune(path, "other")
    ctx.tune = {"rounds": 1}
    assert ctx.load_tune(path, "mix") and ctx.tune == a
    assert ctx.load_tune(path, "other") and ctx.tune == b
    ctx.tune = b
    ctx.save_tune(path, "mix")  # replaces, keeps "other"
    lines = open(path).read().splitlines()
    dev = "".join("_" if c in " \t|" else c for c in ctx.name)
    assert len(lines) == 2 and all(l.startswith(dev + "|") for l in lines)
    assert not ctx.load_tune(path, "absent")
    with pytest.raises(TcError):
        ctx.save_tune(path, "bad key")
