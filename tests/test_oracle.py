"""Pins the CPU oracle (test infrastructure) before anything is checked
against it: closed forms, brute force, library routines and the golden
fixtures written by networkx/scipy (tests/golden/make_golden.py)."""
import itertools
import json
import math
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dag_of_pairs(n, src, dst):
    rp, ci = oracle.symmetric_csr(n, src, dst)
    return oracle.orient(rp, ci)


def count_pairs(n, src, dst, method=0):
    rp, ci = dag_of_pairs(n, src, dst)
    return oracle.tc_count(rp, ci, method=method)


def clique(n, base=0):
    s, d = zip(*itertools.combinations(range(base, base + n), 2)) if n > 1 else ((), ())
    return list(s), list(d)


@pytest.mark.parametrize("n", [3, 4, 5, 10, 31, 64, 100])
def test_clique_closed_form(n):
    s, d = clique(n)
    assert count_pairs(n, s, d) == math.comb(n, 3)
    assert count_pairs(n, s, d, method=1) == math.comb(n, 3)


@pytest.mark.parametrize("k", [4, 5, 9, 50])
def test_wheel(k):
    # hub 0 + rim cycle of k vertices: k triangles (k >= 4)
    rim = list(range(1, k + 1))
    s = [0] * k + rim
    d = rim + rim[1:] + rim[:1]
    assert count_pairs(k + 1, s, d) == k


def test_triangle_free_families():
    a, b = 7, 9  # complete bipartite
    s = [i for i in range(a) for _ in range(b)]
    d = [a + j for _ in range(a) for j in range(b)]
    assert count_pairs(a + b, s, d) == 0
    n = 40  # even cycle, path, star, binary tree
    assert count_pairs(n, list(range(n)), [(i + 1) % n for i in range(n)]) == 0
    assert count_pairs(n, list(range(n - 1)), list(range(1, n))) == 0
    assert count_pairs(n, [0] * (n - 1), list(range(1, n))) == 0
    assert count_pairs(n, [(i - 1) // 2 for i in range(1, n)], list(range(1, n))) == 0
    assert count_pairs(3, [0, 1, 2], [1, 2, 0]) == 1  # C3


def test_disjoint_union_adds():
    sizes = [3, 4, 6, 11, 17]
    s, d, base = [], [], 0
    for k in sizes:
        cs, cd = clique(k, base)
        s += cs
        d += cd
        base += k
    assert count_pairs(base, s, d) == sum(math.comb(k, 3) for k in sizes)


@pytest.mark.parametrize("seed", range(8))
def test_brute_force_small(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 64))
    m = int(rng.integers(0, n * 6))
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    want = oracle.brute_force_triangles(n, s, d)
    assert count_pairs(n, s, d) == want
    assert count_pairs(n, s, d, method=1) == want


@pytest.mark.parametrize("seed", range(5))
def test_intersections_match_set_semantics(seed):
    rng = np.random.default_rng(100 + seed)
    for _ in range(50):
        a = np.unique(rng.integers(0, 500, rng.integers(0, 120))).astype(np.int32)
        b = np.unique(rng.integers(0, 500, rng.integers(0, 120))).astype(np.int32)
        want = len(set(a.tolist()) & set(b.tolist()))
        assert oracle.intersect_merge(a, b) == want
        assert oracle.intersect_bsearch(a, b) == want
        for key in rng.integers(-5, 505, 10):
            assert oracle.binary_search(a, key) == (int(key) in set(a.tolist()))


def test_orientation_rule_and_invariants():
    rp, ci = oracle.gen_rmat(2048, 12, 9)
    drp, dci = oracle.orient(rp, ci)
    nrp, nci = oracle.orient_numpy(rp, ci)
    assert np.array_equal(drp, nrp) and np.array_equal(dci, nci)
    assert dci.size * 2 == ci.size  # every undirected edge kept exactly once
    deg = np.diff(rp)
    src = np.repeat(np.arange(rp.size - 1), np.diff(drp))
    key_s = deg[src] * (1 << 32) + src
    key_d = deg[dci] * (1 << 32) + dci
    assert (key_d > key_s).all()  # strictly up the (degree, id) order => acyclic
    for u in range(0, rp.size - 1, 97):  # rows stay sorted
        row = dci[drp[u]:drp[u + 1]]
        assert (np.diff(row) > 0).all()


def test_orientation_of_regular_graph_is_upper_triangle():
    n = 30
    s, d = clique(n)
    rp, ci = dag_of_pairs(n, s, d)
    for u in range(n):
        assert ci[rp[u]:rp[u + 1]].tolist() == list(range(u + 1, n))
    cyc_rp, cyc_ci = dag_of_pairs(10, list(range(10)), [(i + 1) % 10 for i in range(10)])
    src = np.repeat(np.arange(10), np.diff(cyc_rp))
    assert (cyc_ci > src).all()


def test_golden_fixtures_small():
    gold = json.load(open(os.path.join(GOLDEN, "small.json")))
    for case in gold["cases"]:
        kind, n, p, seed = case["kind"], case["n"], case["param"], case["seed"]
        rp, ci = oracle.gen_er(n, p, seed) if kind == "er" else oracle.gen_rmat(n, p, seed)
        assert oracle.csr_digest(rp, ci) == case["sha256_sym"], case
        drp, dci = oracle.orient(rp, ci)
        assert oracle.tc_count(drp, dci) == case["triangles"], case
        assert oracle.tc_count(drp, dci, method=1) == case["triangles"], case
        if "per_vertex" in case:
            assert oracle.tc_per_vertex(drp, dci).tolist() == case["per_vertex"]


def test_golden_batch_specs():
    gold = json.load(open(os.path.join(GOLDEN, "small.json")))
    for spec in gold["batch_specs"]:
        kind, n, deg, seed = oracle.batch_spec(1, spec["index"])
        assert (kind, n, seed) == (spec["kind"], spec["n"], spec["seed"])
        assert deg == spec["avg_degree"]


def test_count_is_relabel_and_chunk_invariant():
    rng = np.random.default_rng(3)
    rp, ci = oracle.gen_rmat(3000, 10, 4)
    n = rp.size - 1
    src = np.repeat(np.arange(n), np.diff(rp))
    want = oracle.tc_count(*oracle.orient(rp, ci))
    perm = rng.permutation(n)
    rp2, ci2 = oracle.symmetric_csr(n, perm[src], perm[ci])
    assert oracle.tc_count(*oracle.orient(rp2, ci2)) == want
    drp, dci = oracle.orient(rp, ci)
    for chunk, threads in ((1, 1), (7, 3), (4096, 0)):
        assert oracle.tc_count(drp, dci, chunk=chunk, threads=threads) == want
    assert oracle.tc_count_rows(drp, dci, np.arange(n)) == want
    # row ranges add up
    cut = n // 3
    assert oracle.tc_count(drp, dci, 0, cut) + oracle.tc_count(drp, dci, cut, n) == want


def test_per_vertex_sums_to_three_times_total():
    drp, dci = oracle.orient(*oracle.gen_er(4000, 30, 5))
    per = oracle.tc_per_vertex(drp, dci)
    assert int(per.sum()) == 3 * oracle.tc_count(drp, dci)


@pytest.mark.parametrize("k", [2, 3, 4])
def test_pushpull_attribution_partitions_the_total(k):
    drp, dci = oracle.orient(*oracle.gen_rmat(4096, 16, 12))
    total = oracle.tc_count(drp, dci)
    assert sum(oracle.tc_pushpull_pe(drp, dci, k, pe) for pe in range(k)) == total


# Six vertices, ten edges, six triangles ({0,2,4} {0,2,5} {0,4,5} {0,1,5}
# {2,4,5} {2,3,4}).  Undirected degrees 4 2 4 2 4 4 orient it to
# N+ = 0:[2,4,5] 1:[0,5] 2:[4,5] 3:[2,4] 4:[5] 5:[], d+ = 3 2 2 2 1 0.
# Per edge |N+(u) ∩ N+(v)| in CSR order: (0,2) 2 (0,4) 1 (0,5) 0 (1,0) 1
# (1,5) 0 (2,4) 1 (2,5) 0 (3,2) 1 (3,4) 0 (4,5) 0.  Split {0} | {1..5}:
# (0,2) and (0,4) cross with d+(0) >= d+(v): pulled by part 0 (3); (1,0)
# crosses with d+(1) = 2 < d+(0) = 3: pushed to 0's part (1); (2,4), (3,2)
# stay in part 1 (2).
PUSH_EDGES = [(0, 5), (3, 4), (0, 2), (0, 1), (2, 3), (1, 5), (2, 4), (0, 4), (2, 5), (4, 5)]


def test_edge_counts_and_push_attribution_by_hand():
    s, d = zip(*PUSH_EDGES)
    rp, ci = dag_of_pairs(6, s, d)
    assert ci.tolist() == [2, 4, 5, 0, 5, 4, 5, 2, 4, 5]
    assert oracle.tc_edge_counts(rp, ci).tolist() == [2, 1, 0, 1, 0, 1, 0, 1, 0, 0]
    assert oracle.tc_pushpull_bounds(rp, ci, [0, 1, 6]).tolist() == [4, 2]
    assert oracle.tc_per_vertex(rp, ci).tolist() == [4, 1, 4, 1, 4, 4]


@pytest.mark.parametrize("n,d,seed", [(300, 12, 1), (2000, 20, 2)])
def test_edge_counts_match_set_intersection(n, d, seed):
    rp, ci = oracle.orient(*oracle.gen_er(n, d, seed))
    got = oracle.tc_edge_counts(rp, ci)
    want = []
    for u in range(n):
        a = ci[rp[u]:rp[u + 1]]
        for v in a:
            want.append(np.intersect1d(a, ci[rp[v]:rp[v + 1]]).size)
    assert got.tolist() == want
    assert int(got.sum()) == oracle.tc_count(rp, ci)


def test_pushpull_bounds_on_disjoint_cliques():
    # cut at the clique boundaries every edge is local: part p sees its own clique
    sizes = [5, 7, 4, 9]
    src, dst, base = [], [], 0
    for k in sizes:
        s, d = clique(k, base)
        src += s
        dst += d
        base += k
    rp, ci = dag_of_pairs(base, src, dst)
    bounds = np.concatenate([[0], np.cumsum(sizes)])
    assert oracle.tc_pushpull_bounds(rp, ci, bounds).tolist() == [math.comb(k, 3) for k in sizes]
    assert oracle.tc_pushpull_bounds(rp, ci, [0, base]).tolist() == [sum(math.comb(k, 3) for k in sizes)]


@pytest.mark.parametrize("k", [2, 3, 5])
def test_pushpull_bounds_matches_the_chunked_split(k):
    rp, ci = oracle.orient(*oracle.gen_rmat(4096, 16, 12))
    n = rp.size - 1
    chunk = -(-n // k)
    bounds = [min(p * chunk, n) for p in range(k + 1)]
    got = oracle.tc_pushpull_bounds(rp, ci, bounds)
    assert got.tolist() == [oracle.tc_pushpull_pe(rp, ci, k, p) for p in range(k)]


def test_planted_mistake_is_caught():
    """A plausible orientation bug (>= instead of > on equal degrees keeps both
    directions of a tie) must fail the closed form the real rule passes."""
    n = 12
    s, d = clique(n)
    rp, ci = oracle.symmetric_csr(n, s, d)
    deg = np.diff(rp)
    src = np.repeat(np.arange(n), deg)
    keep = deg[ci] >= deg[src]  # planted mistake
    brp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(src[keep], minlength=n), out=brp[1:])
    bad = oracle.tc_count(brp, ci[keep])
    assert oracle.tc_count(*oracle.orient(rp, ci)) == math.comb(n, 3)
    assert bad != math.comb(n, 3)
