"""The tc_run driver keeps the reference's command line (main.cc:11-36)."""
import os
import re
import subprocess

import pytest

import oracle
from hetegeneous_parallel_programming_amd import Graph
from hetegeneous_parallel_programming_amd._native import LIB_DIR

TC_RUN = os.path.join(LIB_DIR, "tc_run")


def run(*args):
    return subprocess.run([TC_RUN, *args], capture_output=True, text=True, timeout=300)


def test_usage_without_arguments():
    p = run()
    assert p.returncode == 1 and "Usage:" in p.stdout and "[num_gpu(1)]" in p.stdout


def test_bad_graph_is_a_clean_error():
    p = run("/nonexistent/graph")
    assert p.returncode == 2 and "tc_run:" in p.stderr


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="only meaningful where no GPU is visible")
def test_no_device_is_a_clean_error():
    p = run("er:1000:8:1")
    assert p.returncode == 2 and "status 3" in p.stderr


@pytest.mark.gpu
def test_counts_graphminer_file_and_generated_graph(tmp_path):
    g = Graph.rmat(1 << 14, 16, 5)
    prefix = str(tmp_path / "graph")
    g.write(prefix)
    d = g.oriented()
    want = oracle.tc_count(d.out_rowptr(), d.out_colidx())
    for args in ((prefix,), (prefix, "1", "1024", "0"), ("rmat:14:16:5",)):
        p = run(*args)
        assert p.returncode == 0, p.stderr
        m = re.search(r"total_num_triangles = (\d+)", p.stdout)
        assert m and int(m.group(1)) == want, p.stdout
    p = run(prefix, "2")
    assert p.returncode == 2 and "exactly one GPU" in p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("balance", ["vertices", "partner"])
def test_per_pe_attribution_lines(balance):
    from hetegeneous_parallel_programming_amd import PartitionedGraph

    d = Graph.rmat(1 << 14, 16, 5).oriented()
    rp, ci = d.out_rowptr(), d.out_colidx()
    bounds = PartitionedGraph(d, 3, balance).bounds()
    want = oracle.tc_pushpull_bounds(rp, ci, bounds)
    p = run("rmat:14:16:5", "--parts", "3", "--balance", balance)
    assert p.returncode == 0, p.stderr
    got = re.findall(r"pe\[(\d+)\] ids \[(\d+), (\d+)\) count = (\d+)", p.stdout)
    assert [(int(a), int(b), int(c), int(x)) for a, b, c, x in got] == [
        (i, int(bounds[i]), int(bounds[i + 1]), int(want[i])) for i in range(3)]
    m = re.search(r"total_num_triangles = (\d+)", p.stdout)
    assert int(m.group(1)) == int(want.sum())


@pytest.mark.gpu
def test_metis_part_files(tmp_path):
    import numpy as np

    d = Graph.rmat(1 << 13, 16, 2).oriented()
    rp, ci = d.out_rowptr(), d.out_colidx()
    k = 3
    part = (np.arange(d.V(), dtype=np.int64) * 7919) % k
    prefix = str(tmp_path / "pgraph")
    d.write_metis(prefix, part)
    p = run(f"metis:{prefix}:{k}")
    assert p.returncode == 0, p.stderr
    got = re.findall(r"pe\[(\d+)\] n_real = (\d+) count = (\d+)", p.stdout)
    want = [(i, int((part == i).sum()), oracle.tc_count_rows(rp, ci, np.flatnonzero(part == i)))
            for i in range(k)]
    assert [(int(a), int(b), int(c)) for a, b, c in got] == want
    m = re.search(r"total_num_triangles = (\d+)", p.stdout)
    assert int(m.group(1)) == oracle.tc_count(rp, ci)
    assert run(f"metis:{prefix}:{k + 1}").returncode == 2   # PE 3's files do not exist
