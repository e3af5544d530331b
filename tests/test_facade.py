"""The header-only C++ facade include/tc_graph.hpp stands in for the
reference's graph.h + TCSolver (successful_push_pull/main.cc:5,11-36,
graph.h:14-188, graph_partition.h:11-68): a driver written against those
names compiles and links against the facade and the two libraries alone.

Without a GPU the program stops at TCSolver with the device error the facade
throws; with one, its count must equal the oracle's."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from hetegeneous_parallel_programming_amd import Graph
from hetegeneous_parallel_programming_amd import _native as nat
from hetegeneous_parallel_programming_amd._native import LIB_DIR, TC_EDEVICE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# A driver in the shape of the reference's main: load a GraphMiner graph
# (oriented when asked), print its meta data, walk the CSR through the
# accessors, split it the reference's way, and count.
DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tc_graph.hpp"
using namespace hpp_tc;

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  Graph g(argv[1], std::atoi(argv[2]) != 0);
  g.print_meta_data();
  eidType walked = 0;
  vidType longest = 0;
  for (vidType v = 0; v < g.V(); ++v) {
    walked += g.edge_end(v) - g.edge_begin(v);
    if (g.out_degree(v) > longest) longest = g.out_degree(v);
    for (vidType i = 1; i < g.out_degree(v); ++i)
      if (g.N(v, i - 1) >= g.N(v, i)) return 4;  // rows must stay sorted
  }
  std::printf("walked=%lld longest=%d max=%d\n", (long long)walked, longest, g.get_max_degree());
  PartitionedGraph pg(&g, 4);
  pg.edgecut_induced_partition1D();
  vidType owned = 0;
  for (int i = 0; i < pg.get_num_subgraphs(); ++i) owned += pg.get_local_end(i) - pg.get_local_begin(i);
  std::printf("subgraphs=%d owned=%d\n", pg.get_num_subgraphs(), owned);
  uint64_t total = 0;
  try {
    TCSolver(g, total, 1, 1024);
  } catch (const Error& e) {
    std::printf("device error %d\n", (int)e.status);
    return 3;
  }
  std::printf("total_num_triangles = %llu\n", (unsigned long long)total);
  return 0;
}
"""


def build_driver(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "driver"
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-L", LIB_DIR,
           "-ltc_hip", "-ltc_host", f"-Wl,-rpath,{LIB_DIR}", "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    return str(exe)


def test_reference_style_driver_builds_and_runs(tmp_path):
    exe = build_driver(tmp_path)
    g = Graph.rmat(1 << 12, 8, 9)
    prefix = str(tmp_path / "graph")
    g.write(prefix)
    d = g.oriented()
    p = subprocess.run([exe, prefix, "1"], capture_output=True, text=True, timeout=300)
    lines = p.stdout.splitlines()
    assert lines[0] == f"|V|: {d.V()}, |E|: {d.E()}, Max Degree: {d.get_max_degree()}"
    longest = int(np.diff(d.out_rowptr()).max())
    assert lines[1] == f"walked={d.E()} longest={longest} max={d.get_max_degree()}"
    assert lines[2] == f"subgraphs=4 owned={d.V()}"
    if p.returncode == 3:  # no device here: the facade throws instead of exiting
        assert lines[3] == f"device error {TC_EDEVICE}"
    else:
        assert p.returncode == 0, p.stderr
        want = oracle.tc_count(d.out_rowptr(), d.out_colidx())
        assert lines[3] == f"total_num_triangles = {want}"


# A driver in the shape of the reference's per-PE program (main.cc under
# USE_METIS): the PE id from initialize_nvshmem, PE i's METIS part files
# through the 8-argument Graph constructor, the PE count as n_gpus, the
# reference's own prototypes beside the drop-in graph.h.
METIS_DRIVER = r"""
#include "graph.h"

void TCSolver(Graph &g, uint64_t &total, int n_gpu, int chunk_size);
int initialize_nvshmem();

int main(int argc, char *argv[]) {
  if (argc < 3) return 2;
  int mype = initialize_nvshmem();
  Graph g(argv[1], false, false, false, false, false, false, mype);
  int n_devices = std::atoi(argv[2]);
  g.print_meta_data();
  vidType* vl = g.get_vertex_list();
  hpp_tc::idx_t* part = g.get_metis_part();
  eidType owned_edges = 0;
  for (vidType x = 0; x < g.get_n_real_vertices(); ++x) {
    if (part[vl[x]] != mype) return 5;
    owned_edges += g.get_degree(vl[x]);
    VertexSet s = g.N(vl[x]);
    for (vidType i = 0; i < s.size(); ++i)
      if (s[i] != g.getEdgeDst(g.edge_begin(vl[x]) + i) || s[i] != g.adj_ptr(vl[x])[i]) return 6;
  }
  std::printf("pe=%d n_real=%d owned_edges=%lld max_subg_nv=%d max_subg_ne=%lld\n", mype,
              g.get_n_real_vertices(), (long long)owned_edges, g.get_max_subg_nv(),
              (long long)g.get_max_subg_ne());
  std::fflush(stdout);
  uint64_t total = 0;
  TCSolver(g, total, n_devices, 1024);
  std::cout << "total_num_triangles = " << total << "\n";
  return 0;
}
"""


def build_with(tmp_path, name, source, defines=()):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / f"{name}.cpp"
    src.write_text(source)
    exe = tmp_path / name
    cmd = ["g++", "-std=c++17", "-O1", *[f"-D{d}" for d in defines],
           "-I", os.path.join(ROOT, "include", "compat"), "-I", os.path.join(ROOT, "include"), str(src),
           "-L", LIB_DIR, "-ltc_hip", "-ltc_host", f"-Wl,-rpath,{LIB_DIR}", "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    return str(exe)


def metis_files(tmp_path, k=3, seed=5):
    d = Graph.rmat(1 << 12, 8, seed).oriented()
    part = np.random.default_rng(seed).integers(0, k, d.V()).astype(np.int64)
    (tmp_path / "metis").mkdir(exist_ok=True)
    prefix = str(tmp_path / "metis" / "pgraph")
    d.write_metis(prefix, part)
    return d, part, prefix


def run_pe(exe, prefix, k, pe):
    env = dict(os.environ, TC_PE=str(pe), TC_LOCAL_PE="0")
    return subprocess.run([exe, prefix, str(k)], capture_output=True, text=True, timeout=300, env=env)


def test_metis_mode_driver_loads_each_pe(tmp_path):
    """Host side of the per-PE program on every PE: meta data from the part
    files, the PE's vertex_list and part array, its rows through N(v),
    getEdgeDst and adj_ptr.  Without a device the solver exits with its
    error, as the reference's does."""
    from hetegeneous_parallel_programming_amd import MetisPart

    exe = build_with(tmp_path, "metis_driver", METIS_DRIVER, ["USE_METIS"])
    d, part, prefix = metis_files(tmp_path)
    k = 3
    for pe in range(k):
        m = MetisPart(prefix, pe)
        p = run_pe(exe, prefix, k, pe)
        lines = p.stdout.splitlines()
        assert lines[0] == f"|V|: {d.V()}, |E|: {m.nnz}, Max Degree: {m.max_degree}", p.stderr
        assert lines[1] == (f"pe={pe} n_real={int((part == pe).sum())} owned_edges={m.nnz} "
                            f"max_subg_nv={m.max_subg_nv} max_subg_ne={m.max_subg_ne}")
        if p.returncode == 1:  # no device here
            assert "TCSolver:" in p.stderr and f"(status {TC_EDEVICE})" in p.stderr
        else:
            assert p.returncode == 0, p.stderr


def test_metis_mode_rejects_a_wrong_pe_count(tmp_path):
    exe = build_with(tmp_path, "metis_driver", METIS_DRIVER, ["USE_METIS"])
    _, _, prefix = metis_files(tmp_path)
    p = run_pe(exe, prefix, 2, 0)  # part values reach 2: three PEs wrote these files
    assert p.returncode == 1 and f"(status {nat.TC_EINVAL})" in p.stderr, (p.returncode, p.stderr)
    p = run_pe(exe, prefix, 4, 0)  # PE 3's files do not exist
    assert p.returncode == 1 and "TCSolver:" in p.stderr


@pytest.mark.gpu
def test_metis_mode_driver_counts_each_pe_on_the_gpu(tmp_path):
    """Per PE the CLI prints the count of its vertex_list pivots (local and
    remote lists alike), which the oracle gives by counting exactly those
    rows; the PEs' counts add up to the graph's total."""
    from hetegeneous_parallel_programming_amd import MetisPart

    exe = build_with(tmp_path, "metis_driver", METIS_DRIVER, ["USE_METIS"])
    d, part, prefix = metis_files(tmp_path)
    rp, ci = d.out_rowptr(), d.out_colidx()
    k, got = 3, []
    for pe in range(k):
        p = run_pe(exe, prefix, k, pe)
        assert p.returncode == 0, p.stderr
        m = MetisPart(prefix, pe)  # its arrays are views of memory the object frees
        want = oracle.tc_count_rows(rp, ci, m.vertex_list.copy())
        assert p.stdout.splitlines()[-1] == f"total_num_triangles = {want}"
        got.append(want)
    assert sum(got) == oracle.tc_count(rp, ci)


REF_MAIN = "/root/reference/successful_push_pull/main.cc"


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="reference tree not present")
@pytest.mark.parametrize("defines", [[], ["USE_METIS"]], ids=["whole_graph", "metis"])
def test_reference_main_compiles_unchanged(tmp_path, defines):
    """The reference's own main.cc, read where it lies (not copied) and fed
    to the compiler on stdin -- so its #include "graph.h" resolves to the
    drop-in graph.h, as once the maintainer swaps that header -- links
    against the two libraries; run without a device, it loads the graph and
    exits in the solver with the device error."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "ref_main"
    cmd = ["g++", "-std=c++17", "-O1", *[f"-D{x}" for x in defines],
           "-I", os.path.join(ROOT, "include", "compat"), "-I", os.path.join(ROOT, "include"), "-x", "c++", "-",
           "-x", "none", "-L", LIB_DIR, "-ltc_hip", "-ltc_host", f"-Wl,-rpath,{LIB_DIR}", "-o", str(exe)]
    with open(REF_MAIN) as src:
        p = subprocess.run(cmd, stdin=src, capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    assert p.returncode == 0, p.stderr
    if defines:
        d, _, prefix = metis_files(tmp_path)
        args, env = [prefix, "3"], dict(os.environ, TC_PE="1")
    else:
        g = Graph.rmat(1 << 10, 8, 3)
        prefix = str(tmp_path / "graph")
        g.write(prefix)
        d = g.oriented()
        args, env = [prefix, "1"], dict(os.environ)
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=300, env=env)
    assert "Triangle Counting" in r.stdout
    if r.returncode != 0:  # no device here
        assert r.returncode == 1 and f"(status {TC_EDEVICE})" in r.stderr, (r.returncode, r.stderr)
    elif not defines:
        want = oracle.tc_count(d.out_rowptr(), d.out_colidx())
        assert r.stdout.splitlines()[-1] == f"total_num_triangles = {want}"


def test_host_intersect_num_and_vertex_set():
    """Graph::intersect_num (graph.cc:376-391) on the host library: |N(v) ∩ N(u)|
    of two sorted rows, against numpy on random pairs; an out-of-range id is an
    error, not a read past the arrays."""
    import ctypes

    g = Graph.rmat(1 << 10, 8, 4)
    rp, ci = g.out_rowptr(), g.out_colidx()
    rng = np.random.default_rng(0)
    host = nat.host()
    for _ in range(200):
        v, u = (int(x) for x in rng.integers(0, g.V(), 2))
        want = np.intersect1d(ci[rp[v]:rp[v + 1]], ci[rp[u]:rp[u + 1]]).size
        assert host.tch_intersect_num(ctypes.byref(g.view()), v, u, None, 0) == want
    labels = rng.integers(0, 3, g.V()).astype(np.uint8)
    lp = labels.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    v, u = 0, 1
    common = np.intersect1d(ci[rp[v]:rp[v + 1]], ci[rp[u]:rp[u + 1]])
    for lab in range(3):
        assert host.tch_intersect_num(ctypes.byref(g.view()), v, u, lp, lab) == int((labels[common] == lab).sum())
    assert host.tch_intersect_num(ctypes.byref(g.view()), g.V(), 0, None, 0) == -1
