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
