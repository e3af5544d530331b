// graph.h -- drop-in for the reference's successful_push_pull/graph.h on the
// triangle-counting path.  It puts the facade's Graph, PartitionedGraph and
// TCSolver in the global namespace with the reference's signatures, so the
// reference's per-process driver (successful_push_pull/main.cc:1-36: its own
// prototypes of TCSolver and, under USE_METIS, initialize_nvshmem included)
// compiles unchanged with
//     -I <repo>/include/compat -I <repo>/include [-DUSE_METIS]
// and links against libtc_hip.so and libtc_host.so in place of the CUDA /
// NVSHMEM objects (successful_push_pull/Makefile:46-48).
//
// As in the reference, a failing solver prints the error and exits (the
// reference's CUDA_SAFE_CALL and device-count checks exit,
// multigpu_nvshmem_pushpull.cu:33-36); the namespaced facade in tc_graph.hpp
// throws hpp_tc::Error instead.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <string>

#include "tc_graph.hpp"

#ifndef USE_DAG
#define USE_DAG true  // main.cc's whole-graph branch loads the graph as a degree-oriented DAG
#endif

using hpp_tc::AccType;
using hpp_tc::eidType;
using hpp_tc::Graph;
using hpp_tc::PartitionedGraph;
using hpp_tc::VertexSet;
using hpp_tc::vidType;

inline void TCSolver(Graph& g, uint64_t& total, int n_gpu, int chunk_size) {
  try {
    hpp_tc::tc_solve(g, total, n_gpu, chunk_size);
  } catch (const hpp_tc::Error& e) {
    std::cerr << "TCSolver: " << e.what() << " (status " << (int)e.status << ")\n";
    std::exit(1);
  }
}

#ifdef USE_METIS
inline int initialize_nvshmem() { return hpp_tc::initialize_nvshmem(); }
#endif
