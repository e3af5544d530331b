// tc_run: command-line driver with the reference's interface
// (successful_push_pull/main.cc:11-36):
//
//   tc_run <graph> [num_gpu(1)] [chunk_size(1024)] [adj_sorted(1)]
//
// <graph> is a GraphMiner prefix (prefix.meta.txt / .vertex.bin / .edge.bin)
// loaded as a DAG, or "rmat:<scale>:<edge_factor>:<seed>" /
// "er:<n>:<avg_degree>:<seed>" for a seeded synthetic graph.  Prints the
// same lines as the reference, ending in "total_num_triangles = <count>".
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "tc_graph.hpp"

using namespace hpp_tc;

static Graph load(const std::string& spec) {
  if (spec.rfind("rmat:", 0) == 0) {
    int scale = 0;
    double ef = 0;
    unsigned long long seed = 0;
    if (std::sscanf(spec.c_str(), "rmat:%d:%lf:%llu", &scale, &ef, &seed) != 3 || scale < 1 || scale > 30)
      throw Error(TC_EINVAL, "expected rmat:<scale>:<edge_factor>:<seed>");
    Graph g = Graph::rmat(1 << scale, ef, seed);
    g.orientation();
    return g;
  }
  if (spec.rfind("er:", 0) == 0) {
    int n = 0;
    double d = 0;
    unsigned long long seed = 0;
    if (std::sscanf(spec.c_str(), "er:%d:%lf:%llu", &n, &d, &seed) != 3)
      throw Error(TC_EINVAL, "expected er:<n>:<avg_degree>:<seed>");
    Graph g = Graph::er(n, d, seed);
    g.orientation();
    return g;
  }
  return Graph(spec, /*use_dag=*/true);
}

int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::cout << "Usage: " << argv[0] << " <graph> [num_gpu(1)] [chunk_size(1024)] [adj_sorted(1)]\n";
    std::cout << "Example: " << argv[0] << " /graph_inputs/mico/graph\n";
    return 1;
  }
  std::cout << "Triangle Counting: we assume the neighbor lists are sorted.\n";
  try {
    Graph g = load(argv[1]);
    int n_devices = 1, chunk_size = 1024, adj_sorted = 1;
    if (argc > 2) n_devices = std::atoi(argv[2]);
    if (argc > 3) chunk_size = std::atoi(argv[3]);
    g.print_meta_data();
    if (argc > 4) adj_sorted = std::atoi(argv[4]);
    if (!adj_sorted) g.sort_neighbors();
    uint64_t total = 0;
    auto t0 = std::chrono::steady_clock::now();
    TCSolver(g, total, n_devices, chunk_size);
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "runtime[gpu 0] = " << s << " sec\n";
    std::cout << "total_num_triangles = " << total << "\n";
  } catch (const Error& e) {
    std::cerr << "tc_run: " << e.what() << " (status " << e.status << ")\n";
    return 2;
  }
  return 0;
}
