// tc_run: command-line driver with the reference's interface
// (successful_push_pull/main.cc:11-36):
//
//   tc_run <graph> [num_gpu(1)] [chunk_size(1024)] [adj_sorted(1)]
//          [--parts K] [--balance vertices|edges|wedges|partner]
//
// <graph> is a GraphMiner prefix (prefix.meta.txt / .vertex.bin / .edge.bin)
// loaded as a DAG, or "rmat:<scale>:<edge_factor>:<seed>" /
// "er:<n>:<avg_degree>:<seed>" for a seeded synthetic graph.  Prints the
// same lines as the reference, ending in "total_num_triangles = <count>".
// With --parts K it also prints, per PE of a K-way split, the sum the
// reference's push/pull kernel reports for that PE
// (successful_push_pull/bs_warp_vertex_nvshmem.cuh:127-136), as
// "pe[<p>] ids [<begin>, <end>) count = <sum>".
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

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

static int balance_of(const std::string& name) {
  if (name == "vertices") return TCH_BALANCE_VERTICES;
  if (name == "edges") return TCH_BALANCE_EDGES;
  if (name == "wedges") return TCH_BALANCE_WEDGES;
  if (name == "partner") return TCH_BALANCE_PARTNER;
  throw Error(TC_EINVAL, "--balance takes vertices, edges, wedges or partner");
}

int main(int argc, char* argv[]) {
  // positional arguments as main.cc:11-36; options anywhere after <graph>
  std::vector<std::string> pos;
  int parts = 0;
  std::string balance = "vertices";
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if ((a == "--parts" || a == "--balance") && i + 1 < argc) {
      (a == "--parts" ? (void)(parts = std::atoi(argv[++i])) : (void)(balance = argv[++i]));
    } else {
      pos.push_back(a);
    }
  }
  if (pos.empty()) {
    std::cout << "Usage: " << argv[0] << " <graph> [num_gpu(1)] [chunk_size(1024)] [adj_sorted(1)]"
              << " [--parts K] [--balance vertices|edges|wedges|partner]\n";
    std::cout << "Example: " << argv[0] << " /graph_inputs/mico/graph\n";
    return 1;
  }
  std::cout << "Triangle Counting: we assume the neighbor lists are sorted.\n";
  try {
    if (parts < 0 || parts > 1024) throw Error(TC_EINVAL, "--parts must be in [1, 1024]");
    const int bal = balance_of(balance);
    Graph g = load(pos[0]);
    int n_devices = 1, chunk_size = 1024, adj_sorted = 1;
    if (pos.size() > 1) n_devices = std::atoi(pos[1].c_str());
    if (pos.size() > 2) chunk_size = std::atoi(pos[2].c_str());
    g.print_meta_data();
    if (pos.size() > 3) adj_sorted = std::atoi(pos[3].c_str());
    if (!adj_sorted) g.sort_neighbors();
    uint64_t total = 0;
    auto t0 = std::chrono::steady_clock::now();
    TCSolver(g, total, n_devices, chunk_size);
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "runtime[gpu 0] = " << s << " sec\n";
    if (parts > 0) {
      const std::vector<int32_t> b = PartitionedGraph(&g, parts, bal).bounds();
      std::vector<uint64_t> pe;
      Device dev(0);
      dev.count_local(g, nullptr, nullptr, &b, &pe);
      for (int p = 0; p < parts; ++p)
        std::cout << "pe[" << p << "] ids [" << b[(size_t)p] << ", " << b[(size_t)p + 1] << ") count = " << pe[(size_t)p]
                  << "\n";
    }
    std::cout << "total_num_triangles = " << total << "\n";
  } catch (const Error& e) {
    std::cerr << "tc_run: " << e.what() << " (status " << e.status << ")\n";
    return 2;
  }
  return 0;
}
