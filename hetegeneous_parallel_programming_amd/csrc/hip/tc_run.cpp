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
// "metis:<prefix>:<k>" reads the reference's per-PE METIS part files
// <prefix>0.* .. <prefix>{k-1}.* (graph_partition.cc:501-566) and counts each
// PE's own rows against the remote lists they reach, printing
// "pe[<i>] n_real = <owned> count = <sum>" per PE before the total.
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

// Counts every PE of a METIS file set; returns the total.
static uint64_t count_metis(const std::string& spec) {
  const size_t c = spec.rfind(':');
  if (c == std::string::npos || c <= 6) throw Error(TC_EINVAL, "expected metis:<prefix>:<k>");
  const std::string prefix = spec.substr(6, c - 6);
  const int k = std::atoi(spec.c_str() + c + 1);
  if (k < 1) throw Error(TC_EINVAL, "expected metis:<prefix>:<k> with k >= 1");
  std::vector<tch_metis*> pes((size_t)k, nullptr);
  auto free_all = [&]() {
    for (tch_metis* m : pes) tch_metis_free(m);
  };
  Device dev(0);
  uint64_t total = 0;
  try {
    for (int i = 0; i < k; ++i) host_check(tch_metis_read(prefix.c_str(), i, &pes[(size_t)i]));
    for (int i = 0; i < k; ++i) {
      tch_metis_view mv;
      host_check(tch_metis_get(pes[(size_t)i], &mv));
      tch_csr* sub = nullptr;
      std::vector<uint8_t> owned((size_t)mv.n);
      host_check(tch_metis_assemble(pes.data(), k, i, &sub, owned.data()));
      tc_csr_view v;
      tch_csr_view(sub, &v);
      tc_graph* dg = nullptr;
      uint64_t cnt = 0;
      tc_status st = tc_graph_upload_rows(dev.get(), &v, owned.data(), TC_CHECK_SHAPE, &dg);
      if (st == TC_OK) st = tc_count(dev.get(), dg, &cnt, nullptr);
      tc_graph_free(dev.get(), dg);
      tch_csr_free(sub);
      dev_check(st);
      std::cout << "pe[" << i << "] n_real = " << mv.n_real << " count = " << cnt << "\n";
      total += cnt;
    }
  } catch (...) {
    free_all();
    throw;
  }
  free_all();
  return total;
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
    if (pos[0].rfind("metis:", 0) == 0) {
      auto t0 = std::chrono::steady_clock::now();
      const uint64_t total = count_metis(pos[0]);
      double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      std::cout << "runtime[gpu 0] = " << s << " sec\n";
      std::cout << "total_num_triangles = " << total << "\n";
      return 0;
    }
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
