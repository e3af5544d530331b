// tc_graph.hpp -- header-only C++ facade over the C ABI (tc_host.h, tc.h)
// with the reference's host-side names, so that a caller of the reference's
// triangle-counting path can switch by changing an include:
//
//   Graph            successful_push_pull/graph.h:14-188 (GraphMiner loader
//                    graph.cc:4-142, orientation graph.cc:252-298,
//                    sort_neighbors graph.cc:156-164, accessors graph.h:75-93)
//   PartitionedGraph successful_push_pull/graph_partition.h:11-68
//                    (edgecut_induced_partition1D graph_partition.cc:85-135)
//   TCSolver         void TCSolver(Graph&, uint64_t& total, int n_gpus,
//                    int chunk_size)  successful_push_pull/main.cc:5,
//                    multigpu_nvshmem_pushpull.cu:29-85
//
// Errors throw hpp_tc::Error (the reference exits or asserts instead).
#pragma once

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tc.h"

namespace hpp_tc {

using vidType = int32_t;
using eidType = int64_t;
using AccType = uint64_t;

struct Error : std::runtime_error {
  tc_status status;
  Error(tc_status st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

inline void host_check(tc_status st) {
  if (st != TC_OK) throw Error(st, tch_last_error());
}
inline void dev_check(tc_status st) {
  if (st != TC_OK) throw Error(st, tc_last_error(nullptr));
}

class Graph {
 public:
  // Loads prefix.{meta.txt,vertex.bin,edge.bin}; use_dag orients it.
  explicit Graph(const std::string& prefix, bool use_dag = false) {
    tch_csr* h = nullptr;
    host_check(tch_read_graphminer(prefix.c_str(), use_dag ? 1 : 0, &h));
    reset(h);
  }
  static Graph rmat(int32_t n, double edge_factor, uint64_t seed) {
    tch_csr* h = nullptr;
    host_check(tch_gen_rmat(n, edge_factor, 0.57, 0.19, 0.19, seed, &h));
    return Graph(h);
  }
  static Graph er(int32_t n, double avg_degree, uint64_t seed) {
    tch_csr* h = nullptr;
    host_check(tch_gen_er(n, avg_degree, seed, &h));
    return Graph(h);
  }
  static Graph adopt(tch_csr* h) { return Graph(h); }

  vidType V() const { return view_.n_rows; }
  eidType E() const { return view_.nnz; }
  vidType num_vertices() const { return V(); }
  eidType num_edges() const { return E(); }
  vidType get_max_degree() const { return tch_csr_max_degree(h_.get()); }
  vidType out_degree(vidType v) const { return (vidType)(view_.rowptr[v + 1] - view_.rowptr[v]); }
  vidType get_degree(vidType v) const { return out_degree(v); }
  eidType edge_begin(vidType v) const { return view_.rowptr[v]; }
  eidType edge_end(vidType v) const { return view_.rowptr[v + 1]; }
  const vidType* N(vidType v) const { return view_.colidx + view_.rowptr[v]; }
  vidType N(vidType v, vidType i) const { return view_.colidx[view_.rowptr[v] + i]; }
  const eidType* out_rowptr() const { return view_.rowptr; }
  const vidType* out_colidx() const { return view_.colidx; }
  const tc_csr_view& view() const { return view_; }
  tch_csr* handle() const { return h_.get(); }

  // In place, as Graph::orientation (graph.cc:252-298).
  void orientation() {
    tch_csr* d = nullptr;
    host_check(tch_orient(&view_, &d));
    reset(d);
  }
  void sort_neighbors() { host_check(tch_sort_neighbors(h_.get())); refresh(); }
  void print_meta_data() const {
    std::printf("|V|: %d, |E|: %lld, Max Degree: %d\n", V(), (long long)E(), get_max_degree());
  }

 private:
  struct Free {
    void operator()(tch_csr* p) const { tch_csr_free(p); }
  };
  std::unique_ptr<tch_csr, Free> h_;
  tc_csr_view view_{};

  explicit Graph(tch_csr* h) { reset(h); }
  void reset(tch_csr* h) {
    h_.reset(h);
    refresh();
  }
  void refresh() { host_check(tch_csr_view(h_.get(), &view_)); }
};

// One contiguous-range part of a DAG with its halo cache, or the reference's
// vertex-induced subgraphs (graph_partition.cc:85-135).  `balance`
// (TCH_BALANCE_*) picks the halo partitions' split points; the induced
// subgraphs always use the reference's ceil(n/k) split.
class PartitionedGraph {
 public:
  PartitionedGraph(Graph* g, int nc, int balance = TCH_BALANCE_VERTICES) : g_(g), nc_(nc), balance_(balance) {
    if (nc < 1) throw Error(TC_EINVAL, "need at least one vertex chunk");
  }
  // Split points b[0..nc]: part i owns ids [b[i], b[i+1]).
  std::vector<int32_t> bounds() const {
    std::vector<int32_t> b((size_t)nc_ + 1);
    host_check(tch_partition_bounds(&g_->view(), nc_, balance_, b.data()));
    return b;
  }
  void edgecut_induced_partition1D() {
    subs_.clear();
    lb_.clear();
    le_.clear();
    idx_map_.clear();
    for (int i = 0; i < nc_; ++i) {
      tch_csr* s = nullptr;
      int32_t lb = 0, le = 0;
      int32_t* ids = nullptr;
      host_check(tch_partition_induced(&g_->view(), nc_, i, &s, &lb, &le, &ids));
      subs_.push_back(std::make_unique<Graph>(Graph::adopt(s)));
      lb_.push_back(lb);
      le_.push_back(le);
      idx_map_.emplace_back(ids, ids + subs_.back()->V());
      tch_free_ids(ids);
    }
  }
  int get_num_subgraphs() const { return (int)subs_.size(); }
  Graph* get_subgraph(int i) { return subs_.at(i).get(); }
  vidType get_local_begin(int i) const { return lb_.at(i); }
  vidType get_local_end(int i) const { return le_.at(i); }
  const std::vector<vidType>& idx_map(int i) const { return idx_map_.at(i); }

  // Owned rows of part i plus every remote list they reach (caller frees).
  tch_partition* halo_partition(int i) const {
    tch_partition* p = nullptr;
    if (balance_ == TCH_BALANCE_VERTICES) {
      host_check(tch_partition_range(&g_->view(), nc_, i, &p));
    } else {
      const std::vector<int32_t> b = bounds();
      if (i < 0 || i >= nc_) throw Error(TC_EINVAL, "part index out of range");
      host_check(tch_partition_span(&g_->view(), b[(size_t)i], b[(size_t)i + 1], &p));
    }
    return p;
  }

 private:
  Graph* g_;
  int nc_;
  int balance_;
  std::vector<std::unique_ptr<Graph>> subs_;
  std::vector<vidType> lb_, le_;
  std::vector<std::vector<vidType>> idx_map_;
};

// RAII device context.
class Device {
 public:
  explicit Device(int dev = 0) { dev_check(tc_ctx_create(dev, &ctx_)); }
  ~Device() { tc_ctx_destroy(ctx_); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
  tc_ctx* get() const { return ctx_; }

  // Count the DAG `g` on the device (binned kernels).
  AccType count(const Graph& g, tc_stats* st = nullptr) {
    tc_graph* dg = nullptr;
    dev_check(tc_graph_upload(ctx_, &g.view(), TC_CHECK_SHAPE, &dg));
    uint64_t total = 0;
    tc_status s = tc_count(ctx_, dg, &total, st);
    tc_graph_free(ctx_, dg);
    dev_check(s);
    return total;
  }

  // Count `g` and attribute its triangles (tc_count_local): per DAG edge
  // |N+(u) ∩ N+(v)|, per vertex the triangles at it, and, given split points
  // (bounds.size() - 1 parts), the reference push/pull kernel's per-PE sums.
  // Null outputs are skipped.
  AccType count_local(const Graph& g, std::vector<uint32_t>* edge_counts, std::vector<uint64_t>* vertex_counts,
                      const std::vector<int32_t>* bounds = nullptr, std::vector<uint64_t>* part_counts = nullptr) {
    tc_local lo{};
    if (edge_counts) {
      edge_counts->assign((size_t)g.E(), 0);
      lo.edge_counts = edge_counts->data();
    }
    if (vertex_counts) {
      vertex_counts->assign((size_t)g.V(), 0);
      lo.vertex_counts = vertex_counts->data();
    }
    if (bounds && part_counts && bounds->size() > 1) {
      lo.parts = (int32_t)bounds->size() - 1;
      lo.bounds = bounds->data();
      part_counts->assign(bounds->size() - 1, 0);
      lo.part_counts = part_counts->data();
    }
    tc_graph* dg = nullptr;
    dev_check(tc_graph_upload(ctx_, &g.view(), TC_CHECK_SHAPE, &dg));
    uint64_t total = 0;
    tc_status s = tc_count_local(ctx_, dg, &total, &lo, nullptr);
    tc_graph_free(ctx_, dg);
    dev_check(s);
    return total;
  }

 private:
  tc_ctx* ctx_ = nullptr;
};

// The reference's boundary.  `g` must be oriented (Graph(prefix, true) or
// orientation()); n_gpus must be 1; chunk_size is accepted and unused, as
// in the reference's GPU solvers (main.cc:25,28).  total is overwritten.
inline void TCSolver(Graph& g, uint64_t& total, int n_gpus = 1, int chunk_size = 1024) {
  (void)chunk_size;
  if (n_gpus != 1) throw Error(TC_EINVAL, "this build drives exactly one GPU (n_gpus must be 1)");
  Device dev(0);
  total = dev.count(g);
}

}  // namespace hpp_tc
