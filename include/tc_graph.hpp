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
#include <cstdlib>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
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

// Read-only view of one adjacency list: the part of GraphMiner's VertexSet
// (absent from the reference tree) that Graph::N(v) callers use -- size,
// indexing, iteration, data.
class VertexSet {
 public:
  VertexSet(const vidType* p, vidType n) : p_(p), n_(n) {}
  vidType size() const { return n_; }
  vidType operator[](vidType i) const { return p_[i]; }
  const vidType* data() const { return p_; }
  const vidType* begin() const { return p_; }
  const vidType* end() const { return p_ + n_; }

 private:
  const vidType* p_;
  vidType n_;
};

using vlabel_t = uint8_t;
using idx_t = int64_t;  // METIS idx_t: the part array is 8 bytes per id (graph_partition.cc:438)

class Graph {
 public:
  // The reference's constructor (graph.cc:4-142).  Built with USE_METIS, as
  // the reference's per-PE target is (successful_push_pull/Makefile:46-48),
  // it loads PE my_id's METIS part files <prefix><my_id>.{meta.txt,
  // vertex.bin, edge.bin, vertex_list.bin, part.bin} (graph.cc:16-23,33-34,
  // 56-59): the PE's DAG rows with global ids, its vertex_list and the part
  // array.  Otherwise it loads the GraphMiner files prefix.{meta.txt,
  // vertex.bin, edge.bin} and orients them when use_dag.  directed and
  // need_reverse change nothing on this path; vertex labels are read from
  // prefix.vlabel.bin when use_vlabel (one byte per vertex); edge labels and
  // bipartite graphs are not on the triangle-counting path and throw.
  explicit Graph(const std::string& prefix, bool use_dag = false, bool directed = false, bool use_vlabel = false,
                 bool use_elabel = false, bool need_reverse = false, bool bipartite = false, int my_id = 0) {
    (void)directed;
    (void)need_reverse;
    if (use_elabel || bipartite)
      throw Error(TC_EINVAL, "edge labels and bipartite graphs are not supported by this triangle-counting build");
#ifdef USE_METIS
    // part files already hold DAG rows (the partitioner orients first)
    if (use_dag) throw Error(TC_EINVAL, "METIS part files are already oriented: use_dag must be false");
    load_metis(prefix, my_id);
    if (use_vlabel) load_vlabels(prefix + std::to_string(my_id));
#else
    (void)my_id;
    tch_csr* h = nullptr;
    host_check(tch_read_graphminer(prefix.c_str(), use_dag ? 1 : 0, &h));
    reset(h);
    if (use_vlabel) load_vlabels(prefix);
#endif
  }
  // PE my_id's METIS part files, whatever USE_METIS says.
  static Graph metis(const std::string& prefix, int my_id) {
    Graph g;
    g.load_metis(prefix, my_id);
    return g;
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
  Graph(Graph&&) = default;
  Graph& operator=(Graph&&) = default;

  // meta information (graph.h:74-86)
  vidType V() const { return view_.n_rows; }
  eidType E() const { return view_.nnz; }
  vidType num_vertices() const { return V(); }
  eidType num_edges() const { return E(); }
  vidType size() const { return V(); }
  eidType sizeEdges() const { return E(); }
  vidType get_max_degree() const { return m_ ? max_degree_ : tch_csr_max_degree(h_.get()); }
  std::string get_name() const { return name_; }
  std::string get_inputfile_path() const { return path_; }

  // topology (graph.h:88-104); N(v) views the list, as GraphMiner's VertexSet
  vidType out_degree(vidType v) const { return (vidType)(view_.rowptr[v + 1] - view_.rowptr[v]); }
  vidType get_degree(vidType v) const { return out_degree(v); }
  eidType edge_begin(vidType v) const { return view_.rowptr[v]; }
  eidType edge_end(vidType v) const { return view_.rowptr[v + 1]; }
  const vidType* adj_ptr(vidType v) const { return view_.colidx + view_.rowptr[v]; }
  VertexSet N(vidType v) const { return VertexSet(adj_ptr(v), out_degree(v)); }
  vidType N(vidType v, vidType i) const { return view_.colidx[view_.rowptr[v] + i]; }
  vidType getEdgeDst(eidType e) const { return view_.colidx[e]; }
  const eidType* out_rowptr() const { return view_.rowptr; }
  const vidType* out_colidx() const { return view_.colidx; }
  const tc_csr_view& view() const { return view_; }
  tch_csr* handle() const { return h_.get(); }

  // |N(v) ∩ N(u)| by a merge on the host (graph.cc:376-391); the labelled
  // form counts only common ids w with label(w) == label and needs labels
  // (use_vlabel).
  vidType intersect_num(vidType v, vidType u) const { return merge_count(v, u, nullptr, 0); }
  vidType intersect_num(vidType v, vidType u, vlabel_t label) const {
    if (vlabels_.empty()) throw Error(TC_EINVAL, "intersect_num with a label needs vertex labels (use_vlabel)");
    return merge_count(v, u, vlabels_.data(), label);
  }
  bool has_vlabel() const { return !vlabels_.empty(); }
  vlabel_t get_vlabel(vidType v) const { return vlabels_.at((size_t)v); }

  // METIS mode (graph.h:174-182): this PE's id and part files, its owned
  // ids (vertex_list, n_real of them) and the part of every id.  For a
  // GraphMiner graph n_real is V() unless set.
  bool is_metis() const { return m_ != nullptr; }
  const tch_metis* metis_handle() const { return m_.get(); }
  int my_id() const { return my_id_; }
  const std::string& metis_prefix() const { return prefix_; }
  void set_n_real_vertices(vidType n) { n_real_ = n; }
  vidType get_n_real_vertices() const { return n_real_ >= 0 ? n_real_ : V(); }
  vidType get_max_subg_nv() const { return max_subg_nv_; }
  eidType get_max_subg_ne() const { return max_subg_ne_; }
  vidType* get_vertex_list() { return vertex_list_.data(); }
  idx_t* get_metis_part() { return part_.data(); }

  // In place, as Graph::orientation (graph.cc:252-298).
  void orientation() {
    if (m_) throw Error(TC_EINVAL, "a METIS part holds one PE's rows; orient the whole graph before partitioning");
    tch_csr* d = nullptr;
    host_check(tch_orient(&view_, &d));
    reset(d);
  }
  void sort_neighbors() {
    if (m_) throw Error(TC_EINVAL, "a METIS part is loaded read-only");
    host_check(tch_sort_neighbors(h_.get()));
    refresh();
  }
  void print_meta_data() const {
    std::printf("|V|: %d, |E|: %lld, Max Degree: %d\n", V(), (long long)E(), get_max_degree());
  }

 private:
  struct Free {
    void operator()(tch_csr* p) const { tch_csr_free(p); }
  };
  struct FreeMetis {
    void operator()(tch_metis* p) const { tch_metis_free(p); }
  };
  std::unique_ptr<tch_csr, Free> h_;
  std::unique_ptr<tch_metis, FreeMetis> m_;
  tc_csr_view view_{};
  std::string name_, path_, prefix_;
  int my_id_ = 0;
  vidType n_real_ = -1;
  vidType max_degree_ = 0;
  vidType max_subg_nv_ = 0;
  eidType max_subg_ne_ = 0;
  std::vector<vidType> vertex_list_;
  std::vector<idx_t> part_;
  std::vector<vlabel_t> vlabels_;

  Graph() = default;
  explicit Graph(tch_csr* h) { reset(h); }
  void reset(tch_csr* h) {
    h_.reset(h);
    refresh();
  }
  void refresh() { host_check(tch_csr_view(h_.get(), &view_)); }
  void load_metis(const std::string& prefix, int my_id) {
    tch_metis* m = nullptr;
    host_check(tch_metis_read(prefix.c_str(), my_id, &m));
    m_.reset(m);
    tch_metis_view mv;
    host_check(tch_metis_get(m, &mv));
    view_ = tc_csr_view{mv.rowptr, mv.colidx, mv.n, mv.nnz};
    prefix_ = prefix;
    my_id_ = my_id;
    n_real_ = mv.n_real;
    max_degree_ = (vidType)mv.max_degree;
    max_subg_nv_ = (vidType)mv.max_subg_nv;
    max_subg_ne_ = mv.max_subg_ne;
    vertex_list_.assign(mv.vertex_list, mv.vertex_list + mv.n_real);
    part_.assign(mv.part, mv.part + mv.n);
    // graph.cc:10-20: the name is the directory above .../metis/
    const size_t slash = prefix.rfind('/');
    path_ = slash == std::string::npos ? std::string() : prefix.substr(0, slash);
    const size_t up = path_.rfind('/');
    const std::string parent = up == std::string::npos ? std::string() : path_.substr(0, up);
    const size_t up2 = parent.rfind('/');
    name_ = up2 == std::string::npos ? parent : parent.substr(up2 + 1);
  }
  void load_vlabels(const std::string& file_prefix) {
    std::FILE* f = std::fopen((file_prefix + ".vlabel.bin").c_str(), "rb");
    if (!f) throw Error(TC_EIO, "cannot open " + file_prefix + ".vlabel.bin");
    vlabels_.assign((size_t)V(), 0);
    const size_t got = V() > 0 ? std::fread(vlabels_.data(), 1, vlabels_.size(), f) : 0;
    std::fclose(f);
    if (got != vlabels_.size()) throw Error(TC_EIO, file_prefix + ".vlabel.bin must hold one byte per vertex");
  }
  vidType merge_count(vidType v, vidType u, const uint8_t* labels, vlabel_t label) const {
    const int64_t c = tch_intersect_num(&view_, v, u, labels, label);
    if (c < 0) throw Error(TC_EINVAL, tch_last_error());
    return (vidType)c;
  }
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
    uint64_t total = 0;
    dev_check(tc_count_csr(ctx_, &g.view(), TC_CHECK_SHAPE, &total, st));
    return total;
  }

  // Count only the rows u of `g` with owned[u] != 0 (tc_graph_upload_rows).
  AccType count_rows(const Graph& g, const uint8_t* owned, tc_stats* st = nullptr) {
    tc_graph* dg = nullptr;
    dev_check(tc_graph_upload_rows(ctx_, &g.view(), owned, TC_CHECK_SHAPE, &dg));
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

// The PE id and node-local rank of this process.  The reference starts
// NVSHMEM, binds the PE to its node-local GPU and returns the PE id
// (successful_push_pull/multigpu_nvshmem_pushpull.cu:20-26).  Here a
// process drives one GPU and needs no communication library: the ids come
// from the launcher's environment -- TC_PE / TC_LOCAL_PE, else Open MPI's,
// MPICH's, Slurm's or torchrun's variables -- and are 0 when none is set.
inline int env_first_int(std::initializer_list<const char*> names) {
  for (const char* n : names)
    if (const char* v = std::getenv(n))
      if (*v) return std::atoi(v);
  return 0;
}
inline int& pe_device() {
  static int dev = 0;
  return dev;
}
inline int initialize_nvshmem() {
  pe_device() = env_first_int({"TC_LOCAL_PE", "OMPI_COMM_WORLD_LOCAL_RANK", "MPI_LOCALRANKID", "SLURM_LOCALID",
                               "LOCAL_RANK"});
  return env_first_int({"TC_PE", "OMPI_COMM_WORLD_RANK", "PMI_RANK", "SLURM_PROCID", "RANK"});
}

// The reference's boundary (successful_push_pull/main.cc:5,33;
// multigpu_nvshmem_pushpull.cu:29-85).  total is overwritten; chunk_size is
// accepted and unused, as in the reference's GPU solvers (main.cc:25,28).
//  * A graph loaded whole (Graph(prefix, use_dag) etc.) must be oriented;
//    n_gpus must be 1: the count of the whole DAG on this process's GPU.
//  * A METIS part (USE_METIS build, or Graph::metis): n_gpus is the PE
//    count k, as in the reference's per-PE program.  The count is this PE's
//    share: every pivot u of its vertex_list (bs_warp_vertex_nvshmem.cuh:
//    25-28), over all v in N+(u), |N+(u) ∩ N+(v)| -- local v and remote v
//    alike, the remote lists read from their owners' part files (the lists
//    the reference pulls, :110-136), on this process's one GPU.  Summed over
//    the k PEs it is the graph's total.  (The reference splits a PE's sum by
//    its push/pull rule instead, :127-136; the totals agree.)
inline void tc_solve(Graph& g, uint64_t& total, int n_gpus, int chunk_size) {
  (void)chunk_size;
  if (!g.is_metis()) {
    if (n_gpus != 1)
      throw Error(TC_EINVAL, "this library drives exactly one GPU per process: a whole graph needs n_gpus = 1 (METIS parts take the PE count)");
    Device dev(pe_device());
    total = dev.count(g);
    return;
  }
  const int k = n_gpus, me = g.my_id();
  if (k < 1 || me < 0 || me >= k) throw Error(TC_EINVAL, "n_gpus must be the PE count and exceed this PE's id");
  struct FreeMetis {
    void operator()(tch_metis* p) const { tch_metis_free(p); }
  };
  std::vector<std::unique_ptr<tch_metis, FreeMetis>> held((size_t)k);
  std::vector<const tch_metis*> pes((size_t)k, nullptr);
  for (int p = 0; p < k; ++p) {  // this PE's files are loaded already; the others' are read here
    if (p == me) {
      pes[(size_t)p] = g.metis_handle();
      continue;
    }
    tch_metis* m = nullptr;
    host_check(tch_metis_read(g.metis_prefix().c_str(), p, &m));
    held[(size_t)p].reset(m);
    pes[(size_t)p] = m;
  }
  tch_csr* csr = nullptr;
  std::vector<uint8_t> owned((size_t)g.V());
  host_check(tch_metis_assemble(pes.data(), k, me, &csr, owned.data()));
  Graph assembled = Graph::adopt(csr);
  Device dev(pe_device());
  total = dev.count_rows(assembled, owned.data());
}

// TCSolver(g, total, n_gpus, chunk_size) -> tc_solve.  A template, so that
// a caller's own non-template TCSolver of the same signature (the
// reference's prototype, main.cc:5, defined by compat/graph.h) wins overload
// resolution over this one, which argument-dependent lookup also finds.
template <class G>
inline void TCSolver(G& g, uint64_t& total, int n_gpus = 1, int chunk_size = 1024) {
  static_assert(std::is_same<G, Graph>::value, "TCSolver counts a hpp_tc::Graph");
  tc_solve(g, total, n_gpus, chunk_size);
}

}  // namespace hpp_tc
