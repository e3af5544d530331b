// Internal state of the device library shared by tc_runtime.cpp,
// tc_batch.cpp, tc_server.cpp and tc_tune.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "tc.h"
#include "tc_kernels.hpp"

namespace tcr {

// Records the calling thread's error text and returns `st`.
tc_status fail(tc_status st, const std::string& msg);
// Called from a catch block: turns the exception in flight into a status, so
// that no C++ exception crosses the C ABI (a failed host allocation is
// TC_ENOMEM).
tc_status caught() noexcept;

#define TC_HIP_TRY(expr)                                                              \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return ::tcr::fail(TC_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Device allocation with an out-of-memory status instead of an abort.
tc_status dev_alloc(void** p, size_t bytes, const char* what);
void dev_free(void* p);

// Plan scratch (binning output and counters) sized for P pivots / k outputs.
struct PlanScratch {
  uint16_t* bin_of = nullptr;
  uint64_t* items = nullptr;
  uint32_t* blk = nullptr;
  int64_t* part = nullptr;
  int64_t cap = 0;
  uint32_t* counters = nullptr;
  uint64_t* outs = nullptr;
  int64_t outs_cap = 0;
  tc_status ensure(int64_t pivots, int64_t nouts);
  void release();
  tck::PlanBufs bufs() const { return tck::PlanBufs{bin_of, items, blk, part, counters, outs}; }
  // Device bytes per plan index, for footprint estimates.
  static int64_t bytes_for(int64_t pivots);
};

// Partner structure of pooled graphs (see tck::GraphDesc): partner list
// references (as many as the colidx pool has ids), partner-row slots (and
// one more row offset), rows in use per graph.
struct Partners {
  tck::ListRef* pci = nullptr;
  int64_t* vrp = nullptr;
  tck::VRow* vrow = nullptr;
  int32_t* nv = nullptr;
  int64_t bytes = 0;
  tc_status alloc(int64_t ci_slots, int64_t row_slots, int32_t ngraphs);
  static int64_t vrow_bytes(int64_t row_slots) { return (int64_t)sizeof(tck::VRow) * row_slots; }
  void release();
  void into(tck::LaunchArgs& a) const {
    a.pci = pci;
    a.vrp = vrp;
    a.vrow = vrow;
    a.nv = nv;
  }
  static int64_t bytes_for(int64_t ci_slots, int64_t row_slots, int32_t ngraphs);
};

// Buffers of a local count of one graph: the partner entries' DAG edge
// positions and partner vertices, per-edge and per-vertex counts, and
// per-part sums with their split points (a call with part sums alone
// allocates only those two; the partner rows' owners belong to the graph).
struct LocalBufs {
  int64_t* pedge = nullptr;
  int32_t* pvid = nullptr;
  uint32_t* ecount = nullptr;
  unsigned long long* vcount = nullptr;
  unsigned long long* pe = nullptr;  // [tck::kMaxParts]
  int32_t* bounds = nullptr;         // [tck::kMaxParts + 1]
  int32_t* esrc = nullptr;           // [nnz] row u of each DAG edge (u, v) | (d+(u) < d+(v)) << 31
  int64_t bytes = 0;
  void release();
};

// Temporary arrays of a partner build over RP rowptr slots and M DAG ids
// (the ids' part -- 24 bytes per id -- serves the bucketed transpose; with
// ids = 0 only the per-edge atomic path is available).
struct BuildScratch {
  char* mem = nullptr;
  int64_t cap = 0;
  char* emem = nullptr;  // the radix path's pair id positions (local rebuilds)
  int64_t ecap = 0;
  tck::BuildBufs b{};
  tc_status ensure(int64_t rp_slots, int64_t ids);
  // After ensure: room for the pairs' id positions (b.e1, b.e2) of ids ids.
  tc_status ensure_edges(int64_t ids);
  void release();
  static int64_t bytes_for(int64_t rp_slots, int64_t ids);
};

// Builds the partners of the graphs of `a` (descs on the device, pools and
// P set; RP row slots, M ids in all) on stream s.  Stream-ordered, no sync.
// loc: null, or where to write each partner entry's DAG edge position and
// partner vertex and each partner row's owner (pedge, pvid, pown).
tc_status launch_build(const tck::LaunchArgs& a, int64_t RP, int64_t M, BuildScratch& bs, hipStream_t s,
                       const LocalBufs* loc = nullptr, int32_t* pown = nullptr);

// Makes ctx's device current for the scope of one API call.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

tck::Tune to_kernel_tune(const tc_tune& t);
tc_status check_tune(const tc_tune& t);
tc_tune default_tune();

}  // namespace tcr

struct tc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // Bin groups 1.. of a count run on side streams beside group 0 on
  // `stream` (default priority); join[i]: side stream i done, fork: side
  // streams may start (events without timing).
  static constexpr int kSides = 3;
  hipStream_t side[kSides] = {};
  hipEvent_t join[kSides] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int cus = 0;
  int count_grid = 0;   // full-residency grid of the persistent kernel
  int plain_grid = 0;   // full-residency grid of the plain kernel
  tc_tune tune{};
  std::recursive_mutex mu;
  std::string name;
  tcr::PlanScratch scratch;
  uint64_t* pinned = nullptr;  // 64 B count + 64 B bin counters
  void* noalloc = nullptr;     // device address of the static scratch of tc_count_noalloc
  // device: the single-launch count's two words, used in turn; each launch
  // clears the other one for the next call (plain_dirty: a call that did
  // not finish left them unknown)
  uint64_t* plain_out = nullptr;
  int plain_turn = 0;
  bool plain_dirty = false;
  // tc_count_csr's copy of a small graph: rowptr, then colidx at the next
  // 256-byte boundary (grow-only, kept until destroy)
  uint8_t* once_dev = nullptr;
  size_t once_cap = 0;  // bytes
  int grid() const { return tune.grid > 0 ? tune.grid : count_grid; }
};

struct tc_graph {
  tc_ctx* ctx = nullptr;
  int64_t* rp = nullptr;
  int32_t* ci = nullptr;
  int32_t n = 0;
  int64_t nnz = 0;
  int32_t row_begin = 0, row_end = 0;
  int64_t edge_begin = 0, edge_end = 0;  // DAG edge positions of rows [row_begin, row_end)
  uint8_t* own = nullptr;                 // null, or per row 1 if it is a pivot row
  int32_t* pown = nullptr;  // [nv_max + 1] owner vertex of each partner row, written by every build
  int32_t max_out = 0;
  tcr::Partners partners;
  tcr::LocalBufs local;
  int32_t nv_max = 0;   // partner rows it may need (nv_max + 1 slots = plan indices)
  int32_t nv = 0;       // partner rows in use
  double build_ms = 0;  // partner build time on the device
  tck::GraphDesc* d_desc = nullptr;
  int64_t bytes = 0;
  // The binned count's plan, kept across calls: valid for plan_tune; its
  // rows per bin and tiles per bin (which size the launches) on the host.
  tcr::PlanScratch plan;
  bool plan_valid = false;
  tc_tune plan_tune{};
  int64_t plan_bins[tck::kNumBins] = {};
  int64_t plan_tiles[tck::kNumBins] = {};
};
