// Device-side layout shared by the kernels (tc_kernels.hip) and the host
// runtime (tc_runtime.cpp, tc_batch.cpp).  Plain structs only: the runtime
// never sees kernel bodies, it calls the launchers declared at the bottom.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace tck {

constexpr int kBlock = 256;   // threads per workgroup = 4 wave64s
constexpr int kWaves = kBlock / 64;

// Degree bins, in the order tiles are handed out (heaviest first).
enum Bin : int {
  BIN_GLOBAL = 0,  // d+(u) > kCapBlock: whole workgroup, pivot list read from HBM
  BIN_BLOCK = 1,   // whole workgroup, pivot list in LDS
  BIN_WAVE = 2,    // one wave64 per pivot
  BIN_G32 = 3,     // 32-lane groups
  BIN_G16 = 4,     // 16-lane groups
  BIN_G8 = 5,      // 8-lane groups
  kNumBins = 6,
  kNoBin = 0xFF    // pivot cannot close a triangle (d+ < 2 or no wedge)
};

// Largest pivot list (ids + 1) of each bin: the tree of a bin's pivot fits
// kCap words.
constexpr int kCapG8 = 32, kCapG16 = 64, kCapG32 = 256, kCapWave = 1024, kCapBlock = 4096;
// LDS words of one group's pivot-set area per bin (hash tables or tree);
// every bin's areas fill the same 4096-word region of a workgroup.
constexpr int kAreaG8 = 128, kAreaG16 = 256, kAreaG32 = 512, kAreaWave = 1024, kAreaBlock = 4096;

// Counter words of a plan (uint32).
constexpr int kCntBinCount = 0;                   // [kNumBins]
constexpr int kCntBinCursor = kNumBins;           // [kNumBins]
constexpr int kCntTileHead = 2 * kNumBins;        // [kNumBins] tile queue heads
constexpr int kNumCounters = 3 * kNumBins + 2;

// One graph of a (possibly single-graph) count.  Its CSR lives at
// rowptr_pool + rp_off (rowptr values relative to colidx_pool + ci_off);
// its pivots are rows [row_begin, row_end), numbered from plan_base in the
// plan index space shared by all graphs of the launch.
struct GraphDesc {
  int64_t rp_off;
  int64_t ci_off;
  int64_t plan_base;
  int32_t row_begin;
  int32_t row_end;
};

struct Tune {
  int32_t t_g8, t_g16, t_g32, t_wave, rounds;
  int32_t tree_only;   // 1: pivot sets are always LDS trees
  int32_t hash_kicks;  // evictions per hash insert before falling back to the tree
};

// Per-planning-block bin counts, then their item offsets (kBlkStride words
// per block of kBlock consecutive plan indices).
constexpr int kBlkStride = 8;

struct PlanBufs {
  uint8_t* bin_of;      // [P] bin of each pivot
  uint64_t* items;      // [P] (graph << 32 | row), ordered by (bin, plan index)
  uint32_t* blk;        // [ceil(P/kBlock) * kBlkStride]
  uint32_t* counters;   // [kNumCounters]
  uint64_t* outs;       // [ngraphs] u64 counts
};

struct LaunchArgs {
  const GraphDesc* descs;
  int32_t ngraphs;
  int64_t P;            // total pivots over all graphs
  const int64_t* rp;    // rowptr pool
  const int32_t* ci;    // colidx pool
  Tune tune;
  PlanBufs bufs;
};

// Launchers (stream-ordered; return the launch status).
hipError_t launch_zero(void* p, int64_t words32, hipStream_t s);
hipError_t launch_plan(const LaunchArgs& a, hipStream_t s);
// One full-residency persistent launch per bin (grid > 0 overrides the size).
hipError_t launch_count(const LaunchArgs& a, int cus, int grid, hipStream_t s);
hipError_t launch_plain(const int64_t* rp, const int32_t* ci, int32_t row_begin, int32_t row_end,
                        uint64_t* out, int grid, hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, int64_t bytes, int grid, hipStream_t s);
// Resident workgroups per CU of the persistent and plain kernels.
int count_blocks_per_cu();
int plain_blocks_per_cu();

}  // namespace tck
