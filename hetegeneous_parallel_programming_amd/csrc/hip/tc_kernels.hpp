// Device-side layout shared by the kernels (tc_kernels.hip) and the host
// runtime (tc_runtime.cpp, tc_batch.cpp).  Plain structs only: the runtime
// never sees kernel bodies, it calls the launchers declared at the bottom.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "tc_host.h"

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

// Partner rows per tile of each bin (rounds = rows per group in light bins).
__host__ __device__ inline int tile_pivots(int bin, int rounds) {
  switch (bin) {
    case BIN_GLOBAL:
    case BIN_BLOCK: return 1;
    case BIN_WAVE: return kBlock / 64;
    case BIN_G32: return (kBlock / 32) * rounds;
    case BIN_G16: return (kBlock / 16) * rounds;
    default: return (kBlock / 8) * rounds;
  }
}

// Counter words of a plan (uint32).
constexpr int kCntBinCount = 0;                   // [kNumBins]
constexpr int kCntBinCursor = kNumBins;           // [kNumBins]
constexpr int kCntTileHead = 2 * kNumBins;        // [kNumBins] tile queue heads
constexpr int kCntFallback = 3 * kNumBins + 2;    // hash builds that fell back to the tree
constexpr int kNumCounters = 3 * kNumBins + 3;

// Partner rows.  Every DAG edge (u, v) is counted from the endpoint whose
// out-list is longer: its "owner" x keeps N+(x) as the LDS set and its
// "partner" p has N+(p) streamed through it, so each edge streams
// min(d+(u), d+(v)) ids instead of d+(v).  Owner of (u, v): u when
// d+(v) <= d+(u) (partner v), else v (partner u).  The partners of one
// owner x are consecutive in the partner-id array; x's list is cut into
// ceil(count / kPartnersPerRow) "partner rows" (vrow[k] = x, partner ids
// pci[vrp[k] .. vrp[k+1])), the unit of work of the count.
constexpr int kPartnersPerRow = TCH_PARTNERS_PER_ROW;

// One graph of a (possibly single-graph) count.
//   DAG: rowptr at rp_off of the rowptr pool (values relative to ci_off of
//        the colidx pool), n_rows rows; only rows [own_begin, own_end) have
//        their out-edges counted (a partition's owned range), other rows are
//        lists read through them (halo).
//   Partners: the graph's partner list references at pci_off of the
//        partner pool (as many slots as the DAG has ids).  Its partner rows
//        take slots [pr_base, pr_base + nv_max + 1) of the launch's VRow and
//        row-offset arrays; those slots are also the plan indices.
//   Build: the graph's n_rows + 1 slots of a partner build's per-row
//        scratch start at row_base, and its ids' slots of the build's
//        per-id scratch at eb (graphs in order, contiguous).
// List references (set and partners) and row offsets are absolute in their
// pools, so planning and counting never read this descriptor.
struct GraphDesc {
  int64_t rp_off;
  int64_t ci_off;
  int64_t pci_off;
  int64_t row_base;
  int64_t pr_base;
  int32_t n_rows;
  int32_t nv_max;
  int32_t own_begin;
  int32_t own_end;
  int64_t eb;
};

// Partner rows a graph of n rows and m DAG ids may need (it takes this
// many slots plus one).
inline int32_t partner_rows_max(int64_t n, int64_t m) {
  return (int32_t)(n + m / kPartnersPerRow + 1);
}

struct Tune {
  int32_t t_g8, t_g16, t_g32, t_wave, rounds;
  int32_t tree_only;   // 1: pivot sets are always LDS trees; 2: for sub-wave groups
  int32_t hash_kicks;  // evictions per hash insert before falling back to the tree
};

struct PlanBufs {
  uint16_t* bin_of;     // [P] work class << 3 | bin of each pivot (kNoBin: none)
  uint64_t* items;      // [P] (graph << 32 | partner row), ordered by (bin, plan index)
  uint32_t* blk;        // [kNumBins * nblk + 1] per-(bin, planning block) counts -> offsets
  int64_t* part;        // [scan_partials(kNumBins * nblk + 1)] scan partials
  uint32_t* counters;   // [kNumCounters]
  uint64_t* outs;       // [ngraphs] u64 counts
};

// A list of the DAG as one word: offset in the colidx pool in the high 39
// bits, length in the low 25 (out-degrees are below 2^25 by the upload
// checks).
typedef uint64_t ListRef;
constexpr int kLenBits = 25;
__host__ __device__ inline ListRef list_ref(int64_t start, int64_t len) {
  return ((uint64_t)start << kLenBits) | (uint64_t)len;
}
__host__ __device__ inline int64_t list_start(ListRef r) { return (int64_t)(r >> kLenBits); }
__host__ __device__ inline int32_t list_len(ListRef r) { return (int32_t)(r & ((1u << kLenBits) - 1)); }

// A partner row: the owner's list (the set; 0 for an unused slot), the
// total length of its partners' lists (saturated at 2^32 - 1; only compared
// with the bin thresholds) and its graph.
struct VRow {
  ListRef set;
  uint32_t work;
  int32_t g;
};

// Outputs of a local count (one graph only).  pedge / pvid: DAG edge
// position and partner vertex of every partner entry (parallel to the
// partner pool); pown: owner vertex of every partner row (per plan index);
// ecount: per DAG edge position, |N+(u) ∩ N+(v)| of edge (u, v); vcount: per
// vertex, the triangles with a corner there.  All null in a plain count.
// Attribution modes of a local count: none (the plain count); every
// triangle to its DAG edge, its three corners (slot hit counters for the
// third) and, through the edges, the parts of a split (full); to its DAG
// edge only (edges); or only to the part of a split that computes it, one
// sum per partner row (parts: no per-edge or per-vertex output).
constexpr int kLocalNone = 0, kLocalFull = 1, kLocalEdges = 2, kLocalParts = 3;

struct LocalOut {
  const int64_t* pedge;
  const int32_t* pvid;
  const int32_t* pown;
  uint32_t* ecount;
  unsigned long long* vcount;
  const int32_t* bounds = nullptr;  // kLocalParts: split points [parts + 1]
  int32_t parts = 0;
  uint64_t* pe = nullptr;           // kLocalParts: per-part sums [parts], zeroed by the caller
};

struct LaunchArgs {
  const GraphDesc* descs;
  int32_t ngraphs;
  int64_t P;            // partner-row slots = plan indices (sum of nv_max + 1)
  const int64_t* rp;    // DAG rowptr pool
  const int32_t* ci;    // DAG colidx pool
  const ListRef* pci;   // partner pool: each partner's list
  const int64_t* vrp;   // [P + 1] partner-row offsets into the partner pool
  const VRow* vrow;     // [P] partner rows
  const int32_t* nv;    // [ngraphs] partner rows in use
  Tune tune;
  PlanBufs bufs;
  LocalOut local;
  const uint8_t* own;   // one graph only: null, or per row 1 if it is a pivot row
  int32_t wide_ids;     // 1 when ids reach 2^24: hash sets fold the high id bits into the bucket
  int32_t gbase;        // partner build of a segment of graphs: index of descs[0] among all graphs
};

// Ids at or above this need the folded hash (the bucket multiply reads 24 bits).
constexpr int64_t kNarrowIds = (int64_t)1 << 24;

// Planning blocks of a launch over P plan indices.
inline int64_t plan_blocks(int64_t P) { return (P + 255) / 256; }

// Scratch of a partner build over RP row slots (the graphs' n_rows + 1
// slots, concatenated in graph order from row_base 0).
// Partners a row owns through in-edges are counted in kXcds stripes:
// workgroup i updates stripe i mod kXcds, and the dispatcher places it on
// XCD i mod kXcds, so each stripe's counters and cursors are touched by one
// XCD only (a hub's in-edges then spread over kXcds cache lines).  Both
// build passes use the same grid, so an edge falls in the same stripe in
// each whatever the placement.
constexpr int kXcds = 8;

struct BuildBufs {
  int32_t* c1;          // [RP] atomic path: partners owned through out-edges
  int32_t* c2;          // [kXcds * RP] atomic path: partners owned through in-edges, per stripe;
                        // radix path: pairs per owner of the big fine buckets
  unsigned long long* cur;  // [kXcds * RP] atomic path: fill cursor of each stripe;
                            // radix path: pool cursor of each owner of the big fine buckets
  int64_t* off;         // [RP] partner offsets (exclusive scan over the row slots)
  int32_t* kc;          // [RP] partner rows per DAG row
  int64_t* voff;        // [RP] partner-row offsets (pool-global exclusive scan)
  int64_t* part;        // [scan_partials(max(RP, ...))] scan partials
  int64_t* pedge;       // null, or [partner pool]: DAG edge position of each partner entry
  int32_t* pvid;        // with pedge: [partner pool] partner vertex of each entry
  int32_t* pown;        // null, or (one graph) [P] owner vertex of each partner row
  // Radix transpose (null when the scratch has no per-id part).  Build space:
  // M ids, graph g's at eb; row slots: RP.
  int32_t* gid;         // [RP] graph of each row slot; ~graph when the row is not split
  int64_t* erow;        // [RP + 1] first build-space id of each row slot; erow[RP] = M
  ListRef* lref;        // [RP] each row slot's own list
  uint16_t* deg16;      // [RP] out-degree of each row slot, saturated at 65535
  int32_t* key;         // [M] owner of each id's pair (-1: none); after level 2, fine-bucket order
  ListRef* ref;         // [M] partner list of each pair, in the order of key
  int32_t* pkey;        // [M] level-1 (coarse-bucket) order of key
  ListRef* pref;        // [M] and of ref
  int64_t* bq;          // [nmp + 1] rows consumed before each split block
  uint32_t* cnt1;       // [kCoarse * nmp + 1] pairs per (coarse bucket, split block) -> positions
  uint32_t* cnt2;       // [kFine * nt2 + 1] pairs per (fine digit, level-2 tile) -> positions
  int64_t* tstart;      // [kCoarse + 1] first level-2 tile of each coarse bucket
  int64_t* chstart;     // [kMaxFineBuckets + 1] first chunk of each big fine bucket
  int64_t* fst;         // [kMaxFineBuckets + 1] first pair of each fine bucket
  unsigned long long* wq;   // [RP] radix path: each owner's work (its partners' list lengths); else null
  // Radix path with pedge (a local rebuild of one graph): each pair's DAG id
  // position rides along the two levels (e1, e2) and lands in pedge beside
  // the partner's list; esrc[e] = the row of DAG id e gives the partner
  // vertices (pvid).  Either path of a local rebuild (the radix split, or
  // the atomic fill) sets bit 31 of esrc[e] where the edge's head owns it.
  // Null otherwise.
  uint32_t* e1;         // [M] level-1 order of the pairs' id positions
  uint32_t* e2;         // [M] level-2 (fine-bucket) order
  int32_t* esrc;        // [M]
  int64_t rows;         // capacity: row slots
  int64_t ids;          // capacity: ids
};
int64_t scan_partials(int64_t n);

// Radix transpose geometry.  Owner slots fall in fine buckets of 2^fs slots
// (at most kMaxFineBuckets), 64 fine buckets to a coarse bucket.  Split
// blocks take kMP merge-path items (row ends and ids), level-2 tiles kTile
// pairs.  A fine bucket of at most `cap` pairs is sorted whole in LDS; a
// bigger one goes by chunks of `chunk` pairs.  Dynamic LDS of the final
// passes stays within kRadixLds.
constexpr int kMP = 4096;
constexpr int kMPPer = kMP / kBlock;
constexpr int kTile = 4096;
constexpr int kTilePer = kTile / kBlock;
constexpr int kCoarse = 160;
constexpr int kFine = 64;
constexpr int kMaxFineBuckets = kCoarse * kFine;
constexpr int kMinFineShift = 5;
constexpr int kMaxFineShift = 11;
constexpr int kMaxFineOwners = 1 << kMaxFineShift;
constexpr int64_t kFineCapMax = 8192;
constexpr int64_t kChunkMax = 4096;
constexpr int kFinPer = (int)(kFineCapMax / kBlock);  // pairs per thread of a small bucket
constexpr int kChunkPer = (int)(kChunkMax / kBlock);  // pairs per thread of a chunk
constexpr int64_t kRadixLds = 64512;  // 63 KB: room for the kernels' static LDS within 64 KB
inline size_t radix_final_lds(int fs, int cap) { return 20 * ((size_t)1 << fs) + 4 * (size_t)cap; }
inline size_t radix_place_lds(int fs, int chunk) { return 24 * ((size_t)1 << fs) + 4 * (size_t)chunk; }
struct RadixGeom {
  int fs, nf, nc;
  int64_t nmp, nt2, nch;
  int cap, chunk;
};
// False when RP or M is beyond the radix path (the atomic path then builds).
inline bool radix_geom(int64_t RP, int64_t M, RadixGeom& x) {
  if (RP <= 0 || RP >= ((int64_t)1 << 31) || M < 0 || M >= ((int64_t)1 << 32)) return false;
  int fs = kMinFineShift;
  while (fs <= kMaxFineShift && ((RP + ((int64_t)1 << fs) - 1) >> fs) > kMaxFineBuckets) ++fs;
  if (fs > kMaxFineShift) return false;
  const int64_t no = (int64_t)1 << fs;
  x.fs = fs;
  x.nf = (int)((RP + no - 1) >> fs);
  x.nc = (x.nf + kFine - 1) / kFine;
  x.nmp = (RP + M + kMP - 1) / kMP;
  x.nt2 = (M + kTile - 1) / kTile + x.nc;
  const int64_t cap = (kRadixLds - 20 * no) / 4 / 256 * 256, chunk = (kRadixLds - 24 * no) / 4 / 256 * 256;
  x.cap = (int)(cap < kFineCapMax ? cap : kFineCapMax);
  x.chunk = (int)(chunk < kChunkMax ? chunk : kChunkMax);
  x.nch = (M + x.chunk - 1) / x.chunk + x.nf;
  return true;
}
// Upper bounds of the per-build arrays for up to `rows` slots and `ids` ids.
inline int64_t radix_split_blocks(int64_t rows, int64_t ids) { return (rows + ids + kMP - 1) / kMP; }
inline int64_t radix_tiles2(int64_t ids) { return (ids + kTile - 1) / kTile + kCoarse; }

// Builds the partner lists, rows and counts of every graph of `a` (descs,
// rp, ci, P as for a count; pci/vrp/vrow/nv are written); RP = sum of
// n_rows + 1 over the graphs, M = sum of their ids.  With the per-id
// scratch (b.key) and no local outputs (b.pedge null) the in-edge partners
// are placed by a bucketed transpose; otherwise by per-edge atomics.
hipError_t launch_partners(const LaunchArgs& a, int64_t RP, int64_t M, const BuildBufs& b, hipStream_t s);

// Launchers (stream-ordered; return the launch status).
hipError_t launch_zero(void* p, int64_t words32, hipStream_t s);

// One request's CSR on its way from pinned host memory (src, mapped for the
// device) to its extent of the device pool (dst): `bytes`, a multiple of 4,
// both ends 16-byte aligned; chunks [first, first + ceil(bytes / kStageChunk))
// of the staging copy move it.
struct StageCopy {
  const char* src;
  char* dst;
  int64_t bytes;
  int64_t first;
};
constexpr int64_t kStageChunk = (int64_t)64 << 10;  // bytes per chunk of the staging copy
inline int64_t stage_blocks(int64_t bytes) { return (bytes + kStageChunk - 1) / kStageChunk; }
// All copies of a launch in one kernel (`chunks` = the sum of their chunks;
// copies on the device, ordered by first), on at most kStageGrid workgroups.
constexpr int64_t kStageGrid = 128;
hipError_t launch_stage_copy(const StageCopy* copies, int32_t n, int64_t chunks, hipStream_t s);
// Re-arms a kept plan for another count: tile queue heads and the fallback
// counter to 0 (bin counts kept), outs[0..nouts) to 0.
hipError_t launch_rearm(uint32_t* counters, uint64_t* outs, int32_t nouts, hipStream_t s);
hipError_t launch_plan(const LaunchArgs& a, hipStream_t s);
// One full-residency persistent launch per bin (grid > 0 overrides the size).
hipError_t launch_count(const LaunchArgs& a, int cus, int grid, hipStream_t s);
// Bins first..last only, each at full residency.
hipError_t launch_count_bins(const LaunchArgs& a, int cus, int grid, int first, int last, hipStream_t s);
// Same, with each bin's tile count known on the host (tiles[kNumBins]):
// empty bins are skipped and no grid exceeds its bin's tiles.
hipError_t launch_count_sized(const LaunchArgs& a, int cus, int grid, const int64_t* tiles, hipStream_t s);
// Bins order[0..n) in that order, each at full residency or, with tiles
// (tiles[kNumBins], known on the host), sized as launch_count_sized does.
hipError_t launch_count_order(const LaunchArgs& a, int cus, int grid, const int* order, int n,
                              const int64_t* tiles, hipStream_t s);

// Stream groups of the bins: the groups run side by side (group 0 on the
// library stream, the others on side streams), each group's bins one after
// another in the order written.  Persistent workgroups hold their CUs until
// their bin drains, so a bin alone on the device at the end of a count runs
// at its own issue rate: the groups start the sub-wave bins, whose waves
// stall the most, beside the others and leave wave-wide work for last.
#ifndef TC_BIN_ORDER
#define TC_BIN_ORDER "01|2|3|45"
#endif
constexpr int kBinGroups = 4;
struct BinGroups {
  int n = 0;                        // groups in use
  int len[kBinGroups] = {};         // bins per group
  int bin[kBinGroups][kNumBins] = {};
  bool ok = false;                  // every bin exactly once, at most kBinGroups groups
};
constexpr BinGroups parse_bin_groups(const char* t) {
  BinGroups g{};
  int seen = 0, cur = 0;
  bool ok = true;
  for (int i = 0; t[i] != 0; ++i) {
    if (t[i] == '|') {
      ok = ok && g.len[cur] > 0;
      ++cur;
      ok = ok && cur < kBinGroups;
      if (!ok) break;
      continue;
    }
    const int b = t[i] - '0';
    ok = ok && b >= 0 && b < kNumBins && !(seen >> b & 1);
    if (!ok) break;
    seen |= 1 << b;
    g.bin[cur][g.len[cur]++] = b;
  }
  g.n = cur + 1;
  g.ok = ok && g.len[cur] > 0 && seen == (1 << kNumBins) - 1;
  return g;
}
constexpr BinGroups kBinOrder = parse_bin_groups(TC_BIN_ORDER);
static_assert(kBinOrder.ok, "TC_BIN_ORDER: bins 0-5 once each, in at most 4 '|'-separated groups");
// The groups of the edges-only local count (without vertex counts).  The
// workgroup bin goes first on its own: the wave bin then waits for CUs
// behind it and the 32-/16-lane bins, and runs at the end with the whole
// device, beside the short 8-lane bin (RMAT-22 edge and part counts on
// MI355X: 11.8 ms, against 12.2-12.6 ms with TC_BIN_ORDER, 11.9 ms with the
// 8-lane bin after the wave bin).  The full local count keeps TC_BIN_ORDER,
// which measured ~0.25 ms faster for it.
#ifndef TC_EDGES_BIN_ORDER
#define TC_EDGES_BIN_ORDER "1|02|3|45"
#endif
constexpr BinGroups kEdgesBinOrder = parse_bin_groups(TC_EDGES_BIN_ORDER);
static_assert(kEdgesBinOrder.ok, "TC_EDGES_BIN_ORDER: bins 0-5 once each, in at most 4 '|'-separated groups");
// own: null, or per row 1 if it is a pivot row (rows outside
// [row_begin, row_end) are never pivots).  out must hold 0; zero_next (null,
// or another word) is set to 0 by the kernel.
hipError_t launch_plain(const int64_t* rp, const int32_t* ci, int32_t row_begin, int32_t row_end,
                        const uint8_t* own, uint64_t* out, int grid, hipStream_t s, uint64_t* zero_next = nullptr);
// Count of one graph that also fills a.local: same plan, local variants of
// bins first..last.  mode kLocalFull: ecount and vcount (zeroed by the
// caller); kLocalEdges: ecount only (the kernels keep no slot hit counters:
// less LDS, more waves per SIMD); kLocalParts: pe only (bounds, parts, pe
// set; the plain count's kernels, each row's hits added to its owner's part).
hipError_t launch_count_local(const LaunchArgs& a, int cus, int first, int last, int mode, hipStream_t s);
// Per-graph table of launch_local_parts over rows [0, n): esrc[e] = the row
// u of DAG edge e = (u, v); the local rebuild (launch_build with pedge and
// esrc) then sets bit 31 where v owns the edge, i.e. d+(u) < d+(v).
hipError_t launch_local_prep(const int64_t* rp, int64_t n, int32_t* esrc, hipStream_t s);
// After launch_count_local: sums the ecount of every edge in [e0, e1) (the
// counted rows' edges; uncounted ones hold 0) into pe[owner] by the
// reference push/pull rule over the split bounds[0..parts] (1 <= parts <=
// kMaxParts, bounds[0] = 0, bounds[parts] = n, non-decreasing; pe zeroed by
// the caller).
constexpr int kMaxParts = 1024;
hipError_t launch_local_parts(const int32_t* ci, const int32_t* esrc, int64_t n, int64_t e0, int64_t e1,
                              const uint32_t* ecount, const int32_t* bounds, int32_t parts, unsigned long long* pe,
                              int cus, hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, int64_t bytes, int grid, hipStream_t s);
// sink: grid * kBlock words (never written for a byte-filled source).
hipError_t launch_read(const void* src, int64_t bytes, uint32_t* sink, int grid, hipStream_t s);
// Resident workgroups per CU of the persistent and plain kernels.
int count_blocks_per_cu();
int plain_blocks_per_cu();

// Static plan scratch of the allocation-free count (tc_count_noalloc): a
// count over at most kNoAllocSlots plan indices keeps its bins, items, scan
// and counters there instead of in allocated memory.
constexpr int64_t kNoAllocSlots = (int64_t)1 << 21;
// Byte size of the scratch and the plan buffers laid out in it (base = its
// device address, from noalloc_scratch_address).
int64_t noalloc_scratch_bytes();
PlanBufs noalloc_plan_bufs(void* base);
// Device address of the scratch on the current device.
hipError_t noalloc_scratch_address(void** ptr);

}  // namespace tck
