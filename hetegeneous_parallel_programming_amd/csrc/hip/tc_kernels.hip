// Triangle-counting kernels for gfx950 (CDNA4, wave64).
//
// The counted quantity is the reference's bs_warp_vertex loop nest
// (bs_warp_vertex.cuh:13-22): sum over pivots u, over v in N+(u), of
// |N+(u) ∩ N+(v)| on the degree-ordered DAG, i.e. every triangle once.
//
// Partner scheme.  The same sum is grouped by edge owner: edge (u, v) is
// owned by the endpoint with the longer out-list (u when d+(v) <= d+(u)),
// which keeps that list as the set; the other endpoint (its partner) has its
// list streamed through it.  Each edge then streams min(d+(u), d+(v)) ids
// instead of d+(v); the intersection is symmetric, so every edge contributes
// the same term as in the reference.  A build pass (once per upload) writes
// each owner's partners as list references and cuts owners into partner
// rows of at most kPartnersPerRow partners -- the unit of work.
//
// Work unit.  A group of G lanes (G = 8, 16, 32 or 64, or a whole workgroup)
// takes one partner row: it builds the owner's set in LDS (a two-choice
// cuckoo hash, or an Eytzinger search tree when the hash does not fit or
// does not converge; sets too large for LDS are searched in HBM), then walks
// the partners in batches of G: a group scan of the batch's list lengths
// gives each partner a slice of one flat index space, and the lanes stride
// that space with coalesced int32 loads, each id looked up in the set.
// Several lookups per lane (TC_IN_FLIGHT) stay in flight so that the HBM/L2
// reads and the LDS probes overlap.
//
// Scheduling.  A planning pass bins partner rows by set size and work (the
// total length of their partners' lists) into workgroup, wave and 32/16/8-
// lane groups; one persistent kernel per bin pulls tiles (runs of rows of
// that bin) from a device queue, heaviest bins first.  Counts are u64 sums,
// so the result does not depend on tile order.  Local variants of the bin
// kernels also attribute every hit to its edge and corners.
#include <climits>

#include "tc_kernels.hpp"

namespace tck {

// Orders LDS traffic between lanes of one wave (a wave's LDS operations are
// performed in issue order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exclusive scan of x over the G lanes of a group; total = group sum.
template <int G>
__device__ __forceinline__ int group_excl_scan(int x, int gl, int& total) {
  int incl = x;
#pragma unroll
  for (int d = 1; d < G; d <<= 1) {
    const int y = __shfl_up(incl, d, G);
    if (gl >= d) incl += y;
  }
  total = __shfl(incl, G - 1, G);
  return incl - x;
}

// First k in [0, G) with sEnd[k] > t (sEnd non-decreasing, sEnd[G-1] > t).
template <int G>
__device__ __forceinline__ int find_edge(const int32_t* sEnd, int t) {
  int k = 0;
#pragma unroll
  for (int step = G / 2; step >= 1; step >>= 1)
    if (sEnd[k + step - 1] <= t) k += step;
  return k;
}

// Pivot lists in LDS are implicit complete binary search trees in BFS
// ("Eytzinger") order: node k has children 2k and 2k+1, node 1 is the root,
// word 0 holds -1 (never an id) and leaves past the list hold INT_MAX.  An
// in-order walk gives the sorted list, so a search is still a binary search,
// but one descent step is just k = 2k + (slot[k] < w) -- no bounds, no
// midpoint arithmetic -- and the upper levels are consecutive words, read
// without bank conflicts.  A tree of h levels holds up to 2^h - 1 ids in a
// slot of 2^h words.

// Levels of the smallest complete tree holding n >= 1 ids.
__device__ __forceinline__ int tree_levels(int n) { return 32 - __builtin_clz((unsigned)n); }

// BFS position of in-order index x (0-based) in a complete tree of h levels.
__device__ __forceinline__ int tree_pos(int x, int h) {
  const int i1 = x + 1;
  const int tz = __builtin_ctz((unsigned)i1);
  return (1 << (h - 1 - tz)) + (i1 >> (tz + 1));
}

#ifndef TC_IN_FLIGHT
#define TC_IN_FLIGHT 8
#endif
constexpr int kInFlight = TC_IN_FLIGHT;  // lookups per lane per iteration
// Edge ends a lane reads together when it steps past its edge, by group
// width, and how many such reads a wave-wide group makes before it falls
// back to a binary search (sub-wave groups: TC_STEP_STAGES_GROUP).
#ifndef TC_STEP_LA64
#define TC_STEP_LA64 2
#endif
#ifndef TC_STEP_LA32
#define TC_STEP_LA32 2
#endif
#ifndef TC_STEP_LA16
#define TC_STEP_LA16 2
#endif
#ifndef TC_STEP_LA8
#define TC_STEP_LA8 1
#endif
#ifndef TC_STEP_STAGES
#define TC_STEP_STAGES 2
#endif
#ifndef TC_STEP_STAGES_GROUP
#define TC_STEP_STAGES_GROUP 1
#endif
#ifndef TC_IN_FLIGHT_GROUP
#define TC_IN_FLIGHT_GROUP 6
#endif
// Sub-wave groups keep per-group state in vector registers; fewer lookups
// in flight leave room for it.
constexpr int kInFlightGroup = TC_IN_FLIGHT_GROUP;
// Steps whose every slot holds an id skip the per-slot bounds checks: bit 0
// for wave-wide groups, bit 1 for sub-wave groups (code size and registers
// grow with the extra copy of the step).
#ifndef TC_SPLIT_FULL
#define TC_SPLIT_FULL 1
#endif

// ---------------------------------------------------------------------------
// Pivot sets: how a group holds N+(u) for membership tests
// ---------------------------------------------------------------------------
// kSetHash: a two-choice cuckoo hash in LDS.  Two tables of B buckets, two
//   slots (8 bytes) per bucket; id w may sit only in bucket h0(w) of table 0
//   or bucket h1(w) of table 1, so a lookup is two 8-byte LDS reads and four
//   compares -- a fixed cost, independent of d+(u), with no data-dependent
//   branch.  h0/h1 are the top log2(B) bits of w * A0 / w * A1 modulo 2^24
//   (one full-rate v_mul_u32_u24 and one v_bfe_u32 each): multiplicative
//   (Fibonacci) hashing with A / 2^24 near 1/phi and 1/sqrt(2), so ids that
//   differ by a few land far apart -- neighbour lists often hold runs of
//   nearby ids.  Empty slots hold -1, which is never an id.
// kSetTree: the Eytzinger tree above, used when the hash does not fit the
//   group's LDS area or its construction did not converge.
// kSetHbm: the sorted list in HBM, binary-searched (pivots whose list does
//   not fit a workgroup's LDS).
// kSetHashWide: the same hash over keys w ^ (w >> 24), for graphs whose ids
//   reach 2^24 (the 24-bit multiply would otherwise put ids equal modulo 2^24
//   in the same buckets of both tables, and sets holding more than four of
//   them could not be built).
enum SetMode : int { kSetHash = 0, kSetTree = 1, kSetHbm = 2, kSetHashWide = 3 };
constexpr int kModeHash = 1 << kSetHash, kModeTree = 1 << kSetTree, kModeHbm = 1 << kSetHbm;

struct PivotSet {
  const int32_t* A;  // LDS slot area (hash tables or tree) or the HBM list
  int mode;
  int p0;            // hash: log2(B); tree: levels; hbm: list length
  int p1;            // hash: byte offset of table 1 (8B)
};

// An int32 in global memory.  Edge bases pass through LDS as plain pointer
// values; typing them as global keeps the element loads global_load (a
// generic pointer would make them flat loads, which also wait on LDS).
typedef __attribute__((address_space(1))) const int32_t gint32;

constexpr int32_t kEmpty = -1;
constexpr uint32_t kHashA0 = 0x9E3779u;  // odd 24-bit multipliers: 2^24 / phi,
constexpr uint32_t kHashA1 = 0xB504F3u;  // 2^24 / sqrt(2)

// Bucket of key w in a table of 2^lb buckets: bits [24 - lb, 24) of w * mul.
__device__ __forceinline__ uint32_t hash_bucket(int w, uint32_t mul, int lb) {
  return __builtin_amdgcn_ubfe((uint32_t)__umul24((uint32_t)w, mul), (uint32_t)(24 - lb), (uint32_t)lb);
}
// The key hashed for id w: w itself, or with its bits 24..30 folded into the
// low bits (a bijection on ids, so distinct ids stay distinct keys).
template <bool kWide>
__device__ __forceinline__ int hash_key(int w) {
  return kWide ? (w ^ (w >> 24)) : w;
}

// log2 of the buckets per table for a list of da ids in an LDS area of
// `area` words (4B words for both tables): load <= 1/2 when the area allows,
// at most 4/5; 0 when the list needs more than that.
__device__ __forceinline__ int hash_log2_buckets(int da, int area) {
  const int want = (da + 1) >> 1;  // B >= da/2
  int lb = want <= 2 ? 1 : 32 - __builtin_clz((unsigned)(want - 1));
  const int maxlb = 31 - __builtin_clz((unsigned)(area >> 2));
  lb = lb < maxlb ? lb : maxlb;
  return 5 * da <= (16 << lb) ? lb : 0;
}

// Inserts `key` into the two tables (B = 1 << lb buckets each, table 1 at
// word 2B).  LDS atomics make concurrent inserts by the lanes of a group
// safe: every id is always either in exactly one slot or carried by exactly
// one lane.  Returns false when an evicted id could not be placed within
// `kicks` moves (the set is then incomplete and must not be used).
template <bool kWide>
__device__ __forceinline__ bool hash_insert(int32_t* T, int lb, int key, int kicks) {
  const int B = 1 << lb;
  int t = 0;
  for (int it = 0; it < kicks; ++it) {
    const int hk = hash_key<kWide>(key);
    int32_t* b = T + (t ? 2 * B : 0) + 2 * (int)hash_bucket(hk, t ? kHashA1 : kHashA0, lb);
    if (atomicCAS(&b[0], kEmpty, key) == kEmpty) return true;
    if (atomicCAS(&b[1], kEmpty, key) == kEmpty) return true;
    if (it == 0) {  // the other table's bucket before evicting anyone
      int32_t* c = T + 2 * B + 2 * (int)hash_bucket(hk, kHashA1, lb);
      if (atomicCAS(&c[0], kEmpty, key) == kEmpty) return true;
      if (atomicCAS(&c[1], kEmpty, key) == kEmpty) return true;
    }
    key = atomicExch(&b[(it ^ key) & 1], key);  // evicted id moves to its other table
    if (key == kEmpty) return true;
    t ^= 1;
  }
  return false;
}

// Lays N+(u) out as an LDS tree with `n` cooperating lanes (coalesced reads).
__device__ __forceinline__ void stage_tree(int32_t* slot, const int32_t* __restrict__ src, int da,
                                           int lane, int n) {
  const int h = tree_levels(da);
  const int nodes = (1 << h) - 1;
  for (int x = lane; x < nodes; x += n) slot[tree_pos(x, h)] = x < da ? src[x] : INT_MAX;
  if (lane == 0) slot[0] = -1;
}

// True on every lane of the group when `ok` holds on all of its G lanes
// (G lanes of one wave, or the whole workgroup for G == kBlock).  Orders the
// group's preceding LDS operations before its following ones.
template <int G>
__device__ __forceinline__ bool group_all(bool ok) {
  if constexpr (G == kBlock) {
    return __syncthreads_and(ok) != 0;
  } else {
    wave_sync();
    const uint64_t bad = __ballot(!ok);
    if constexpr (G == 64) {
      return bad == 0;
    } else {
      const int first = (int)(threadIdx.x & 63) & ~(G - 1);
      return ((bad >> first) & ((1ull << G) - 1)) == 0;
    }
  }
}

template <int G>
__device__ __forceinline__ void group_sync() {
  if constexpr (G == kBlock) __syncthreads();
  else wave_sync();
}

// Attribution of a count: the kLocal parameter of the count kernels
// (kLocalNone .. kLocalParts, declared with LocalOut).  Inside a row's
// intersections the parts mode works as the plain count does.
constexpr int edge_local(int kLocal) { return kLocal == kLocalParts ? kLocalNone : kLocal; }

// Builds the set of N+(u) (src[0..da), sorted) in `slot` (area words) with
// the G lanes of a group (gl = lane index in the group): a cuckoo hash when
// it fits and converges (and the tune allows it), else the tree (da must
// then fit the tree cap).
// kLocalFull: C (area / 2 words: a 16-bit hit counter per slot word, two
// per word) gets zeroed over the slots the set uses.
// kWide: ids reach 2^24 (kSetHashWide).  fallbacks: counts the groups whose
// hash build did not converge (one increment per fallen-back set).
template <int G, int kLocal, bool kWide>
__device__ __forceinline__ PivotSet build_pivot_set(int32_t* slot, int area, const int32_t* __restrict__ src,
                                                    int da, int gl, const Tune& tune, uint32_t* fallbacks,
                                                    uint32_t* C = nullptr) {
  PivotSet S{slot, kSetTree, 0, 0};
  // tree_only: 1 = every set a tree; 2 = trees for sub-wave groups only
  const bool tree = tune.tree_only == 1 || (tune.tree_only == 2 && G < 64);
  const int lb = tree ? 0 : hash_log2_buckets(da, area);
  if (lb > 0) {
    const int B = 1 << lb;
    for (int x = gl; x < 4 * B; x += G) {
      slot[x] = kEmpty;
      if constexpr (kLocal == kLocalFull)
        if ((x & 1) == 0) C[x >> 1] = 0;
    }
    group_sync<G>();
    bool ok = true;
    for (int j = gl; j < da; j += G) ok = hash_insert<kWide>(slot, lb, src[j], tune.hash_kicks) && ok;
    if (group_all<G>(ok)) {
      S.mode = kWide ? kSetHashWide : kSetHash;
      S.p0 = lb;
      S.p1 = 8 * B;
      return S;
    }
    if (gl == 0 && fallbacks) atomicAdd(fallbacks, 1u);
  }
  stage_tree(slot, src, da, gl, G);
  S.p0 = tree_levels(da);
  if constexpr (kLocal == kLocalFull)
    for (int x = gl; x < (1 << S.p0) / 2; x += G) C[x] = 0;
  group_sync<G>();
  return S;
}

// Per-hit attribution of a local count.  C: LDS hit counters of the set's
// slots, 16 bits each (a slot is hit at most once per partner, and a row has
// at most kPartnersPerRow partners); null for an HBM set, whose hits go to
// vcount directly.  EC: LDS hit counters of the G edges of the current
// batch.  pedge / pvid: DAG edge position and partner vertex of each of the
// row's partners.  ecount / vcount: the outputs.
struct LocalSink {
  uint32_t* C;
  uint32_t* EC;
  const int64_t* pedge;
  const int32_t* pvid;
  uint32_t* ecount;
  unsigned long long* vcount;
};
static_assert(kPartnersPerRow < 65536, "16-bit slot hit counters");

__device__ __forceinline__ void count_slot_hit(uint32_t* C, int idx) {
  atomicAdd(&C[idx >> 1], 1u << ((idx & 1) << 4));
}

// Adds one lookup step's hits (f) to the counters EC of their batch edges
// (kq).  The lanes of a wide group stream one or two edges per step, so
// per-hit atomics would pile onto one LDS word; instead the wave adds once
// per distinct hit counter (ballot, one lane adds the popcount; counters
// told apart by LDS address, so the groups of a wave are kept apart).
// Call from wave-uniform control flow.  Groups narrower than
// TC_EDGE_HITS_MIN_G spread over as many edges as groups and add per hit,
// and so do full local counts (their per-hit branch already bumps a slot
// counter; a second, per-edge loop beside it costs more than it saves).
#ifndef TC_EDGE_HITS_MIN_G
#define TC_EDGE_HITS_MIN_G 64
#endif
template <int G, int kLocal>
__device__ __forceinline__ void edge_hits(uint32_t* EC, int kq, bool f) {
  if constexpr (kLocal == kLocalEdges && G >= TC_EDGE_HITS_MIN_G) {
    const uint32_t key = (uint32_t)(uintptr_t)(EC + kq);
    uint64_t rem = __ballot(f);
    while (rem) {
      const int lead = __builtin_ctzll(rem);
      const uint64_t same = rem & __ballot(key == (uint32_t)__builtin_amdgcn_readlane((int)key, lead));
      if ((int)(threadIdx.x & 63) == lead) atomicAdd(&EC[kq], (uint32_t)__builtin_popcountll(same));
      rem &= ~same;
    }
  } else {
    if (f) atomicAdd(&EC[kq], 1u);
  }
}

// Adds the set's per-slot hit counters to vcount of the ids in those slots
// (empty hash slots hold -1, tree padding INT_MAX, tree word 0 -1).
template <int G>
__device__ __forceinline__ void flush_set_hits(const int32_t* slot, const uint32_t* C, const PivotSet& S, int gl,
                                               unsigned long long* vcount) {
  const int used = (S.mode == kSetHash || S.mode == kSetHashWide) ? (4 << S.p0) : (1 << S.p0);
  for (int x = gl; x < used; x += G) {
    const uint32_t c = (C[x >> 1] >> ((x & 1) << 4)) & 0xFFFFu;
    const int key = slot[x];
    if (c != 0 && key >= 0 && key != INT_MAX) atomicAdd(&vcount[key], (unsigned long long)c);
  }
}

// Sum of x over the G lanes of a group (G <= 64), on every lane.
template <int G>
__device__ __forceinline__ uint32_t group_sum(uint32_t x) {
#pragma unroll
  for (int o = G / 2; o >= 1; o >>= 1) x += __shfl_xor(x, o, G);
  return x;
}

// NF lookups of ids w[] in a hash set: two 8-byte LDS reads (one bucket of
// each table) and four compares per id.  kLocal: a hit also bumps the hit
// counter of its batch edge kq[] and (kLocalFull) of its slot.
template <int G, int NF, int kLocal, bool kWide>
__device__ __forceinline__ void hash_probe(const PivotSet& S, const int (&w)[NF], const int (&kq)[NF],
                                           uint32_t& hits, const LocalSink& ls) {
  const char* Tb = (const char*)S.A;
  const int lb = S.p0;
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    const int hk = hash_key<kWide>(w[q]);
    const uint32_t b0 = hash_bucket(hk, kHashA0, lb), b1 = hash_bucket(hk, kHashA1, lb);
    const int2 x = *(const int2*)(Tb + (b0 << 3));
    const int2 y = *(const int2*)(Tb + S.p1 + (b1 << 3));
    const bool f = (x.x == w[q]) | (x.y == w[q]) | (y.x == w[q]) | (y.y == w[q]);
    hits += f ? 1u : 0u;
    if constexpr (kLocal == kLocalFull) {
      if (f) {
        const int t1 = S.p1 >> 2;  // word offset of table 1
        const int idx = x.x == w[q] ? 2 * (int)b0
                      : x.y == w[q] ? 2 * (int)b0 + 1
                      : y.x == w[q] ? t1 + 2 * (int)b1 : t1 + 2 * (int)b1 + 1;
        count_slot_hit(ls.C, idx);
      }
    }
    if constexpr (kLocal) edge_hits<G, kLocal>(ls.EC, kq[q], f);
  }
}

// Counts, for the edges j = jfirst, jfirst + jstride, ... of pivot u (list
// elist[0..da) in HBM, set S), the elements of N+(v) found in S.
// sEnd/sBase are this group's G-entry LDS scratch.  kModes: the set modes
// this instantiation handles.  Returns this lane's share of the hits.
//
// Flat index t of a batch maps to edge k (sEnd[k-1] <= t < sEnd[k]) and to
// element t of the int32 array at sBase[k] (N+(v) shifted back by the edge's
// first flat index); each lane keeps its current edge's end and base in
// registers and searches sEnd only when it steps past that edge.
// Where the streamed lists of a pivot's edges come from: ids of a DAG row
// (one rowptr lookup each), or partner list references.
struct DagEdges {
  const int64_t* rp;
  const int32_t* ids;
  __device__ __forceinline__ void operator()(int j, int64_t& b, int& c) const {
    const int v = ids[j];
    b = rp[v];
    c = (int)(rp[v + 1] - b);
  }
};
struct RefEdges {
  const ListRef* refs;
  __device__ __forceinline__ void operator()(int j, int64_t& b, int& c) const {
    const ListRef r = refs[j];
    b = list_start(r);
    c = list_len(r);
  }
};

// kWide: hash sets hold keys folded for ids past 2^24 (kSetHashWide); one
// instantiation per hash kind keeps a single lookup path in the loop.
template <int G, int kModes, int NF = kInFlight, int kLocal = kLocalNone, bool kWide = false, class Src>
__device__ __forceinline__ uint32_t pivot_edges(const int32_t* __restrict__ ci, const Src src, int da,
                                                const PivotSet S, int jfirst, int jstride,
                                                int32_t* sEnd, gint32** sBase, int gl,
                                                const LocalSink& ls = LocalSink{}) {
  uint32_t hits = 0;
  const int32_t* A = S.A;
  for (int j0 = jfirst; j0 < da; j0 += jstride) {
    const int j = j0 + gl;
    int c = 0;
    gint32* base = (gint32*)ci;
    if (j < da) {
      int64_t b;
      src(j, b, c);
      base = (gint32*)ci + b;
    }
    // kLocal: the batch edge's DAG position, loaded before the batch is
    // streamed so that the store after it does not wait for the load
    // (RMAT-22 edges-only local count ~0.05 ms faster on MI355X)
    int64_t ej = 0;
    if constexpr (kLocal)
      if (j < da) ej = ls.pedge[j];
    int total;
    const int excl = group_excl_scan<G>(c, gl, total);
    sEnd[gl] = excl + c;
    sBase[gl] = base - excl;
    wave_sync();
    int cur_k = -1;   // the lane's current edge
    int cur_end = 0;  // t < cur_end  <=>  t lies in the lane's current edge
    gint32* cur_base = (gint32*)ci;
    // A lane's flat index grows by G per id, so its next edge is one of the
    // next few: a step reads the next kLookAhead ends together (lists of a
    // few dozen ids make a wave-wide group's lanes pass a couple of edges
    // per step, a narrow group's one), then, for lanes still past them, up
    // to kStages - 1 more such reads, and searches only past those.  Each
    // read moves k over the run of ends at or below t.  t < total =
    // sEnd[G - 1] keeps cur_k + 1 inside the batch and ends every run inside
    // the group, so words read past the group's G entries (other groups'
    // ends, or the Base array after End: LDS of this block) never count.
    constexpr int kLookAhead = G >= 64 ? TC_STEP_LA64 : G >= 32 ? TC_STEP_LA32 : G >= 16 ? TC_STEP_LA16 : TC_STEP_LA8;
    constexpr int kStages = G >= 64 ? TC_STEP_STAGES : TC_STEP_STAGES_GROUP;
    auto step_to = [&](int t) {
      int k = cur_k, e = cur_end;  // e <= t
#pragma unroll
      for (int st = 0; st < kStages; ++st) {
        if (st > 0 && e > t) break;
        int ev[kLookAhead];
#pragma unroll
        for (int i = 0; i < kLookAhead; ++i) ev[i] = sEnd[k + 1 + i];
        k += 1;
        e = ev[0];
        bool run = true;
#pragma unroll
        for (int i = 0; i + 1 < kLookAhead; ++i) {
          run = run && ev[i] <= t;
          k += run ? 1 : 0;
          e = run ? ev[i + 1] : e;
        }
      }
      if (e <= t) {
        k = find_edge<G>(sEnd, t);
        e = sEnd[k];
      }
      cur_k = k;
      cur_end = e;
      cur_base = sBase[k];
    };
    for (int t0 = 0; t0 < total; t0 += G * NF) {
      int w[NF];
      int kq[NF];  // kLocal: batch edge of each id
      if ((TC_SPLIT_FULL & (G == 64 ? 1 : 2)) && t0 + G * NF <= total) {  // every slot holds an id
#pragma unroll
        for (int q = 0; q < NF; ++q) {
          const int t = t0 + q * G + gl;
          if (t >= cur_end) step_to(t);
          w[q] = cur_base[t];
          if constexpr (kLocal) kq[q] = cur_k;
        }
      } else {
#pragma unroll
        for (int q = 0; q < NF; ++q) {
          const int t = t0 + q * G + gl;
          w[q] = INT_MIN;  // ids are >= 0: never found
          if constexpr (kLocal) kq[q] = 0;
          if (t < total) {
            if (t >= cur_end) step_to(t);
            w[q] = cur_base[t];
            if constexpr (kLocal) kq[q] = cur_k;
          }
        }
      }
      if ((kModes & kModeHash) && S.mode == (kWide ? kSetHashWide : kSetHash)) {
        hash_probe<G, NF, kLocal, kWide>(S, w, kq, hits, ls);
      } else if ((kModes & kModeTree) && ((kModes == kModeTree) || S.mode == kSetTree)) {
        // h descent steps; then the lower bound is k with its trailing one
        // bits and the following zero shifted out (0 = past the end -> -1).
        // The node index is carried as a byte offset (4k): one compare, one
        // select and one shift-add per step.  ~k is never 0 (k < 2^h).
        const int h = S.p0;
        const char* Ab = (const char*)A;
        int kb[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q) kb[q] = 4;
        for (int s = 0; s < h; ++s) {
          int val[NF];
#pragma unroll
          for (int q = 0; q < NF; ++q) val[q] = *(const int32_t*)(Ab + kb[q]);
#pragma unroll
          for (int q = 0; q < NF; ++q) kb[q] = 2 * kb[q] + (val[q] < w[q] ? 4 : 0);
        }
#pragma unroll
        for (int q = 0; q < NF; ++q) {
          const unsigned k = (unsigned)kb[q] >> 2;
          const unsigned lo = (k >> __builtin_ctz(~k)) >> 1;
          const bool f = A[lo] == w[q];
          hits += f ? 1u : 0u;
          if constexpr (kLocal == kLocalFull)
            if (f) count_slot_hit(ls.C, (int)lo);
          if constexpr (kLocal) edge_hits<G, kLocal>(ls.EC, kq[q], f);
        }
      } else if (kModes & kModeHbm) {
        int lo[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q) lo[q] = 0;
        for (int len = S.p0; len > 1;) {
          const int half = len >> 1;
#pragma unroll
          for (int q = 0; q < NF; ++q) lo[q] = (A[lo[q] + half] <= w[q]) ? lo[q] + half : lo[q];
          len -= half;
        }
#pragma unroll
        for (int q = 0; q < NF; ++q) {
          const bool f = A[lo[q]] == w[q];
          hits += f ? 1u : 0u;
          if constexpr (kLocal) {
            if constexpr (kLocal == kLocalFull)
              if (f) atomicAdd(&ls.vcount[w[q]], 1ull);
            if constexpr (kLocal) edge_hits<G, kLocal>(ls.EC, kq[q], f);
          }
        }
      }
    }
    if constexpr (kLocal) {
      // every edge of the batch is counted in this batch only: store it,
      // and (kLocalFull) credit its partner endpoint
      wave_sync();
      if (j < da) {
        const uint32_t c = ls.EC[gl];
        ls.ecount[ej] = c;
        if constexpr (kLocal == kLocalFull)
          if (c) atomicAdd(&ls.vcount[ls.pvid[j]], (unsigned long long)c);
      }
      ls.EC[gl] = 0;
    }
    wave_sync();
  }
  return hits;
}
// Values the compiler cannot prove wave-uniform (read from LDS, or derived
// from threadIdx / 64) but that are: moved to scalar registers.
__device__ __forceinline__ int uniform32(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t uniform64(int64_t x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ uint64_t shfl_down_u64(uint64_t x, int d) {
  const uint32_t lo = __shfl_down((uint32_t)x, d, 64);
  const uint32_t hi = __shfl_down((uint32_t)(x >> 32), d, 64);
  return ((uint64_t)hi << 32) | lo;
}

// Sum over the workgroup; result valid in thread 0.  All threads must call.
__device__ __forceinline__ uint64_t block_sum(uint64_t x, uint64_t* red) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += shfl_down_u64(x, d);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = x;
  __syncthreads();
  uint64_t s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kWaves; ++i) s += red[i];
  return s;
}

// ---------------------------------------------------------------------------
// zero fill (used instead of memset nodes inside captured graphs)
// ---------------------------------------------------------------------------
__global__ void zero_kernel(uint32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// ---------------------------------------------------------------------------
// planning: bin every pivot, then scatter (graph, row) items grouped by bin
// ---------------------------------------------------------------------------
// Largest g with row_base <= q (build row slots are in graph order).
// Graph of a row slot from the radix build's gid (~graph marks a row that is
// not split).
__device__ __forceinline__ int gid_graph(int32_t x) { return x >= 0 ? x : ~x; }

__device__ __forceinline__ int find_graph_rows(const GraphDesc* descs, int ngraphs, int64_t q) {
  int lo = 0, hi = ngraphs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].row_base <= q) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// exclusive scan over n values (three passes: block sums, one-workgroup scan
// of the sums, block scans); loads and stores go through functors so one
// implementation serves uint32 plan counts and int64 partner offsets.
// ---------------------------------------------------------------------------
constexpr int kScanPer = 8;                        // values per thread
constexpr int kScanChunk = kBlock * kScanPer;      // values per block
constexpr int kScanTop = 1024;                     // threads of the top pass

int64_t scan_partials(int64_t n) { return (n + kScanChunk - 1) / kScanChunk + 1; }

// Exclusive scan of x over the workgroup; *total = workgroup sum (all threads).
__device__ __forceinline__ int64_t block_excl_scan(int64_t x, int64_t* s_wave, int64_t* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int64_t incl = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) s_wave[wid] = incl;
  __syncthreads();
  int64_t before = 0, all = 0;
  for (int w = 0; w < (int)(blockDim.x / 64); ++w) {
    if (w < wid) before += s_wave[w];
    all += s_wave[w];
  }
  __syncthreads();
  *total = all;
  return before + incl - x;
}

template <class Load>
__global__ __launch_bounds__(kBlock) void scan_reduce_kernel(Load in, int64_t n, int64_t* part) {
  __shared__ int64_t s_wave[kWaves];
  const int64_t base = (int64_t)blockIdx.x * kScanChunk;
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
    if (i < n) s += in(i);
  }
  int64_t total;
  block_excl_scan(s, s_wave, &total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// part[0..nb) -> exclusive offsets; part[nb] = grand total.
__global__ __launch_bounds__(kScanTop) void scan_top_kernel(int64_t* part, int64_t nb) {
  __shared__ int64_t s_wave[kScanTop / 64];
  const int64_t per = (nb + kScanTop - 1) / kScanTop;
  const int64_t b0 = min((int64_t)threadIdx.x * per, nb), b1 = min(b0 + per, nb);
  int64_t s = 0;
  for (int64_t b = b0; b < b1; ++b) s += part[b];
  int64_t total;
  int64_t run = block_excl_scan(s, s_wave, &total);
  for (int64_t b = b0; b < b1; ++b) {
    const int64_t c = part[b];
    part[b] = run;
    run += c;
  }
  if (threadIdx.x == 0) part[nb] = total;
}

// Each thread scans kScanPer consecutive values of its block's chunk.
template <class Load, class Store>
__global__ __launch_bounds__(kBlock) void scan_down_kernel(Load in, Store out, int64_t n,
                                                           const int64_t* part) {
  __shared__ int64_t s_wave[kWaves];
  const int64_t i0 = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
  int64_t v[kScanPer];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = (i0 + k < n) ? in(i0 + k) : 0;
    s += v[k];
  }
  int64_t total;
  int64_t run = part[blockIdx.x] + block_excl_scan(s, s_wave, &total);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (i0 + k < n) out(i0 + k, run);
    run += v[k];
  }
}

template <class Load, class Store>
static hipError_t launch_scan(Load in, Store out, int64_t n, int64_t* part, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t nb = (n + kScanChunk - 1) / kScanChunk;
  scan_reduce_kernel<Load><<<(unsigned)nb, kBlock, 0, s>>>(in, n, part);
  scan_top_kernel<<<1, kScanTop, 0, s>>>(part, nb);
  scan_down_kernel<Load, Store><<<(unsigned)nb, kBlock, 0, s>>>(in, out, n, part);
  return hipGetLastError();
}

struct LoadU32 {
  const uint32_t* p;
  __device__ int64_t operator()(int64_t i) const { return p[i]; }
};
struct StoreU32 {
  uint32_t* p;
  __device__ void operator()(int64_t i, int64_t v) const { p[i] = (uint32_t)v; }
};
// c1[i] plus row i's kXcds in-edge counters (stripe-major, stride n).
struct LoadSum2 {
  const int32_t* a;
  const int32_t* b;
  int64_t n;
  __device__ int64_t operator()(int64_t i) const {
    int64_t t = a[i];
#pragma unroll
    for (int k = 0; k < kXcds; ++k) t += b[k * n + i];
    return t;
  }
};
struct LoadI32 {
  const int32_t* p;
  __device__ int64_t operator()(int64_t i) const { return p[i]; }
};
struct LoadI64 {
  const int64_t* p;
  __device__ int64_t operator()(int64_t i) const { return p[i]; }
};
struct LoadU64 {
  const unsigned long long* p;
  __device__ int64_t operator()(int64_t i) const { return (int64_t)p[i]; }
};
struct StoreU64 {
  unsigned long long* p;
  __device__ void operator()(int64_t i, int64_t v) const { p[i] = (unsigned long long)v; }
};
// c1[i] + c2[i]: a row's partners through out-edges and in-edges.
struct LoadAdd {
  const int32_t* a;
  const int32_t* b;
  __device__ int64_t operator()(int64_t i) const { return (int64_t)a[i] + b[i]; }
};
struct StoreI64 {
  int64_t* p;
  __device__ void operator()(int64_t i, int64_t v) const { p[i] = v; }
};

// ---------------------------------------------------------------------------
// partner build (one thread per DAG row slot of the rowptr pool)
// ---------------------------------------------------------------------------
// Passes 1 and 2 give each DAG row a group of kBuildG lanes that stride its
// edges (rows average a few tens of ids; a thread per row would serialise
// the long ones).
constexpr int kBuildG = 16;

// Pass 1: per row u (owned), partners through out-edges (c1) and, through
// atomics, partners through in-edges of each edge's head (c2).
__global__ __launch_bounds__(kBlock) void partner_count_kernel(LaunchArgs a, int64_t RP, BuildBufs b) {
  const int64_t q = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kBuildG;
  const int gl = threadIdx.x % kBuildG;
  if (q >= RP) return;  // the whole group
  const int g = find_graph_rows(a.descs, a.ngraphs, q);
  const GraphDesc d = a.descs[g];
  const int64_t u = q - d.row_base;
  int32_t c1 = 0;
  if (u >= d.own_begin && u < d.own_end && (!a.own || a.own[u])) {
    const int64_t* rp = a.rp + d.rp_off;
    const int32_t* ci = a.ci + d.ci_off;
    const int64_t e0 = rp[u], e1 = rp[u + 1], du = e1 - e0;
    int32_t* c2 = b.c2 + (int64_t)(blockIdx.x % kXcds) * RP;
    for (int64_t e = e0 + gl; e < e1; e += kBuildG) {
      const int v = ci[e];
      if (rp[v + 1] - rp[v] <= du) ++c1;
      else atomicAdd(&c2[d.row_base + v], 1);
    }
  }
#pragma unroll
  for (int o = kBuildG / 2; o >= 1; o >>= 1) c1 += __shfl_xor(c1, o, kBuildG);
  if (gl == 0) b.c1[q] = c1;
}

// Pass 2: partner ids.  Out-edge partners of u fill the front of u's range
// in row order (ballot ranks); in-edge partners take the rest, stripe by
// stripe, in arrival order within a stripe (a cursor per stripe).  Order
// inside a range does not change any count.
__global__ __launch_bounds__(kBlock) void partner_fill_kernel(LaunchArgs a, int64_t RP, BuildBufs b) {
  const int64_t q = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kBuildG;
  const int gl = threadIdx.x % kBuildG;
  if (q >= RP) return;
  const int g = find_graph_rows(a.descs, a.ngraphs, q);
  const GraphDesc d = a.descs[g];
  const int64_t u = q - d.row_base;
  if (u < d.own_begin || u >= d.own_end || (a.own && !a.own[u])) return;
  const int64_t* rp = a.rp + d.rp_off;
  const int32_t* ci = a.ci + d.ci_off;
  ListRef* pci = (ListRef*)a.pci + d.pci_off;
  const int64_t* off = b.off + d.row_base;
  const int64_t base = off[0];
  int64_t w = off[u] - base;
  const int64_t e0 = rp[u], e1 = rp[u + 1], du = e1 - e0;
  const int shift = (int)(threadIdx.x & 63) & ~(kBuildG - 1);
  const uint32_t below = (1u << gl) - 1u;
  const int64_t sbase = (int64_t)(blockIdx.x % kXcds) * RP;
  for (int64_t eb = e0; eb < e1; eb += kBuildG) {
    const int64_t e = eb + gl;
    int v = 0;
    int64_t vb = 0, dv = 0;
    bool own = false;
    if (e < e1) {
      v = ci[e];
      vb = rp[v];
      dv = rp[v + 1] - vb;
      own = dv <= du;
    }
    const uint32_t gm = (uint32_t)(__ballot(own) >> shift) & ((1u << kBuildG) - 1u);
    if (e < e1) {
      int64_t slot;
      if (own) {
        slot = w + __popc(gm & below);
        pci[slot] = list_ref(d.ci_off + vb, dv);
      } else {
        slot = (int64_t)atomicAdd(&b.cur[sbase + d.row_base + v], 1ull);
        pci[slot] = list_ref(d.ci_off + e0, du);
      }
      if (b.pedge) {  // both ways it is edge (u, v)
        b.pedge[d.pci_off + slot] = d.ci_off + e;
        b.pvid[d.pci_off + slot] = own ? v : (int32_t)u;
        if (b.esrc) b.esrc[d.ci_off + e] = (int32_t)((uint32_t)u | (own ? 0u : 0x80000000u));
      }
    }
    w += __popc(gm);
  }
}

// Pass 1b (after the scan): each stripe's fill cursor starts where its
// in-edge partners go: after the row's out-edge partners and the earlier
// stripes' in-edge partners (offsets relative to the graph's first partner).
__global__ __launch_bounds__(kBlock) void partner_cursor_kernel(LaunchArgs a, int64_t RP, BuildBufs b) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= RP) return;
  const GraphDesc d = a.descs[find_graph_rows(a.descs, a.ngraphs, q)];
  int64_t at = b.off[q] - b.off[d.row_base] + b.c1[q];
#pragma unroll
  for (int k = 0; k < kXcds; ++k) {
    b.cur[k * RP + q] = (unsigned long long)at;
    at += b.c2[k * RP + q];
  }
}

// Pass 3: partner rows per DAG row.
__global__ __launch_bounds__(kBlock) void partner_rows_count_kernel(LaunchArgs a, int64_t RP, BuildBufs b) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= RP) return;
  const int g = b.gid ? gid_graph(b.gid[q]) : find_graph_rows(a.descs, a.ngraphs, q);
  const GraphDesc d = a.descs[g];
  const int64_t u = q - d.row_base;
  int32_t k = 0;
  if (u < d.n_rows) {
    const int64_t c = b.off[q + 1] - b.off[q];
    k = (int32_t)((c + kPartnersPerRow - 1) / kPartnersPerRow);
  }
  b.kc[q] = k;
}

// Pass 4: partner rows (owner's list, partner range), split evenly by
// partner count; the graph's row-count slot also writes nv[g] and the end
// of its last row.  Row offsets are absolute in the partner pool.
__global__ __launch_bounds__(kBlock) void partner_rows_fill_kernel(LaunchArgs a, int64_t RP, BuildBufs b) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= RP) return;
  const int g = b.gid ? gid_graph(b.gid[q]) : find_graph_rows(a.descs, a.ngraphs, q);
  const GraphDesc d = a.descs[g];
  const int64_t u = q - d.row_base;
  const int64_t pbase = b.off[d.row_base], vbase = b.voff[d.row_base];
  int64_t* vrp = (int64_t*)a.vrp + d.pr_base;
  VRow* vrow = (VRow*)a.vrow + d.pr_base;
  if (u == d.n_rows) {
    const int64_t nv = b.voff[q] - vbase;
    ((int32_t*)a.nv)[g] = (int32_t)nv;
    vrp[nv] = d.pci_off + b.off[q] - pbase;
    return;
  }
  const int32_t K = b.kc[q];
  if (K == 0) return;
  const int64_t v0 = b.voff[q] - vbase, p0 = d.pci_off + b.off[q] - pbase, c = b.off[q + 1] - b.off[q];
  const int64_t* rp = a.rp + d.rp_off;
  const ListRef set = list_ref(d.ci_off + rp[u], rp[u + 1] - rp[u]);
  // an owner with one partner row and its work known from the build: set it
  // here; the work pass fills rows whose work is still 0
  const uint32_t w = K == 1 && b.wq ? (uint32_t)min(b.wq[q], 0xFFFFFFFFull) : 0u;
  for (int32_t k = 0; k < K; ++k) {
    vrow[v0 + k] = VRow{set, w, g + a.gbase};
    vrp[v0 + k] = p0 + (c * k) / K;
    if (b.pown) b.pown[d.pr_base + v0 + k] = (int32_t)u;
  }
}

// Pass 5: work of every partner row still at 0 (the sum of its partners'
// list lengths), with a group of kBuildG lanes per slot.
__global__ __launch_bounds__(kBlock) void partner_work_kernel(LaunchArgs a) {
  const int64_t p = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kBuildG;
  const int gl = threadIdx.x % kBuildG;
  if (p >= a.P) return;  // the whole group
  VRow* vr = (VRow*)a.vrow + p;
  if (vr->set == 0 || vr->work != 0) return;  // unused slot, or work known from the build
  const int64_t pb = a.vrp[p], pe = a.vrp[p + 1];
  int64_t w = 0;
  constexpr int kU = 4;  // list references in flight per lane
  for (int64_t e0 = pb + gl; e0 < pe; e0 += kU * kBuildG) {
    ListRef r[kU];
#pragma unroll
    for (int j = 0; j < kU; ++j) r[j] = e0 + j * kBuildG < pe ? a.pci[e0 + j * kBuildG] : (ListRef)0;
#pragma unroll
    for (int j = 0; j < kU; ++j) w += list_len(r[j]);
  }
#pragma unroll
  for (int o = kBuildG / 2; o >= 1; o >>= 1) w += __shfl_xor(w, o, kBuildG);
  if (gl == 0) vr->work = (uint32_t)min(w, (int64_t)0xFFFFFFFF);
}

// ---------------------------------------------------------------------------
// partner build by a radix transpose (no per-edge global atomics)
// ---------------------------------------------------------------------------
// Every DAG edge of a split row is one pair (owner row slot, partner row
// slot): the endpoint with the longer out-list owns it, and the other
// endpoint's list joins the owner's partner range.  The partner pool is the
// pairs grouped by owner, so the build is a transpose of the pairs:
//   prep    -- per row slot: its graph (negated when the row is not split),
//              first build-space id, own list and saturated degree;
//   mpath   -- merge-path split points of the sequence (row ends + ids) in
//              blocks of kMP items;
//   split   -- per block of kMP items, whatever the row lengths: each id's
//              pair (an out-list degree gather decides the owner), and the
//              block's pair count per coarse bucket (a local rebuild also
//              marks each edge's row word in esrc where the head owns it);
//   level 1 -- each split block's pairs sorted by coarse bucket in LDS and
//              written as one run per bucket, at positions from the scanned
//              (bucket-major) counts;
//   level 2 -- the same within each coarse bucket, by fine bucket, over
//              tiles of kTile pairs;
//   final   -- a fine bucket of at most `cap` pairs is sorted by owner in
//              LDS and written out whole: its owners' partner ranges are one
//              run of the pool.  A bigger one (hub owners) is counted per
//              owner by chunks, scanned, and then each chunk is sorted in
//              LDS and every owner's run reserved on the owner's cursor.
// Fine bucket = owner >> fs (at most kCoarse * kFine of them), coarse
// bucket = fine bucket >> 6.  Within an owner the partner order is whatever
// the LDS ranks give; no count depends on it.
//
// Positions: pairs are numbered in owner order over all row slots of the
// build; the pool position of pair number p of graph g is pci_off(g) -
// eb(g) + p.  That takes graph g's pairs to start at eb(g): true for a
// single graph (eb = 0) and for launches of several graphs, where every
// row is split (batches and the server), so graph g's pairs are the ids of
// the graphs before it.

constexpr int32_t kDegSat = 65535;  // deg16 saturation

// Exclusive scan of n (<= kBlock * per) values of s (in place); returns the
// sum.  All threads of the block call it.
template <int kPer>
__device__ uint32_t block_scan_lds(uint32_t* s, int n, int64_t* s_wave) {
  const int i0 = threadIdx.x * kPer;
  uint32_t v[kPer];
  uint32_t t = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    v[k] = i0 + k < n ? s[i0 + k] : 0u;
    t += v[k];
  }
  int64_t total;
  uint32_t run = (uint32_t)block_excl_scan((int64_t)t, s_wave, &total);
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    if (i0 + k < n) s[i0 + k] = run;
    run += v[k];
  }
  __syncthreads();
  return (uint32_t)total;
}

// LDS cnt[d] += 1 for every lane of the wave with d >= 0 (all 64 lanes
// call it); returns each such lane's old value (its rank).  Owner-keyed
// digits are skewed (hub owners), and lanes adding to one LDS word
// serialise: the wave's kPeel most common digits (met in lane order) take
// one atomic per digit for all their lanes, the rest one atomic per lane.
template <int kPeel>
__device__ __forceinline__ uint32_t wave_count_add(uint32_t* cnt, int d) {
  const int lane = threadIdx.x & 63;
  const uint64_t below = (1ull << lane) - 1ull;
  bool done = d < 0;
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < kPeel; ++k) {
    const uint64_t act = __ballot(!done);
    if (act == 0) break;
    const int leader = __ffsll((unsigned long long)act) - 1;
    const int dl = __shfl(d, leader);
    const bool mine = !done && d == dl;
    const uint64_t peers = __ballot(mine);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&cnt[dl], (uint32_t)__popcll(peers));
    base = __shfl(base, leader);
    if (mine) {
      r = base + (uint32_t)__popcll(peers & below);
      done = true;
    }
  }
  if (!done) r = atomicAdd(&cnt[d], 1u);
  return r;
}

__global__ __launch_bounds__(kBlock) void radix_prep_kernel(LaunchArgs a, int64_t RP, int64_t M, BuildBufs b) {
  // graphs of this block's first and last slot (uniform), then each lane's
  const int64_t qa = (int64_t)blockIdx.x * kBlock, qz = min(qa + kBlock, RP) - 1;
  int lo = find_graph_rows(a.descs, a.ngraphs, qa), hi = find_graph_rows(a.descs, a.ngraphs, qz);
  const int64_t q = qa + threadIdx.x;
  if (q >= RP) return;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.descs[mid].row_base <= q) lo = mid;
    else hi = mid - 1;
  }
  const GraphDesc d = a.descs[lo];
  const int64_t u = q - d.row_base;
  const int64_t* rp = a.rp + d.rp_off;
  const int64_t e0 = rp[u];  // u <= n_rows
  int64_t du = 0;
  bool split = false;
  if (u < d.n_rows) {
    du = rp[u + 1] - e0;
    split = u >= d.own_begin && u < d.own_end && (!a.own || a.own[u]);
  }
  b.gid[q] = split ? lo : ~lo;
  b.erow[q] = d.eb + e0;
  b.lref[q] = u < d.n_rows ? list_ref(d.ci_off + e0, du) : (ListRef)0;
  b.deg16[q] = (uint16_t)min(du, (int64_t)kDegSat);
  if (q == RP - 1) b.erow[RP] = M;
}

// Rows consumed before merge-path item t * kMP: the largest i with
// erow[i] + i <= t * kMP (erow[i] + i rises strictly with i).
__global__ __launch_bounds__(kBlock) void radix_mpath_kernel(const int64_t* erow, int64_t RP, int64_t M,
                                                             int64_t nmp, int64_t* bq) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t > nmp) return;
  const int64_t k = min(t * kMP, RP + M);
  int64_t lo = max((int64_t)0, k - M), hi = min(k, RP);
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (erow[mid] + mid <= k) lo = mid;
    else hi = mid - 1;
  }
  bq[t] = lo;
}

// A pair's partner before placement: its list, or (kRefSlot set) the
// partner's row slot, whose list the placement reads.  Pool offsets stay
// below 2^38, so bit 63 of a list reference is free.
constexpr ListRef kRefSlot = 1ull << 63;

// Split block blk: the ids of its merge-path range, kMPPer per thread at a
// stride of kBlock (coalesced loads and stores); each id's row by a
// branch-free binary search over the block's rows in LDS.  An edge its head
// owns carries the row's own list; one its row owns carries the head's slot.
__global__ __launch_bounds__(kBlock) void radix_split_kernel(LaunchArgs a, int64_t RP, int64_t M, BuildBufs b,
                                                             int cs, int nc, int64_t nmp) {
  constexpr int kFill = (kMP + 2 + kBlock - 1) / kBlock;
  __shared__ int32_t s_erow[kMP + 2];  // first id of each row, relative to the block's first row
  __shared__ int16_t s_gid[kMP + 1];   // graph - g0 (a block spans at most kMP + 1 graphs); ~ when not split
  __shared__ uint32_t s_hist[kCoarse];
  const int64_t blk = blockIdx.x;
  const int64_t k0 = blk * kMP, k1 = min(k0 + kMP, RP + M);
  const int64_t i0 = b.bq[blk], i1 = b.bq[blk + 1];
  const int64_t j0 = k0 - i0, j1 = k1 - i1;
  // rows i0 .. min(i1, RP - 1) (k0 < RP + M, so i0 < RP) and the end of the last
  const int nr = (int)(min(i1, RP - 1) - i0 + 1);
  const int64_t ebase = b.erow[i0];
  const int g0 = gid_graph(b.gid[i0]);
  const int64_t rb0 = a.descs[g0].row_base, cd0 = a.descs[g0].ci_off - a.descs[g0].eb;
  {
    int64_t ev[kFill];
    int32_t gv[kFill];
#pragma unroll
    for (int k = 0; k < kFill; ++k) {
      const int x = k * kBlock + threadIdx.x;
      ev[k] = x <= nr ? b.erow[i0 + x] : 0;
      gv[k] = x < nr ? b.gid[i0 + x] : 0;
    }
#pragma unroll
    for (int k = 0; k < kFill; ++k) {
      const int x = k * kBlock + threadIdx.x;
      if (x <= nr) s_erow[x] = (int32_t)(ev[k] - ebase);
      if (x < nr) s_gid[x] = (int16_t)(gv[k] >= 0 ? gv[k] - g0 : ~(gid_graph(gv[k]) - g0));
    }
  }
  for (int x = threadIdx.x; x < nc; x += kBlock) s_hist[x] = 0u;
  __syncthreads();
  int32_t R[kMPPer], qv[kMPPer], dv[kMPPer];
  int64_t cix[kMPPer];  // colidx position of each id
#pragma unroll
  for (int s = 0; s < kMPPer; ++s) {
    const int64_t E = j0 + (int64_t)s * kBlock + threadIdx.x;
    R[s] = -1;
    if (E < j1) {
      const int32_t e = (int32_t)(E - ebase);
      int r = 0;  // the last row whose first id is <= e (s_erow[0] = 0 <= e)
#pragma unroll
      for (int bit = 1 << 12; bit > 0; bit >>= 1)
        if (r + bit < nr && s_erow[r + bit] <= e) r += bit;
      R[s] = r;
    }
  }
  int gc = 0;
  int64_t rb = rb0, cd = cd0;
#pragma unroll
  for (int s = 0; s < kMPPer; ++s) {
    qv[s] = -1;
    cix[s] = 0;
    if (R[s] < 0) continue;
    const int64_t E = j0 + (int64_t)s * kBlock + threadIdx.x;
    const int gr = s_gid[R[s]];
    if (gr < 0) {
      b.key[E] = -1;  // the row is not split: no pair
      continue;
    }
    if (gr != gc) {
      gc = gr;
      rb = a.descs[g0 + gr].row_base;
      cd = a.descs[g0 + gr].ci_off - a.descs[g0 + gr].eb;
    }
    cix[s] = E + cd;
    qv[s] = (int32_t)(rb + a.ci[cix[s]]);
  }
#pragma unroll
  for (int s = 0; s < kMPPer; ++s) dv[s] = qv[s] >= 0 ? (int32_t)b.deg16[qv[s]] : 0;
#pragma unroll
  for (int s = 0; s < kMPPer; ++s) {
    int digit = -1;
    if (qv[s] >= 0) {
      const int64_t E = j0 + (int64_t)s * kBlock + threadIdx.x;
      const int r = R[s];
      const int64_t du = s_erow[r + 1] - s_erow[r];
      int64_t dvx = dv[s];
      if (dvx == kDegSat && du >= kDegSat) dvx = b.erow[qv[s] + 1] - b.erow[qv[s]];  // both saturated: exact
      const int32_t q = (int32_t)(i0 + r);
      int32_t owner;
      if (dvx <= du) {  // the row owns it: partner = the head, by slot
        owner = q;
        b.ref[E] = kRefSlot | (ListRef)(uint32_t)qv[s];
      } else {  // the head owns it: partner = this row's list
        owner = qv[s];
        b.ref[E] = list_ref(cix[s] - (E - ebase - s_erow[r]), du);
      }
      b.key[E] = owner;
      // a local rebuild (one graph): the edge's row word, bit 31 when the head owns it
      if (b.esrc) b.esrc[cix[s]] = (int32_t)((uint32_t)(q - rb0) | (owner == q ? 0u : 0x80000000u));
      digit = owner >> cs;
    }
    wave_count_add<2>(s_hist, digit);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < nc; c += kBlock) b.cnt1[(int64_t)c * nmp + blk] = s_hist[c];
}

// A tile's pairs sorted by digit in LDS and written as one run per digit:
// pairs [p0, p1) of (ikey, iref), digit = (key >> shift) & mask (nd digits),
// gpos[d] = the tile's first output position of digit d.  LDS holds each
// sorted key and its place in the tile; the list references are read in
// sorted order from the tile's own range.  kEdge: each pair's id position
// rides along (iedge, or the pair's own index when iedge is null) to oedge.
template <bool kEdge>
__device__ __forceinline__ void tile_sort_write(const int32_t* __restrict__ ikey, const ListRef* __restrict__ iref,
                                                int64_t p0, int64_t p1, int shift, int mask, int nd,
                                                const uint32_t* s_gpos, uint32_t* s_cnt, int32_t* s_k,
                                                uint16_t* s_i, int64_t* s_wave, int32_t* okey, ListRef* oref,
                                                const uint32_t* __restrict__ iedge = nullptr,
                                                uint32_t* oedge = nullptr) {
  int32_t kk[kTilePer];
  uint32_t rk[kTilePer];
#pragma unroll
  for (int s = 0; s < kTilePer; ++s) {
    const int64_t p = p0 + (int64_t)s * kBlock + threadIdx.x;
    kk[s] = p < p1 ? ikey[p] : -1;
  }
#pragma unroll
  for (int s = 0; s < kTilePer; ++s) rk[s] = wave_count_add<2>(s_cnt, kk[s] >= 0 ? (kk[s] >> shift) & mask : -1);
  __syncthreads();
  // s_cnt -> starts
  __shared__ uint32_t s_st[kCoarse + 1];
  for (int x = threadIdx.x; x < nd; x += kBlock) s_st[x] = s_cnt[x];
  __syncthreads();
  const uint32_t n = block_scan_lds<1>(s_st, nd, s_wave);
#pragma unroll
  for (int s = 0; s < kTilePer; ++s) {
    if (kk[s] < 0) continue;
    const uint32_t at = s_st[(kk[s] >> shift) & mask] + rk[s];
    s_k[at] = kk[s];
    s_i[at] = (uint16_t)(s * kBlock + threadIdx.x);
  }
  __syncthreads();
  ListRef r[kTilePer];
  uint32_t ev[kTilePer];
#pragma unroll
  for (int s = 0; s < kTilePer; ++s) {
    const uint32_t p = s * kBlock + threadIdx.x;
    r[s] = p < n ? iref[p0 + s_i[p]] : (ListRef)0;
    if constexpr (kEdge) ev[s] = p < n ? (iedge ? iedge[p0 + s_i[p]] : (uint32_t)(p0 + s_i[p])) : 0u;
  }
#pragma unroll
  for (int s = 0; s < kTilePer; ++s) {
    const uint32_t p = s * kBlock + threadIdx.x;
    if (p >= n) continue;
    const int d = (s_k[p] >> shift) & mask;
    const int64_t at = (int64_t)s_gpos[d] + (p - s_st[d]);
    okey[at] = s_k[p];
    oref[at] = r[s];
    if constexpr (kEdge) oedge[at] = ev[s];
  }
}

// Level 1: split block blk's pairs -> coarse runs of pkey/pref (kEdge: and
// their id positions, the pairs' own indexes, to e1).
template <bool kEdge>
__global__ __launch_bounds__(kBlock) void radix_scatter1_kernel(BuildBufs b, int64_t RP, int64_t M, int cs, int nc,
                                                                int64_t nmp) {
  __shared__ int32_t s_k[kTile];
  __shared__ uint16_t s_i[kTile];
  __shared__ uint32_t s_cnt[kCoarse], s_gpos[kCoarse];
  __shared__ int64_t s_wave[kWaves];
  const int64_t blk = blockIdx.x;
  const int64_t k0 = blk * kMP, k1 = min(k0 + kMP, RP + M);
  const int64_t j0 = k0 - b.bq[blk], j1 = k1 - b.bq[blk + 1];
  for (int c = threadIdx.x; c < nc; c += kBlock) {
    s_cnt[c] = 0u;
    s_gpos[c] = b.cnt1[(int64_t)c * nmp + blk];
  }
  __syncthreads();
  tile_sort_write<kEdge>(b.key, b.ref, j0, j1, cs, 0x7FFFFFFF, nc, s_gpos, s_cnt, s_k, s_i, s_wave, b.pkey, b.pref,
                        nullptr, b.e1);
}

// Coarse bucket c's pairs are [cnt1[c * nmp], cnt1[(c + 1) * nmp]) of the
// level-1 output (cnt1 scanned, cnt1[nc * nmp] = all pairs).
__device__ __forceinline__ int64_t coarse_start(const BuildBufs& b, int c, int64_t nmp) {
  return (int64_t)b.cnt1[(int64_t)c * nmp];
}

// One workgroup: level-2 tiles per coarse bucket, scanned into tstart[0..nc].
__global__ __launch_bounds__(kBlock) void radix_tiles2_kernel(BuildBufs b, int nc, int64_t nmp) {
  __shared__ int64_t s_wave[kWaves];
  const int c = threadIdx.x;  // nc <= kCoarse <= kBlock
  int64_t t = 0;
  if (c < nc) t = (coarse_start(b, c + 1, nmp) - coarse_start(b, c, nmp) + kTile - 1) / kTile;
  int64_t total;
  const int64_t at = block_excl_scan(t, s_wave, &total);
  if (c < nc) b.tstart[c] = at;
  if (c == 0) b.tstart[nc] = total;
}

// Level-2 tile t: its coarse bucket, tile index within it and pair range.
struct Tile2 {
  int c;
  int64_t lt, nt, p0, p1;
};
__device__ __forceinline__ bool tile2_of(const BuildBufs& b, int64_t t, int nc, int64_t nmp, Tile2& x) {
  if (t >= b.tstart[nc]) return false;
  int lo = 0, hi = nc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.tstart[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  x.c = lo;
  x.nt = b.tstart[lo + 1] - b.tstart[lo];
  x.lt = t - b.tstart[lo];
  x.p0 = coarse_start(b, lo, nmp) + x.lt * kTile;
  x.p1 = min(x.p0 + kTile, coarse_start(b, lo + 1, nmp));
  return true;
}
// cnt2 index of (tile, fine digit d): bucket-major within the coarse bucket.
__device__ __forceinline__ int64_t cnt2_at(const BuildBufs& b, const Tile2& x, int d) {
  return b.tstart[x.c] * kFine + (int64_t)d * x.nt + x.lt;
}

__global__ __launch_bounds__(kBlock) void radix_hist2_kernel(BuildBufs b, int fs, int nc, int64_t nmp) {
  __shared__ uint32_t s_h[kFine];
  Tile2 x;
  if (!tile2_of(b, blockIdx.x, nc, nmp, x)) return;
  if (threadIdx.x < kFine) s_h[threadIdx.x] = 0u;
  __syncthreads();
  int dg[kTilePer];
#pragma unroll
  for (int s = 0; s < kTilePer; ++s) {
    const int64_t p = x.p0 + (int64_t)s * kBlock + threadIdx.x;
    dg[s] = p < x.p1 ? (b.pkey[p] >> fs) & (kFine - 1) : -1;
  }
#pragma unroll
  for (int s = 0; s < kTilePer; ++s) wave_count_add<2>(s_h, dg[s]);
  __syncthreads();
  if (threadIdx.x < kFine) b.cnt2[cnt2_at(b, x, threadIdx.x)] = s_h[threadIdx.x];
}

// Level 2: pkey/pref runs of one coarse bucket -> fine runs of key/ref.
template <bool kEdge>
__global__ __launch_bounds__(kBlock) void radix_scatter2_kernel(BuildBufs b, int fs, int nc, int64_t nmp) {
  __shared__ int32_t s_k[kTile];
  __shared__ uint16_t s_i[kTile];
  __shared__ uint32_t s_cnt[kFine], s_gpos[kFine];
  __shared__ int64_t s_wave[kWaves];
  Tile2 x;
  if (!tile2_of(b, blockIdx.x, nc, nmp, x)) return;
  if (threadIdx.x < kFine) {
    s_cnt[threadIdx.x] = 0u;
    s_gpos[threadIdx.x] = b.cnt2[cnt2_at(b, x, threadIdx.x)];
  }
  __syncthreads();
  tile_sort_write<kEdge>(b.pkey, b.pref, x.p0, x.p1, fs, kFine - 1, kFine, s_gpos, s_cnt, s_k, s_i, s_wave, b.key,
                        b.ref, b.e1, b.e2);
}

// Fine bucket f's pairs are [fine_start(f), fine_start(f + 1)) of the
// level-2 output; f >= 64 * nc is past the last pair.  ts = tstart.
__device__ __forceinline__ int64_t fine_start(const BuildBufs& b, const int64_t* ts, int64_t f, int nc, int64_t nmp) {
  const int c = (int)(f / kFine), d = (int)(f % kFine);
  if (c >= nc) return coarse_start(b, nc, nmp);
  const int64_t t0 = ts[c], nt = ts[c + 1] - t0;
  return nt > 0 ? (int64_t)b.cnt2[t0 * kFine + (int64_t)d * nt] : coarse_start(b, c, nmp);
}

// One workgroup: fst[f] = first pair of fine bucket f (f <= nf), and the
// chunks of every fine bucket above cap pairs, scanned into chstart[0..nf].
__global__ __launch_bounds__(kBlock) void radix_chunks_kernel(BuildBufs b, int nf, int nc, int64_t nmp, int cap,
                                                              int chunk) {
  constexpr int kPer = kMaxFineBuckets / kBlock;
  __shared__ int64_t s_wave[kWaves];
  __shared__ int64_t s_ts[kCoarse + 1];
  for (int c = threadIdx.x; c <= nc; c += kBlock) s_ts[c] = b.tstart[c];
  __syncthreads();
  const int f0 = threadIdx.x * kPer;
  int64_t st[kPer + 1];
#pragma unroll
  for (int k = 0; k <= kPer; ++k) st[k] = f0 + k <= nf ? fine_start(b, s_ts, f0 + k, nc, nmp) : 0;
  int64_t v[kPer];
  int64_t t = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    v[k] = 0;
    if (f0 + k < nf) {
      const int64_t n = st[k + 1] - st[k];
      v[k] = n > cap ? (n + chunk - 1) / chunk : 0;
    }
    t += v[k];
  }
  int64_t total;
  int64_t run = block_excl_scan(t, s_wave, &total);
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    if (f0 + k < nf) {
      b.chstart[f0 + k] = run;
      b.fst[f0 + k] = st[k];
    }
    run += v[k];
  }
  if (threadIdx.x == 0) {
    b.chstart[nf] = total;
    b.fst[nf] = fine_start(b, s_ts, nf, nc, nmp);
  }
}

// Chunk x of the big fine buckets: its bucket and pair range.
struct Chunk {
  int f;
  int64_t p0, p1;
};
__device__ __forceinline__ bool chunk_of(const BuildBufs& b, int64_t x, int nf, int chunk, Chunk& c) {
  if (x >= b.chstart[nf]) return false;
  int lo = 0, hi = nf - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.chstart[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  c.f = lo;
  c.p0 = b.fst[lo] + (x - b.chstart[lo]) * chunk;
  c.p1 = min(c.p0 + chunk, b.fst[lo + 1]);
  return true;
}

// Big buckets, pass 1: pairs per owner (one global add per owner and chunk).
__global__ __launch_bounds__(kBlock) void radix_count_big_kernel(BuildBufs b, int fs, int nf, int chunk) {
  extern __shared__ uint32_t s_h[];  // [1 << fs]
  Chunk c;
  if (!chunk_of(b, blockIdx.x, nf, chunk, c)) return;
  const int no = 1 << fs;
  const int32_t q0 = c.f << fs;
  for (int o = threadIdx.x; o < no; o += kBlock) s_h[o] = 0u;
  __syncthreads();
  int32_t k[kChunkPer];
#pragma unroll
  for (int s = 0; s < kChunkPer; ++s) {
    const int64_t p = c.p0 + (int64_t)s * kBlock + threadIdx.x;
    k[s] = p < c.p1 ? b.key[p] - q0 : -1;
  }
#pragma unroll
  for (int s = 0; s < kChunkPer; ++s) wave_count_add<2>(s_h, k[s]);
  __syncthreads();
  for (int o = threadIdx.x; o < no; o += kBlock)
    if (s_h[o]) atomicAdd(&b.c2[q0 + o], (int32_t)s_h[o]);
}

// Pool position of pair number p owned in graph g.
__device__ __forceinline__ int64_t pool_pos(const LaunchArgs& a, int g, int64_t p) {
  return a.descs[g].pci_off - a.descs[g].eb + p;
}

// Adds v into s_w[o] once per run of equal o across the wave's lanes; o is
// non-decreasing over the lanes that hold one (o < 0: none, at the end).
// All 64 lanes call it.
__device__ __forceinline__ void wave_run_add(unsigned long long* s_w, int o, unsigned long long v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long u = __shfl_up(v, d);
    const int ou = __shfl_up(o, d);
    if (lane >= d && ou == o) v += u;
  }
  const int on = __shfl_down(o, 1);
  if (o >= 0 && (lane == 63 || on != o)) atomicAdd(&s_w[o], v);
}

// Writes the n pairs sorted in LDS (pair s_idx[i] of iref, owner s_own[i])
// to the pool at dst(o, i), and adds each partner's list length to its
// owner's work s_w[o]; kUnroll list references in flight per lane (a
// partner given by its slot is read from lref).  kEdge: each pair's id
// position (iedge) goes to pedge at the same place.
template <bool kEdge, class Dst>
__device__ __forceinline__ void write_sorted(ListRef* pci, const ListRef* lref, int64_t n, const ListRef* iref,
                                             const uint16_t* s_idx, const uint16_t* s_own, unsigned long long* s_w,
                                             Dst dst, const uint32_t* iedge = nullptr, int64_t* pedge = nullptr) {
  constexpr int kUnroll = 8;
  for (int64_t i0 = 0; i0 < n; i0 += kUnroll * kBlock) {
    ListRef r[kUnroll];
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) {
      const int64_t i = i0 + (int64_t)j * kBlock + threadIdx.x;
      r[j] = i < n ? iref[s_idx[i]] : (ListRef)0;
    }
#pragma unroll
    for (int j = 0; j < kUnroll; ++j)
      if (r[j] & kRefSlot) r[j] = lref[(uint32_t)r[j]];
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) {
      const int64_t i = i0 + (int64_t)j * kBlock + threadIdx.x;
      int o = -1;
      if (i < n) {
        o = s_own[i];
        const int64_t at = dst(o, i);
        pci[at] = r[j];
        if constexpr (kEdge) pedge[at] = (int64_t)iedge[s_idx[i]];
      }
      wave_run_add(s_w, o, (unsigned long long)list_len(r[j]));
    }
  }
}

// Per fine bucket: partner offsets of its owners; a small bucket is placed
// whole, a big one gets its owners' cursors (absolute pool positions).
template <bool kEdge>
__global__ __launch_bounds__(kBlock) void radix_final_kernel(LaunchArgs a, BuildBufs b, int64_t RP, int fs, int cap) {
  extern __shared__ int64_t s_mem[];
  const int no_max = 1 << fs;
  unsigned long long* s_w = (unsigned long long*)s_mem;  // [no_max] work of each owner
  uint32_t* s_cnt = (uint32_t*)(s_w + no_max);  // [no_max]
  uint32_t* s_cur = s_cnt + no_max;             // [no_max]
  int32_t* s_g = (int32_t*)(s_cur + no_max);    // [no_max]
  uint16_t* s_idx = (uint16_t*)(s_g + no_max);  // [cap] place of each sorted pair in the bucket
  uint16_t* s_own = s_idx + cap;                // [cap]
  __shared__ int64_t s_wave[kWaves];
  const int64_t f = blockIdx.x;
  const int32_t q0 = (int32_t)(f << fs);
  const int no = (int)(min((int64_t)q0 + no_max, RP) - q0);
  const int64_t p0 = b.fst[f], p1 = b.fst[f + 1], n = p1 - p0;
  const bool big = n > cap;
  for (int o = threadIdx.x; o < no_max; o += kBlock) {
    s_cnt[o] = big && o < no ? (uint32_t)b.c2[q0 + o] : 0u;
    s_g[o] = o < no ? gid_graph(b.gid[q0 + o]) : 0;
    s_w[o] = 0ull;
  }
  __syncthreads();
  // a small bucket's owners stay in registers, with their ranks
  int32_t ko[kFinPer];
  uint32_t rk[kFinPer];
  if (!big) {
#pragma unroll
    for (int s = 0; s < kFinPer; ++s) {
      const int64_t p = p0 + (int64_t)s * kBlock + threadIdx.x;
      ko[s] = p < p1 ? b.key[p] - q0 : -1;
    }
#pragma unroll
    for (int s = 0; s < kFinPer; ++s) rk[s] = wave_count_add<2>(s_cnt, ko[s]);
    __syncthreads();
  }
  for (int o = threadIdx.x; o < no_max; o += kBlock) s_cur[o] = s_cnt[o];
  __syncthreads();
  if (no_max <= kBlock) block_scan_lds<1>(s_cur, no_max, s_wave);
  else if (no_max <= 4 * kBlock) block_scan_lds<4>(s_cur, no_max, s_wave);
  else block_scan_lds<kMaxFineOwners / kBlock>(s_cur, no_max, s_wave);
  for (int o = threadIdx.x; o < no; o += kBlock) {
    b.off[q0 + o] = p0 + s_cur[o];
    if (big) {
      b.cur[q0 + o] = (unsigned long long)pool_pos(a, s_g[o], p0 + s_cur[o]);
      b.wq[q0 + o] = 0ull;  // the chunks add to it
    }
  }
  if (big) return;
#pragma unroll
  for (int s = 0; s < kFinPer; ++s) {
    if (ko[s] < 0) continue;
    const uint32_t at = s_cur[ko[s]] + rk[s];
    s_idx[at] = (uint16_t)(s * kBlock + threadIdx.x);
    s_own[at] = (uint16_t)ko[s];
  }
  __syncthreads();
  write_sorted<kEdge>((ListRef*)a.pci, b.lref, n, b.ref + p0, s_idx, s_own, s_w,
                      [&](int o, int64_t i) { return pool_pos(a, s_g[o], p0 + i); }, kEdge ? b.e2 + p0 : nullptr,
                      b.pedge);
  __syncthreads();
  for (int o = threadIdx.x; o < no; o += kBlock) b.wq[q0 + o] = s_w[o];
}

// Big buckets, pass 2: each chunk sorted by owner in LDS; every owner's run
// reserved on its cursor and written out.
template <bool kEdge>
__global__ __launch_bounds__(kBlock) void radix_place_big_kernel(LaunchArgs a, BuildBufs b, int fs, int nf, int chunk) {
  extern __shared__ int64_t s_mem[];
  const int no = 1 << fs;
  unsigned long long* s_base = (unsigned long long*)s_mem;  // [no]
  unsigned long long* s_w = s_base + no;                    // [no] work added by this chunk
  uint32_t* s_cnt = (uint32_t*)(s_w + no);                  // [no]
  uint32_t* s_st = s_cnt + no;                              // [no]
  uint16_t* s_idx = (uint16_t*)(s_st + no);                 // [chunk] place of each sorted pair in the chunk
  uint16_t* s_own = s_idx + chunk;                          // [chunk]
  __shared__ int64_t s_wave[kWaves];
  Chunk c;
  if (!chunk_of(b, blockIdx.x, nf, chunk, c)) return;
  const int32_t q0 = c.f << fs;
  for (int o = threadIdx.x; o < no; o += kBlock) {
    s_cnt[o] = 0u;
    s_w[o] = 0ull;
  }
  __syncthreads();
  int32_t ko[kChunkPer];
  uint32_t rk[kChunkPer];
#pragma unroll
  for (int s = 0; s < kChunkPer; ++s) {
    const int64_t p = c.p0 + (int64_t)s * kBlock + threadIdx.x;
    ko[s] = p < c.p1 ? b.key[p] - q0 : -1;
  }
#pragma unroll
  for (int s = 0; s < kChunkPer; ++s) rk[s] = wave_count_add<2>(s_cnt, ko[s]);
  __syncthreads();
  for (int o = threadIdx.x; o < no; o += kBlock) {
    const uint32_t h = s_cnt[o];
    s_st[o] = h;
    if (h) s_base[o] = atomicAdd(&b.cur[q0 + o], (unsigned long long)h);
  }
  __syncthreads();
  if (no <= kBlock) block_scan_lds<1>(s_st, no, s_wave);
  else if (no <= 4 * kBlock) block_scan_lds<4>(s_st, no, s_wave);
  else block_scan_lds<kMaxFineOwners / kBlock>(s_st, no, s_wave);
#pragma unroll
  for (int s = 0; s < kChunkPer; ++s) {
    if (ko[s] < 0) continue;
    const uint32_t at = s_st[ko[s]] + rk[s];
    s_idx[at] = (uint16_t)(s * kBlock + threadIdx.x);
    s_own[at] = (uint16_t)ko[s];
  }
  __syncthreads();
  write_sorted<kEdge>((ListRef*)a.pci, b.lref, c.p1 - c.p0, b.ref + c.p0, s_idx, s_own, s_w,
                      [&](int o, int64_t i) { return (int64_t)s_base[o] + (i - s_st[o]); },
                      kEdge ? b.e2 + c.p0 : nullptr, b.pedge);
  __syncthreads();
  for (int o = threadIdx.x; o < no; o += kBlock)
    if (s_w[o]) atomicAdd(&b.wq[q0 + o], s_w[o]);
}

// Radix path of a local rebuild (one graph): the partner vertex of every
// entry of partner row k, from the entry's DAG id position e (pedge): the
// DAG edge is (esrc[e] without bit 31, ci[e]); the partner is the end that
// is not the row's owner pown[k].  One wave per row, grid-stride.
__global__ __launch_bounds__(kBlock) void partner_vid_kernel(LaunchArgs a, int64_t P, BuildBufs b) {
  const int lane = threadIdx.x & 63;
  const int64_t w0 = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6, nw = ((int64_t)gridDim.x * kBlock) >> 6;
  for (int64_t k = w0; k < P; k += nw) {
    if (a.vrow[k].set == 0) continue;
    const int32_t x = b.pown[k];
    const int64_t i0 = a.vrp[k], i1 = a.vrp[k + 1];
    for (int64_t i = i0 + lane; i < i1; i += 64) {
      const int64_t e = b.pedge[i];
      const int32_t r = b.esrc[e] & 0x7fffffff;
      b.pvid[i] = r == x ? a.ci[e] : r;
    }
  }
}

// Rows, row offsets and work after the partner pool is filled (both paths).
static hipError_t launch_partner_rows(const LaunchArgs& a, int64_t RP, const BuildBufs& b, hipStream_t s) {
  const unsigned blocks = (unsigned)((RP + kBlock - 1) / kBlock);
  hipError_t e;
  partner_rows_count_kernel<<<blocks, kBlock, 0, s>>>(a, RP, b);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_scan(LoadI32{b.kc}, StoreI64{b.voff}, RP, b.part, s)) != hipSuccess) return e;
  partner_rows_fill_kernel<<<blocks, kBlock, 0, s>>>(a, RP, b);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (a.P > 0) partner_work_kernel<<<(unsigned)((a.P * kBuildG + kBlock - 1) / kBlock), kBlock, 0, s>>>(a);
  return hipGetLastError();
}

template <bool kEdge>
static hipError_t launch_partners_radix(const LaunchArgs& a, int64_t RP, int64_t M, const BuildBufs& b,
                                        const RadixGeom& x, hipStream_t s) {
  const int cs = x.fs + 6;
  const unsigned rblocks = (unsigned)((RP + kBlock - 1) / kBlock);
  hipError_t e = launch_zero((void*)a.vrow, a.P * (int64_t)(sizeof(VRow) / 4), s);  // slots unused
  if (e == hipSuccess) e = launch_zero(b.cnt1 + (int64_t)x.nc * x.nmp, 1, s);
  if (e == hipSuccess) e = launch_zero(b.cnt2, kFine * x.nt2 + 1, s);
  if (e == hipSuccess) e = launch_zero(b.c2, RP, s);
  if (e != hipSuccess) return e;
  radix_prep_kernel<<<rblocks, kBlock, 0, s>>>(a, RP, M, b);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  radix_mpath_kernel<<<(unsigned)((x.nmp + kBlock) / kBlock), kBlock, 0, s>>>(b.erow, RP, M, x.nmp, b.bq);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  radix_split_kernel<<<(unsigned)x.nmp, kBlock, 0, s>>>(a, RP, M, b, cs, x.nc, x.nmp);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int64_t n1 = (int64_t)x.nc * x.nmp + 1;
  if ((e = launch_scan(LoadU32{b.cnt1}, StoreU32{b.cnt1}, n1, b.part, s)) != hipSuccess) return e;
  radix_scatter1_kernel<kEdge><<<(unsigned)x.nmp, kBlock, 0, s>>>(b, RP, M, cs, x.nc, x.nmp);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  radix_tiles2_kernel<<<1, kBlock, 0, s>>>(b, x.nc, x.nmp);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  radix_hist2_kernel<<<(unsigned)x.nt2, kBlock, 0, s>>>(b, x.fs, x.nc, x.nmp);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int64_t n2 = kFine * x.nt2 + 1;
  if ((e = launch_scan(LoadU32{b.cnt2}, StoreU32{b.cnt2}, n2, b.part, s)) != hipSuccess) return e;
  radix_scatter2_kernel<kEdge><<<(unsigned)x.nt2, kBlock, 0, s>>>(b, x.fs, x.nc, x.nmp);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  radix_chunks_kernel<<<1, kBlock, 0, s>>>(b, x.nf, x.nc, x.nmp, x.cap, x.chunk);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const size_t own_lds = (size_t)1 << x.fs;
  radix_count_big_kernel<<<(unsigned)x.nch, kBlock, 4 * own_lds, s>>>(b, x.fs, x.nf, x.chunk);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  radix_final_kernel<kEdge><<<(unsigned)x.nf, kBlock, radix_final_lds(x.fs, x.cap), s>>>(a, b, RP, x.fs, x.cap);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  radix_place_big_kernel<kEdge><<<(unsigned)x.nch, kBlock, radix_place_lds(x.fs, x.chunk), s>>>(a, b, x.fs, x.nf,
                                                                                                x.chunk);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_partner_rows(a, RP, b, s)) != hipSuccess) return e;
  if constexpr (kEdge) {
    if (a.P > 0) {
      const int64_t blocks = min((a.P * 64 + kBlock - 1) / kBlock, (int64_t)65536);
      partner_vid_kernel<<<(unsigned)blocks, kBlock, 0, s>>>(a, a.P, b);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

hipError_t launch_partners(const LaunchArgs& a, int64_t RP, int64_t M, const BuildBufs& bb, hipStream_t s) {
  if (RP <= 0) return hipSuccess;
  RadixGeom x;
  // A local rebuild (pedge) takes the radix path for one graph with the id
  // positions' scratch and every edge's row.
  const bool local = bb.pedge != nullptr;
  const bool radix_local = local && bb.e1 && bb.e2 && bb.esrc && bb.pvid && bb.pown && a.ngraphs == 1;
  if (bb.key && (!local || radix_local) && RP <= bb.rows && M <= bb.ids && radix_geom(RP, M, x)) {
    BuildBufs b = bb;
    b.wq = bb.cur + RP;  // second stripe of the cursors: owners' work
    return local ? launch_partners_radix<true>(a, RP, M, b, x, s) : launch_partners_radix<false>(a, RP, M, b, x, s);
  }
  BuildBufs b = bb;
  b.gid = nullptr;  // rows are found by their graph search (no prep pass here)
  b.wq = nullptr;   // and every row's work comes from the work pass
  const unsigned blocks = (unsigned)((RP + kBlock - 1) / kBlock);
  const unsigned gblocks = (unsigned)((RP * kBuildG + kBlock - 1) / kBlock);
  hipError_t e = launch_zero(b.c2, RP * kXcds, s);
  if (e == hipSuccess) e = launch_zero((void*)a.vrow, a.P * (int64_t)(sizeof(VRow) / 4), s);  // slots unused
  if (e != hipSuccess) return e;
  partner_count_kernel<<<gblocks, kBlock, 0, s>>>(a, RP, b);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_scan(LoadSum2{b.c1, b.c2, RP}, StoreI64{b.off}, RP, b.part, s)) != hipSuccess) return e;
  partner_cursor_kernel<<<blocks, kBlock, 0, s>>>(a, RP, b);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  partner_fill_kernel<<<gblocks, kBlock, 0, s>>>(a, RP, b);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return launch_partner_rows(a, RP, b, s);
}


// ---------------------------------------------------------------------------
// planning: bin every partner row, then scatter (graph, row) items by bin
// ---------------------------------------------------------------------------
// Bin of a partner row: its set is N+(x) (s ids), its partners' lists hold
// `work` ids in all.
__device__ __forceinline__ int choose_bin(int64_t np, int64_t s, int64_t work, const Tune& t) {
  if (s < 1 || np < 1 || work == 0) return kNoBin;
  if (s > kCapBlock - 1) return BIN_GLOBAL;
  if (s > kCapWave - 1 || work > t.t_wave) return BIN_BLOCK;
  if (s > kCapG32 - 1 || work > t.t_g32) return BIN_WAVE;
  if (s > kCapG16 - 1 || work > t.t_g16) return BIN_G32;
  if (s > kCapG8 - 1 || work > t.t_g8) return BIN_G16;
  return BIN_G8;
}

// A plan index's partner row: graph, owner's list, work, partner range.
struct PRow {
  int g;
  ListRef set;  // 0: unused slot
  uint32_t work;
  int64_t pb, pe;
};

// The row's offsets are loaded beside its record, not after it (vrp has
// P + 1 entries, so k + 1 is in bounds for every k < P): the two loads are
// in flight together.  pb and pe mean nothing for an unused slot (set 0).
__device__ __forceinline__ PRow partner_row(const LaunchArgs& a, int64_t k) {
  const VRow vr = a.vrow[k];
  return PRow{vr.g, vr.set, vr.work, a.vrp[k], a.vrp[k + 1]};
}

// Work class of a binned partner row: its estimated cost in quarter octaves
// (4 floor(log2 c) + the two bits below the leading one), capped.  Within a
// planning block a bin's rows are placed in class order, so the groups that
// share a wave (8, 4 or 2 rows in lockstep) get rows of similar cost and
// fewer of their lanes idle while the heaviest row finishes.
// The cost estimate weighs the streamed ids (work) against the per-edge
// batch bookkeeping (np partners) and the set build (s ids).
constexpr int kWorkClasses = 64;
__device__ __forceinline__ int work_class(uint32_t work, int64_t np, int s) {
  const uint64_t c = (uint64_t)work + 8 * (uint64_t)np + 4 * (uint64_t)s;
  const uint64_t x = c > 0 ? c : 1;
  const int e = 63 - __clzll((long long)x);
  const int q = 4 * e + (int)(((x << 2) >> e) & 3);
  return min(kWorkClasses - 1, q);
}

// Pass 1: bin and work class of every plan index (bin_of = class << 3 | bin)
// and, per planning block, the count per bin (bin-major:
// blk[bin * nblk + block]).
__global__ __launch_bounds__(kBlock) void plan_bin_kernel(LaunchArgs a) {
  __shared__ uint32_t s_cnt[kNumBins];
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (threadIdx.x < kNumBins) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  if (p < a.P) {
    const PRow pr = partner_row(a, p);
    const int bin = pr.set ? choose_bin(pr.pe - pr.pb, list_len(pr.set), pr.work, a.tune) : (int)kNoBin;
    a.bufs.bin_of[p] = bin == kNoBin ? (uint16_t)kNoBin
                                     : (uint16_t)(work_class(pr.work, pr.pe - pr.pb, list_len(pr.set)) << 3 | bin);
    if (bin != kNoBin) atomicAdd(&s_cnt[bin], 1u);
  }
  __syncthreads();
  const int64_t nblk = gridDim.x;
  if (threadIdx.x < kNumBins) a.bufs.blk[threadIdx.x * nblk + blockIdx.x] = s_cnt[threadIdx.x];
  if (blockIdx.x == 0 && threadIdx.x == 0) a.bufs.blk[kNumBins * nblk] = 0;  // scanned to the total
}

// Pass 2 (after the scan of blk): bin totals into the counters.
__global__ void plan_totals_kernel(LaunchArgs a, int64_t nblk) {
  const int b = threadIdx.x;
  if (b < kNumBins) a.bufs.counters[kCntBinCount + b] = a.bufs.blk[(b + 1) * nblk] - a.bufs.blk[b * nblk];
}

// Pass 3: scatter of (graph, row) items into their bin segment.  A planning
// block keeps its place in the segment (so items stay grouped by graph); its
// items are ordered by work class (any order within a class).
__global__ __launch_bounds__(kBlock) void plan_scatter_kernel(LaunchArgs a) {
  constexpr int kKeys = kNumBins * kWorkClasses;
  __shared__ uint32_t s_key[kKeys];  // per (bin, class): count, then next rank
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int code = (p < a.P) ? (int)a.bufs.bin_of[p] : (int)kNoBin;
  const bool has = code != kNoBin;
  const int bin = code & 7, key = bin * kWorkClasses + (code >> 3);
  for (int i = threadIdx.x; i < kKeys; i += kBlock) s_key[i] = 0;
  __syncthreads();
  if (has) atomicAdd(&s_key[key], 1u);
  __syncthreads();
  if (threadIdx.x < kNumBins) {  // per bin: exclusive scan over its classes
    uint32_t run = 0;
    for (int c = 0; c < kWorkClasses; ++c) {
      const uint32_t n = s_key[threadIdx.x * kWorkClasses + c];
      s_key[threadIdx.x * kWorkClasses + c] = run;
      run += n;
    }
  }
  __syncthreads();
  if (has) {
    const int64_t nblk = gridDim.x;
    const uint32_t rank = atomicAdd(&s_key[key], 1u);
    a.bufs.items[a.bufs.blk[bin * nblk + blockIdx.x] + rank] = (uint64_t)p;
  }
}

// ---------------------------------------------------------------------------
// persistent count
// ---------------------------------------------------------------------------
struct __align__(16) CountLds {
  int32_t A[kAreaBlock];         // pivot-set areas (one per group of the tile's bin)
  int32_t End[kBlock];           // per-lane edge end in the batch's flat index
  gint32* Base[kBlock];          // per-lane N+(v) base of that edge, shifted by its first index
  uint64_t red[kWaves];
  int64_t tab_cnt[1];
  int64_t cur_first, cur_seg;
};
static_assert(kAreaWave * kWaves <= kAreaBlock, "wave areas exceed LDS region");
static_assert(kAreaG32 * (kBlock / 32) <= kAreaBlock, "g32 areas exceed LDS region");
static_assert(kAreaG16 * (kBlock / 16) <= kAreaBlock, "g16 areas exceed LDS region");
static_assert(kAreaG8 * (kBlock / 8) <= kAreaBlock, "g8 areas exceed LDS region");
static_assert(kCapG8 <= kAreaG8 && kCapG16 <= kAreaG16 && kCapG32 <= kAreaG32 &&
              kCapWave <= kAreaWave && kCapBlock <= kAreaBlock, "a bin's tree must fit its area");

// Per-lane running count; in batch mode it is flushed to outs[g] whenever
// the lane moves to a pivot of another graph.
// Part of vertex u under split points b[0..k] (b[0] = 0 <= u < b[k]): the
// last p with b[p] <= u, which skips empty parts.
__device__ __forceinline__ int part_of(const int32_t* b, int k, int64_t u) {
  int lo = 0, hi = k;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (b[mid] <= u) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Part of a partner row's owner under the call's split (kLocalParts).  The
// PE that computes a DAG edge in the reference's push/pull kernel
// (successful_push_pull/bs_warp_vertex_nvshmem.cuh:127-136: u's part when v
// is local or d+(u) >= d+(v), else v's) is always its owner's part: the
// owner is u exactly when d+(v) <= d+(u), and when v owns it u and v are in
// one part or the edge is pushed to v's.  So a row's hits all go to one
// part.
__device__ __forceinline__ int owner_part(const LaunchArgs& a, int64_t item) {
  return part_of(a.local.bounds, a.local.parts, a.local.pown[item]);
}

struct Acc {
  uint64_t sum = 0;
  int g = -1;
  __device__ __forceinline__ void to_graph(int ng, bool batch, uint64_t* outs) {
    if (batch && ng != g) {
      if (sum) atomicAdd(&outs[g], sum);
      sum = 0;
      g = ng;
    }
  }
};

// LDS of a local count beside CountLds: per-slot hit counters parallel to
// the set areas, per-batch-edge hit counters parallel to End.
template <int kLocal>
struct LocalLds {
  __device__ uint32_t* c() { return nullptr; }
  __device__ uint32_t* ec() { return nullptr; }
};
template <>
struct __align__(16) LocalLds<kLocalFull> {
  uint32_t C[kAreaBlock / 2];
  uint32_t EC[kBlock];
  __device__ uint32_t* c() { return C; }
  __device__ uint32_t* ec() { return EC; }
};
template <>
struct __align__(16) LocalLds<kLocalEdges> {
  uint32_t EC[kBlock];
  __device__ uint32_t* c() { return nullptr; }
  __device__ uint32_t* ec() { return EC; }
};

template <int G, int AREA, int kLocal, bool kWide>
__device__ __forceinline__ void group_tile(const LaunchArgs& a, const uint64_t* items, int64_t first,
                                           int64_t last, CountLds& L, LocalLds<kLocal>& LL, bool batch,
                                           Acc& acc) {
  constexpr int NGRP = kBlock / G;
  // With G = 64 a group is a wave: its pivot and everything derived from it
  // is wave-uniform and lives in scalar registers.
  const int grp = G == 64 ? uniform32(threadIdx.x / 64) : threadIdx.x / G, gl = threadIdx.x % G;
  int32_t* slot = L.A + grp * AREA;
  int32_t* sEnd = L.End + grp * G;
  gint32** sBase = L.Base + grp * G;
  constexpr int kEL = edge_local(kLocal);
  uint32_t* C = kLocal == kLocalFull ? LL.c() + grp * (AREA / 2) : nullptr;
  for (int64_t i = first + grp; i < last; i += NGRP) {
    const int64_t item = (int64_t)items[i];
    const PRow pr = partner_row(a, item);
    const PivotSet S = build_pivot_set<G, kEL, kWide>(slot, AREA, a.ci + list_start(pr.set), list_len(pr.set),
                                                      gl, a.tune, a.bufs.counters + kCntFallback, C);
    const LocalSink ls{C, kEL ? LL.ec() + grp * G : nullptr, kEL ? a.local.pedge + pr.pb : nullptr,
                       kEL ? a.local.pvid + pr.pb : nullptr, a.local.ecount, a.local.vcount};
    const uint32_t h = pivot_edges<G, kModeHash | kModeTree, G == 64 ? kInFlight : kInFlightGroup, kEL, kWide>(
        a.ci, RefEdges{a.pci + pr.pb}, (int)(pr.pe - pr.pb), S, 0, G, sEnd, sBase, gl, ls);
    if constexpr (kLocal == kLocalFull) {
      group_sync<G>();
      flush_set_hits<G>(slot, C, S, gl, a.local.vcount);
      const uint32_t hr = group_sum<G>(h);  // the row's triangles: each has the owner as a corner
      if (gl == 0 && hr) atomicAdd(&a.local.vcount[a.local.pown[item]], (unsigned long long)hr);
    }
    if constexpr (kLocal == kLocalParts) acc.to_graph(owner_part(a, item), true, a.local.pe);
    else acc.to_graph(pr.g, batch, a.bufs.outs);
    acc.sum += h;
    wave_sync();
  }
}

template <bool kLdsPivot, int kLocal, bool kWide>
__device__ __forceinline__ void block_pivot(const LaunchArgs& a, uint64_t item, CountLds& L,
                                            LocalLds<kLocal>& LL, bool batch, Acc& acc) {
  const PRow pr = partner_row(a, (int64_t)item);
  const int32_t* ci = a.ci;
  const int32_t* set = ci + list_start(pr.set);
  const int ds = list_len(pr.set);
  const RefEdges part{a.pci + pr.pb};
  const int np = (int)(pr.pe - pr.pb);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int kEL = edge_local(kLocal);
  uint32_t* EC = kEL ? LL.ec() + wid * 64 : nullptr;
  const int64_t* pedge = kEL ? a.local.pedge + pr.pb : nullptr;
  const int32_t* pvid = kLocal == kLocalFull ? a.local.pvid + pr.pb : nullptr;
  uint32_t h;
  if (kLdsPivot) {
    const PivotSet S = build_pivot_set<kBlock, kEL, kWide>(L.A, kAreaBlock, set, ds, threadIdx.x, a.tune,
                                                           a.bufs.counters + kCntFallback, LL.c());
    const LocalSink ls{LL.c(), EC, pedge, pvid, a.local.ecount, a.local.vcount};
    h = pivot_edges<64, kModeHash | kModeTree, kInFlight, kEL, kWide>(ci, part, np, S, wid * 64, kBlock,
                                                                       L.End + wid * 64, L.Base + wid * 64, lane,
                                                                       ls);
    if constexpr (kLocal == kLocalFull) {
      __syncthreads();
      flush_set_hits<kBlock>(L.A, LL.c(), S, threadIdx.x, a.local.vcount);
    }
  } else {
    const PivotSet S{set, kSetHbm, ds, 0};
    const LocalSink ls{nullptr, EC, pedge, pvid, a.local.ecount, a.local.vcount};
    h = pivot_edges<64, kModeHbm, kInFlight, kEL>(ci, part, np, S, wid * 64, kBlock, L.End + wid * 64,
                                                  L.Base + wid * 64, lane, ls);
  }
  if constexpr (kLocal == kLocalFull) {  // the owner's share, one atomic per wave
    const uint32_t hw = group_sum<64>(h);
    if (lane == 0 && hw) atomicAdd(&a.local.vcount[a.local.pown[item]], (unsigned long long)hw);
  }
  if constexpr (kLocal == kLocalParts) acc.to_graph(owner_part(a, (int64_t)item), true, a.local.pe);
  else acc.to_graph(pr.g, batch, a.bufs.outs);
  acc.sum += h;
}

// Minimum resident waves per SIMD the compiler must allow for (caps VGPRs:
// 8 waves = 64 VGPRs).  Latency hiding dominates this kernel's speed.
#ifndef TC_COUNT_WAVES_PER_EU
#define TC_COUNT_WAVES_PER_EU 8
#endif

// Full local variants hold 1.5x the LDS of a plain count (5 workgroups per
// CU); edges-only ones the plain count's plus 1 KB of edge counters (7).
#ifndef TC_LOCAL_EDGES_WAVES_PER_EU
#define TC_LOCAL_EDGES_WAVES_PER_EU 7
#endif
constexpr int kLocalWavesPerEU = 5;
constexpr int waves_per_eu(int kLocal) {
  return kLocal == kLocalFull ? kLocalWavesPerEU
       : kLocal == kLocalEdges ? TC_LOCAL_EDGES_WAVES_PER_EU : TC_COUNT_WAVES_PER_EU;  // parts: the plain count's
}

// One persistent kernel per bin, specialised so that each keeps only its own
// working set in registers.  Blocks pull tiles from the bin's queue head.
// kLocal: also the per-edge (and, kLocalFull, per-vertex) attribution of
// a.local (one graph only).
// kWide: the instantiation for graphs whose ids reach 2^24 (folded hash keys).
template <int BIN, int kLocal, bool kWide>
__global__ __launch_bounds__(kBlock, waves_per_eu(kLocal)) void count_kernel(
    LaunchArgs a) {
  __shared__ CountLds L;
  __shared__ LocalLds<kLocal> LL;
  // batch: per-graph sums; kLocalParts: per-part sums, kept the same way
  const bool batch = (!kLocal && a.ngraphs > 1) || kLocal == kLocalParts;
  if constexpr (edge_local(kLocal) != kLocalNone) LL.EC[threadIdx.x] = 0;  // ordered by the barrier below
  if (threadIdx.x == 0) {
    int64_t seg = 0;
    for (int b = 0; b < BIN; ++b) seg += a.bufs.counters[kCntBinCount + b];
    L.cur_seg = seg;
    L.tab_cnt[0] = a.bufs.counters[kCntBinCount + BIN];
  }
  __syncthreads();
  const int piv = tile_pivots(BIN, a.tune.rounds > 0 ? a.tune.rounds : 1);
  const int64_t cnt = L.tab_cnt[0];
  const int64_t ntiles = (cnt + piv - 1) / piv;
  const uint64_t* items = a.bufs.items + L.cur_seg;
  Acc acc;
  acc.g = 0;
  for (;;) {
    if (threadIdx.x == 0) L.cur_first = atomicAdd(&a.bufs.counters[kCntTileHead + BIN], 1u);
    __syncthreads();
    const int64_t t = uniform64(L.cur_first);
    if (t >= ntiles) break;
    const int64_t first = t * piv;
    const int64_t last = min(first + (int64_t)piv, cnt);
    if (BIN == BIN_GLOBAL) block_pivot<false, kLocal, kWide>(a, items[first], L, LL, batch, acc);
    else if (BIN == BIN_BLOCK) block_pivot<true, kLocal, kWide>(a, items[first], L, LL, batch, acc);
    else if (BIN == BIN_WAVE) group_tile<64, kAreaWave, kLocal, kWide>(a, items, first, last, L, LL, batch, acc);
    else if (BIN == BIN_G32) group_tile<32, kAreaG32, kLocal, kWide>(a, items, first, last, L, LL, batch, acc);
    else if (BIN == BIN_G16) group_tile<16, kAreaG16, kLocal, kWide>(a, items, first, last, L, LL, batch, acc);
    else group_tile<8, kAreaG8, kLocal, kWide>(a, items, first, last, L, LL, batch, acc);
    if (batch) {
      // Items are ordered by graph (parts: by owner, so by part), so a tile
      // almost always belongs to one: then one block-reduced atomic;
      // otherwise one per lane.
      uint64_t* outs = kLocal == kLocalParts ? a.local.pe : a.bufs.outs;
      const int tg = kLocal == kLocalParts ? owner_part(a, (int64_t)items[first]) : a.vrow[items[first]].g;
      if (!__syncthreads_or(acc.sum != 0 && acc.g != tg)) {
        const uint64_t tot = block_sum(acc.sum, L.red);
        if (threadIdx.x == 0 && tot) atomicAdd(&outs[tg], tot);
      } else if (acc.sum) {
        atomicAdd(&outs[acc.g], acc.sum);
      }
      acc.sum = 0;
    }
    __syncthreads();
  }
  if (!batch) {
    const uint64_t tot = block_sum(acc.sum, L.red);
    if (threadIdx.x == 0 && tot) atomicAdd(&a.bufs.outs[0], tot);
  }
}

// ---------------------------------------------------------------------------
// local count, second pass: per-edge counts to vertices and parts
// ---------------------------------------------------------------------------

// Per-graph table of the per-part pass, first half: esrc[e] = the row u of
// each DAG edge e = (u, v).  The local rebuild then sets bit 31 where v owns
// the edge, i.e. d+(u) < d+(v) (the owner rule), which is when the
// reference's push/pull kernel pushes the edge to v's part.  A wave walks
// the edges of 64 consecutive rows as one flattened range (its lanes stay
// busy whatever the row lengths; a lane finds its row by a binary search
// over the rows' degree prefix in LDS).
__global__ __launch_bounds__(kBlock) void local_prep_kernel(const int64_t* __restrict__ rp, int64_t n,
                                                            int32_t* __restrict__ esrc) {
  __shared__ int64_t s_e0[kBlock];
  __shared__ int32_t s_end[kBlock], s_du[kBlock];
  const int lane = threadIdx.x & 63, w0 = threadIdx.x & ~63;
  const int64_t u = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int32_t deg = 0;
  if (u < n) {
    const int64_t e0 = rp[u];
    deg = (int32_t)(rp[u + 1] - e0);  // a row holds fewer than 2^31 / 64 ids
    s_e0[threadIdx.x] = e0;
  }
  int total;
  const int before = group_excl_scan<64>(deg, lane, total);
  s_end[threadIdx.x] = before + deg;
  s_du[threadIdx.x] = deg;
  wave_sync();
  for (int t = lane; t < total; t += 64) {
    const int r = w0 + find_edge<64>(s_end + w0, t);
    esrc[s_e0[r] + (t - (s_end[r] - s_du[r]))] = (int32_t)(blockIdx.x * kBlock + r);
  }
}

// Every counted DAG edge e = (u, v) in [e0, e1) with count c goes to the
// part that computes it in the reference's push/pull kernel
// (successful_push_pull/bs_warp_vertex_nvshmem.cuh:127-136): u's part when v
// is in the same part or d+(u) >= d+(v) (pull), else v's part (push).
// Uncounted edges hold c = 0.  Workgroups stream chunks of kPartsChunk
// consecutive edges, reading only streams: c, then u and the push bit from
// esrc, v from ci (kPartsUnroll edges per lane at a time, all loads issued
// before any is used).  Parts come from the split points in LDS: the lane
// keeps its last row's part and range in registers (a chunk's rows are
// consecutive), so only a pushed edge's head outside that range is
// searched.  Each lane sums runs of edges with the same owner part before
// one LDS atomic.
constexpr int kPartsChunk = kBlock * 32;
constexpr int kPartsUnroll = 4;
__global__ __launch_bounds__(kBlock) void local_parts_kernel(const int32_t* __restrict__ ci,
                                                             const int32_t* __restrict__ esrc,
                                                             const int32_t* __restrict__ bounds, int32_t parts,
                                                             int64_t e0, int64_t e1,
                                                             const uint32_t* __restrict__ ecount,
                                                             unsigned long long* pe) {
  __shared__ unsigned long long s_pe[kMaxParts];
  __shared__ int32_t s_b[kMaxParts + 1];
  for (int i = threadIdx.x; i < parts; i += kBlock) s_pe[i] = 0;
  for (int i = threadIdx.x; i <= parts; i += kBlock) s_b[i] = bounds[i];
  __syncthreads();
  int run_owner = 0;
  unsigned long long run = 0;
  int pu = 0, lo = 0, hi = 0;  // the last row's part and its vertex range [lo, hi)
  for (int64_t b0 = e0 + (int64_t)blockIdx.x * kPartsChunk; b0 < e1; b0 += (int64_t)gridDim.x * kPartsChunk) {
    const int64_t b1 = min(b0 + (int64_t)kPartsChunk, e1);
    for (int64_t eb = b0 + threadIdx.x; eb < b1; eb += (int64_t)kBlock * kPartsUnroll) {
      uint32_t c[kPartsUnroll], x[kPartsUnroll];
      int v[kPartsUnroll];
#pragma unroll
      for (int j = 0; j < kPartsUnroll; ++j) {
        const int64_t e = eb + (int64_t)j * kBlock;
        const bool in = e < b1;
        c[j] = in ? ecount[e] : 0u;
        x[j] = in ? (uint32_t)esrc[e] : 0u;
        v[j] = in ? ci[e] : 0;
      }
#pragma unroll
      for (int j = 0; j < kPartsUnroll; ++j) {
        if (c[j] == 0) continue;
        const int u = (int)(x[j] & 0x7fffffffu);
        if (u < lo || u >= hi) {
          pu = part_of(s_b, parts, u);
          lo = s_b[pu];
          hi = s_b[pu + 1];
        }
        int owner = pu;
        if ((x[j] >> 31) && (v[j] < lo || v[j] >= hi)) owner = part_of(s_b, parts, v[j]);
        if (owner != run_owner) {
          if (run) atomicAdd(&s_pe[run_owner], run);
          run_owner = owner;
          run = 0;
        }
        run += c[j];
      }
    }
  }
  if (run) atomicAdd(&s_pe[run_owner], run);
  __syncthreads();
  for (int i = threadIdx.x; i < parts; i += kBlock)
    if (s_pe[i]) atomicAdd(&pe[i], s_pe[i]);
}

// ---------------------------------------------------------------------------
// plain kernel: one wave64 per pivot, grid-stride (bs_warp_vertex.cuh:2-35)
// ---------------------------------------------------------------------------
struct __align__(16) PlainLds {
  int32_t A[kWaves][kCapWave];
  int32_t End[kBlock];
  gint32* Base[kBlock];
  uint64_t red[kWaves];
};

__global__ __launch_bounds__(kBlock, TC_COUNT_WAVES_PER_EU) void plain_kernel(const int64_t* __restrict__ rp,
                                                       const int32_t* __restrict__ ci,
                                                       int32_t row_begin, int32_t row_end,
                                                       const uint8_t* own, uint64_t* out,
                                                       uint64_t* zero_next) {
  __shared__ PlainLds L;
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // the next call's word, cleared here instead of by a launch of its own
  if (zero_next && blockIdx.x == 0 && threadIdx.x == 0) *zero_next = 0;
  const int64_t nw = (int64_t)gridDim.x * kWaves;
  uint64_t acc = 0;
  for (int64_t u = row_begin + (int64_t)blockIdx.x * kWaves + wid; u < row_end; u += nw) {
    const int64_t rb = rp[u];
    const int da = (int)(rp[u + 1] - rb);
    if (da < 2 || (own && !own[u])) continue;
    constexpr int kModes = kModeTree | kModeHbm;
    constexpr int kPlainInFlight = 6;
    if (da <= kCapWave - 1) {
      stage_tree(L.A[wid], ci + rb, da, lane, 64);
      wave_sync();
      const PivotSet S{L.A[wid], kSetTree, tree_levels(da), 0};
      acc += pivot_edges<64, kModes, kPlainInFlight>(ci, DagEdges{rp, ci + rb}, da, S, 0, 64, L.End + wid * 64,
                                                     L.Base + wid * 64, lane);
      wave_sync();
    } else {
      const PivotSet S{ci + rb, kSetHbm, da, 0};
      acc += pivot_edges<64, kModes, kPlainInFlight>(ci, DagEdges{rp, ci + rb}, da, S, 0, 64, L.End + wid * 64,
                                                     L.Base + wid * 64, lane);
    }
  }
  const uint64_t tot = block_sum(acc, L.red);
  if (threadIdx.x == 0 && tot) atomicAdd(out, tot);
}

// ---------------------------------------------------------------------------
// plain streaming copy: known HBM bytes for calibrating traffic counters
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                      int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    dst[i] = src[i];
}

// plain streaming read: the practical HBM read rate, the ceiling of a kernel
// that only reads.  Eight 16-byte nontemporal loads in flight per lane (a
// sweep of loads in flight, grid size and the hint on MI355X found this the
// fastest form).  The xor of the words is stored only if it equals a value
// the probe's byte fill cannot produce, which keeps every load and writes
// nothing.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kReadUnroll = 8;
__global__ __launch_bounds__(kBlock) void read_kernel(const u32x4* __restrict__ src, int64_t n,
                                                      uint32_t* __restrict__ sink) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  uint32_t acc = 0;
  for (; i + (kReadUnroll - 1) * stride < n; i += kReadUnroll * stride) {
    u32x4 x[kReadUnroll];
#pragma unroll
    for (int k = 0; k < kReadUnroll; ++k) x[k] = __builtin_nontemporal_load(src + i + k * stride);
#pragma unroll
    for (int k = 0; k < kReadUnroll; ++k) acc ^= x[k].x ^ x[k].y ^ x[k].z ^ x[k].w;
  }
  for (; i < n; i += stride) {
    const u32x4 x = __builtin_nontemporal_load(src + i);
    acc ^= x.x ^ x.y ^ x.z ^ x.w;
  }
  if (acc == 0x9E3779B9u) sink[(int64_t)blockIdx.x * kBlock + threadIdx.x] = acc;
}

// Staging copy of a server launch: a small persistent grid strides over the
// chunks (chunk b of all copies: the copy with the last first <= b, one
// kStageChunk of it).  Each lane keeps 16-byte loads in flight across the
// host link; few workgroups hold the CU slots they wait in, so the count
// kernels of other launches keep their occupancy.
__global__ __launch_bounds__(kBlock) void stage_copy_kernel(const StageCopy* copies, int32_t n, int64_t chunks) {
  for (int64_t b = blockIdx.x; b < chunks; b += gridDim.x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (copies[mid].first <= b) lo = mid;
      else hi = mid - 1;
    }
    const StageCopy c = copies[lo];
    const int64_t c0 = (b - c.first) * kStageChunk;
    if (c0 >= c.bytes) continue;
    const int64_t c1 = min(c0 + kStageChunk, c.bytes);
    const int64_t v1 = c0 + ((c1 - c0) & ~(int64_t)15);  // 16-byte part of this chunk
    constexpr int kU = (int)(kStageChunk / (16 * kBlock));  // 16-byte words per lane
    u32x4 x[kU];
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int64_t i = c0 + ((int64_t)k * kBlock + threadIdx.x) * 16;
      if (i < v1) x[k] = *(const u32x4*)(c.src + i);
    }
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int64_t i = c0 + ((int64_t)k * kBlock + threadIdx.x) * 16;
      if (i < v1) *(u32x4*)(c.dst + i) = x[k];
    }
    for (int64_t i = v1 + 4 * (int64_t)threadIdx.x; i < c1; i += 4 * kBlock)
      *(uint32_t*)(c.dst + i) = *(const uint32_t*)(c.src + i);
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_stage_copy(const StageCopy* copies, int32_t n, int64_t chunks, hipStream_t s) {
  if (n <= 0 || chunks <= 0) return hipSuccess;
  const int64_t grid = chunks < kStageGrid ? chunks : kStageGrid;
  stage_copy_kernel<<<(unsigned)grid, kBlock, 0, s>>>(copies, n, chunks);
  return hipGetLastError();
}

hipError_t launch_zero(void* p, int64_t words32, hipStream_t s) {
  if (words32 <= 0) return hipSuccess;
  int64_t blocks = (words32 + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  zero_kernel<<<(unsigned)blocks, kBlock, 0, s>>>((uint32_t*)p, words32);
  return hipGetLastError();
}

__global__ void rearm_kernel(uint32_t* counters, uint64_t* outs, int32_t nouts) {
  const int i = threadIdx.x;
  if (i < kNumBins) counters[kCntTileHead + i] = 0u;
  if (i == 0) counters[kCntFallback] = 0u;
  for (int j = i; j < nouts; j += blockDim.x) outs[j] = 0ull;
}

hipError_t launch_rearm(uint32_t* counters, uint64_t* outs, int32_t nouts, hipStream_t s) {
  rearm_kernel<<<1, 64, 0, s>>>(counters, outs, nouts);
  return hipGetLastError();
}

hipError_t launch_plan(const LaunchArgs& a, hipStream_t s) {
  if (a.P <= 0) return hipSuccess;
  const int64_t nblk = plan_blocks(a.P);
  plan_bin_kernel<<<(unsigned)nblk, kBlock, 0, s>>>(a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_scan(LoadU32{a.bufs.blk}, StoreU32{a.bufs.blk}, kNumBins * nblk + 1, a.bufs.part, s);
  if (e != hipSuccess) return e;
  plan_totals_kernel<<<1, 64, 0, s>>>(a, nblk);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  plan_scatter_kernel<<<(unsigned)nblk, kBlock, 0, s>>>(a);
  return hipGetLastError();
}

template <int BIN, int kLocal, bool kWide>
static int bin_blocks_per_cu() {
  static const int n = [] {
    int k = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, count_kernel<BIN, kLocal, kWide>, kBlock, 0) !=
        hipSuccess)
      k = 1;
    return k > 0 ? k : 1;
  }();
  return n;
}

// tiles >= 0: the bin's tile count is known on the host; the grid is then
// capped at it, and an empty bin is not launched.
template <int BIN, int kLocal = kLocalNone>
static void launch_bin(const LaunchArgs& a, int cus, int grid, hipStream_t s, int64_t tiles = -1) {
  if (tiles == 0) return;
  const bool wide = a.wide_ids != 0;
  int64_t g = grid > 0 ? grid
                       : (int64_t)(wide ? bin_blocks_per_cu<BIN, kLocal, true>() : bin_blocks_per_cu<BIN, kLocal, false>()) * cus;
  if (tiles > 0 && tiles < g) g = tiles;
  if (wide) count_kernel<BIN, kLocal, true><<<(unsigned)g, kBlock, 0, s>>>(a);
  else count_kernel<BIN, kLocal, false><<<(unsigned)g, kBlock, 0, s>>>(a);
}

hipError_t launch_count(const LaunchArgs& a, int cus, int grid, hipStream_t s) {
  launch_bin<BIN_GLOBAL>(a, cus, grid, s);
  launch_bin<BIN_BLOCK>(a, cus, grid, s);
  launch_bin<BIN_WAVE>(a, cus, grid, s);
  launch_bin<BIN_G32>(a, cus, grid, s);
  launch_bin<BIN_G16>(a, cus, grid, s);
  launch_bin<BIN_G8>(a, cus, grid, s);
  return hipGetLastError();
}

hipError_t launch_count_bins(const LaunchArgs& a, int cus, int grid, int first, int last, hipStream_t s) {
  if (first <= BIN_GLOBAL && BIN_GLOBAL <= last) launch_bin<BIN_GLOBAL>(a, cus, grid, s);
  if (first <= BIN_BLOCK && BIN_BLOCK <= last) launch_bin<BIN_BLOCK>(a, cus, grid, s);
  if (first <= BIN_WAVE && BIN_WAVE <= last) launch_bin<BIN_WAVE>(a, cus, grid, s);
  if (first <= BIN_G32 && BIN_G32 <= last) launch_bin<BIN_G32>(a, cus, grid, s);
  if (first <= BIN_G16 && BIN_G16 <= last) launch_bin<BIN_G16>(a, cus, grid, s);
  if (first <= BIN_G8 && BIN_G8 <= last) launch_bin<BIN_G8>(a, cus, grid, s);
  return hipGetLastError();
}

hipError_t launch_count_sized(const LaunchArgs& a, int cus, int grid, const int64_t* tiles, hipStream_t s) {
  launch_bin<BIN_GLOBAL>(a, cus, grid, s, tiles[BIN_GLOBAL]);
  launch_bin<BIN_BLOCK>(a, cus, grid, s, tiles[BIN_BLOCK]);
  launch_bin<BIN_WAVE>(a, cus, grid, s, tiles[BIN_WAVE]);
  launch_bin<BIN_G32>(a, cus, grid, s, tiles[BIN_G32]);
  launch_bin<BIN_G16>(a, cus, grid, s, tiles[BIN_G16]);
  launch_bin<BIN_G8>(a, cus, grid, s, tiles[BIN_G8]);
  return hipGetLastError();
}

hipError_t launch_count_order(const LaunchArgs& a, int cus, int grid, const int* order, int n,
                              const int64_t* tiles, hipStream_t s) {
  for (int i = 0; i < n; ++i) {
    const int b = order[i];
    const int64_t t = tiles ? tiles[b] : -1;
    switch (b) {
      case BIN_GLOBAL: launch_bin<BIN_GLOBAL>(a, cus, grid, s, t); break;
      case BIN_BLOCK: launch_bin<BIN_BLOCK>(a, cus, grid, s, t); break;
      case BIN_WAVE: launch_bin<BIN_WAVE>(a, cus, grid, s, t); break;
      case BIN_G32: launch_bin<BIN_G32>(a, cus, grid, s, t); break;
      case BIN_G16: launch_bin<BIN_G16>(a, cus, grid, s, t); break;
      default: launch_bin<BIN_G8>(a, cus, grid, s, t); break;
    }
  }
  return hipGetLastError();
}

template <int kLocal>
static void launch_local_bins(const LaunchArgs& a, int cus, int first, int last, hipStream_t s) {
  if (first <= BIN_GLOBAL && BIN_GLOBAL <= last) launch_bin<BIN_GLOBAL, kLocal>(a, cus, 0, s);
  if (first <= BIN_BLOCK && BIN_BLOCK <= last) launch_bin<BIN_BLOCK, kLocal>(a, cus, 0, s);
  if (first <= BIN_WAVE && BIN_WAVE <= last) launch_bin<BIN_WAVE, kLocal>(a, cus, 0, s);
  if (first <= BIN_G32 && BIN_G32 <= last) launch_bin<BIN_G32, kLocal>(a, cus, 0, s);
  if (first <= BIN_G16 && BIN_G16 <= last) launch_bin<BIN_G16, kLocal>(a, cus, 0, s);
  if (first <= BIN_G8 && BIN_G8 <= last) launch_bin<BIN_G8, kLocal>(a, cus, 0, s);
}

hipError_t launch_count_local(const LaunchArgs& a, int cus, int first, int last, int mode, hipStream_t s) {
  if (a.ngraphs != 1 || !a.local.pown) return hipErrorInvalidValue;
  if (mode == kLocalParts) {
    if (!a.local.bounds || !a.local.pe || a.local.parts < 1 || a.local.parts > kMaxParts) return hipErrorInvalidValue;
    launch_local_bins<kLocalParts>(a, cus, first, last, s);
    return hipGetLastError();
  }
  if (!a.local.pedge || !a.local.ecount) return hipErrorInvalidValue;
  if (mode == kLocalFull) {
    if (!a.local.pvid || !a.local.vcount) return hipErrorInvalidValue;
    launch_local_bins<kLocalFull>(a, cus, first, last, s);
  } else if (mode == kLocalEdges) {
    launch_local_bins<kLocalEdges>(a, cus, first, last, s);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_local_prep(const int64_t* rp, int64_t n, int32_t* esrc, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  local_prep_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(rp, n, esrc);
  return hipGetLastError();
}

hipError_t launch_local_parts(const int32_t* ci, const int32_t* esrc, int64_t n, int64_t e0, int64_t e1,
                              const uint32_t* ecount, const int32_t* bounds, int32_t parts, unsigned long long* pe,
                              int cus, hipStream_t s) {
  if (parts < 1 || parts > kMaxParts || !bounds || !pe) return hipErrorInvalidValue;
  if (n <= 0 || e1 <= e0) return hipSuccess;
  const int64_t chunks = (e1 - e0 + kPartsChunk - 1) / kPartsChunk;
  const int64_t pb = min(chunks, (int64_t)cus * 8);
  local_parts_kernel<<<(unsigned)pb, kBlock, 0, s>>>(ci, esrc, bounds, parts, e0, e1, ecount, pe);
  return hipGetLastError();
}

hipError_t launch_plain(const int64_t* rp, const int32_t* ci, int32_t row_begin, int32_t row_end,
                        const uint8_t* own, uint64_t* out, int grid, hipStream_t s, uint64_t* zero_next) {
  plain_kernel<<<grid, kBlock, 0, s>>>(rp, ci, row_begin, row_end, own, out, zero_next);
  return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, int64_t bytes, int grid, hipStream_t s) {
  copy_kernel<<<grid, kBlock, 0, s>>>((const uint4*)src, (uint4*)dst, bytes / 16);
  return hipGetLastError();
}

hipError_t launch_read(const void* src, int64_t bytes, uint32_t* sink, int grid, hipStream_t s) {
  read_kernel<<<grid, kBlock, 0, s>>>((const u32x4*)src, bytes / 16, sink);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// static scratch of the allocation-free count
// ---------------------------------------------------------------------------
// Pieces (256-byte aligned): bin_of [slots] u16, items [slots] u64, blk
// [bins * plan_blocks(slots) + 1] u32, scan partials of blk, counters [64]
// u32, outs [16] u64.
constexpr int64_t nalloc_up(int64_t b) { return (b + 255) & ~(int64_t)255; }
constexpr int64_t kNaBlk = kNumBins * ((kNoAllocSlots + 255) / 256) + 1;
constexpr int64_t kNaBinOf = 0;
constexpr int64_t kNaItems = kNaBinOf + nalloc_up(2 * kNoAllocSlots);
constexpr int64_t kNaBlkOff = kNaItems + nalloc_up(8 * kNoAllocSlots);
constexpr int64_t kNaPart = kNaBlkOff + nalloc_up(4 * kNaBlk);
constexpr int64_t kNaCounters = kNaPart + nalloc_up(8 * ((kNaBlk + kScanChunk - 1) / kScanChunk + 1));
constexpr int64_t kNaOuts = kNaCounters + nalloc_up(4 * 64);
constexpr int64_t kNaBytes = kNaOuts + nalloc_up(8 * 16);

// The scratch itself.  Host code never passes this name to a kernel or a
// copy: it is a host-side stand-in there.  tc_ctx_create takes its device
// address once per context with the context's device current
// (noalloc_scratch_address), and every use goes through that address.  All
// calls on one device share it (overlapping calls race).
static __device__ __align__(256) unsigned char scratch[kNaBytes];

int64_t noalloc_scratch_bytes() { return kNaBytes; }

PlanBufs noalloc_plan_bufs(void* base) {
  char* p = (char*)base;
  return PlanBufs{(uint16_t*)(p + kNaBinOf), (uint64_t*)(p + kNaItems), (uint32_t*)(p + kNaBlkOff),
                  (int64_t*)(p + kNaPart), (uint32_t*)(p + kNaCounters), (uint64_t*)(p + kNaOuts)};
}

hipError_t noalloc_scratch_address(void** ptr) { return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch)); }

int count_blocks_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, count_kernel<BIN_WAVE, false, false>, kBlock, 0) !=
      hipSuccess)
    return 1;
  return n > 0 ? n : 1;
}

int plain_blocks_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, plain_kernel, kBlock, 0) != hipSuccess) return 1;
  return n > 0 ? n : 1;
}

}  // namespace tck
