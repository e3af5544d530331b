// Triangle-counting kernels for gfx950 (CDNA4, wave64).
//
// The counted quantity is the reference's bs_warp_vertex loop nest
// (bs_warp_vertex.cuh:13-22): sum over pivots u, over v in N+(u), of
// |N+(u) ∩ N+(v)| on the degree-ordered DAG, i.e. every triangle once.
//
// Work unit.  A group of G lanes (G = 8, 16, 32 or 64) owns one pivot u.
// It stages N+(u) in LDS, then walks the edges (u, v) in batches of G: lane
// j of the batch reads rowptr[v], rowptr[v+1] for its edge, a group scan of
// the d+(v) gives each edge's slice of a flat index space, and the group's
// lanes then stride that space with coalesced int32 reads of the N+(v)
// lists.  Each read neighbour w is binary-searched in the LDS copy of N+(u)
// (the search of the reference's warp_intersect, successful_push_pull/
// graph_gpu.h:228-252, with the pivot list as the searched side), and every
// element of N+(v) is read exactly once per edge.  Several lookups per lane
// (TC_IN_FLIGHT) are kept in flight so that the HBM/L2 reads and the LDS
// search chains overlap.
//
// Scheduling.  A planning pass bins pivots by d+(u) and by the wedge work
// T(u) = sum of d+(v) over v in N+(u); light pivots go to 8/16/32-lane
// groups, medium ones to a wave, heavy ones to a whole workgroup whose four
// waves share one LDS copy of N+(u).  A persistent grid then pulls tiles
// (bin, run of pivots) from a device queue, heaviest bins first.  Counts are
// u64 sums, so the result does not depend on tile order.
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

// j-th smallest id of a pivot list: from the LDS tree, or from the sorted HBM copy.
template <bool kTree>
__device__ __forceinline__ int list_at(const int32_t* A, int j, int h) {
  return kTree ? A[tree_pos(j, h)] : A[j];
}

#ifndef TC_IN_FLIGHT
#define TC_IN_FLIGHT 6
#endif
constexpr int kInFlight = TC_IN_FLIGHT;  // lookups per lane per iteration

// Counts, for the edges j = jfirst, jfirst + jstride, ... of pivot u (list
// of da ids: an LDS tree, or the sorted list in HBM), the elements of N+(v)
// found in that list.  sEnd/sBase are this group's G-entry LDS scratch.
// Returns this lane's share of the hits.
//
// Flat index t of a batch maps to edge k (sEnd[k-1] <= t < sEnd[k]) and to
// colidx[sBase[k] + t]; each lane keeps its current edge's end and base in
// registers and searches sEnd only when it steps past that edge.
template <int G, bool kTree>
__device__ __forceinline__ uint32_t pivot_edges(const int64_t* __restrict__ rp,
                                                const int32_t* __restrict__ ci, const int32_t* A,
                                                int da, int jfirst, int jstride, int32_t* sEnd,
                                                int64_t* sBase, int gl) {
  uint32_t hits = 0;
  const int h = tree_levels(da);
  for (int j0 = jfirst; j0 < da; j0 += jstride) {
    const int j = j0 + gl;
    int c = 0;
    int64_t off = 0;
    if (j < da) {
      const int v = list_at<kTree>(A, j, h);
      const int64_t b = rp[v];
      c = (int)(rp[v + 1] - b);
      off = b;
    }
    int total;
    const int excl = group_excl_scan<G>(c, gl, total);
    sEnd[gl] = excl + c;
    sBase[gl] = off - excl;
    wave_sync();
    int cur_end = 0;        // t < cur_end  <=>  t lies in the lane's current edge
    int64_t cur_base = 0;
    for (int t0 = 0; t0 < total; t0 += G * kInFlight) {
      int w[kInFlight];
#pragma unroll
      for (int q = 0; q < kInFlight; ++q) {
        const int t = t0 + q * G + gl;
        w[q] = INT_MIN;  // ids are >= 0: never found
        if (t < total) {
          if (t >= cur_end) {
            const int k = find_edge<G>(sEnd, t);
            cur_end = sEnd[k];
            cur_base = sBase[k];
          }
          w[q] = ci[cur_base + t];
        }
      }
      if (kTree) {
        // h descent steps; then the lower bound is k with its trailing one
        // bits and the following zero shifted out (0 = past the end -> -1).
        // The node index is carried as a byte offset (4k): one compare, one
        // select and one shift-add per step.
        const char* Ab = (const char*)A;
        int kb[kInFlight];
#pragma unroll
        for (int q = 0; q < kInFlight; ++q) kb[q] = 4;
        for (int s = 0; s < h; ++s) {
          int val[kInFlight];
#pragma unroll
          for (int q = 0; q < kInFlight; ++q) val[q] = *(const int32_t*)(Ab + kb[q]);
#pragma unroll
          for (int q = 0; q < kInFlight; ++q) kb[q] = 2 * kb[q] + (val[q] < w[q] ? 4 : 0);
        }
#pragma unroll
        for (int q = 0; q < kInFlight; ++q) {
          const int k = kb[q] >> 2;
          hits += (A[k >> __builtin_ffs(~k)] == w[q]) ? 1u : 0u;
        }
      } else {
        int lo[kInFlight];
#pragma unroll
        for (int q = 0; q < kInFlight; ++q) lo[q] = 0;
        for (int len = da; len > 1;) {
          const int half = len >> 1;
#pragma unroll
          for (int q = 0; q < kInFlight; ++q) lo[q] = (A[lo[q] + half] <= w[q]) ? lo[q] + half : lo[q];
          len -= half;
        }
#pragma unroll
        for (int q = 0; q < kInFlight; ++q) hits += (A[lo[q]] == w[q]) ? 1u : 0u;
      }
    }
    wave_sync();
  }
  return hits;
}

// Lays N+(u) out as an LDS tree with `n` cooperating lanes (coalesced reads).
__device__ __forceinline__ void stage_pivot(int32_t* slot, const int32_t* __restrict__ src, int da,
                                            int lane, int n) {
  const int h = tree_levels(da);
  const int nodes = (1 << h) - 1;
  for (int x = lane; x < nodes; x += n) slot[tree_pos(x, h)] = x < da ? src[x] : INT_MAX;
  if (lane == 0) slot[0] = -1;
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
__device__ __forceinline__ int find_graph(const GraphDesc* descs, int ngraphs, int64_t p) {
  int lo = 0, hi = ngraphs - 1;  // largest g with plan_base <= p
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].plan_base <= p) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int choose_bin(const int64_t* rp, const int32_t* ci, int64_t rb,
                                          int64_t da, const Tune& t) {
  if (da < 2) return kNoBin;
  if (da > kCapBlock - 1) return BIN_GLOBAL;
  if (da > kCapWave - 1) return BIN_BLOCK;
  int64_t work = 0;
  for (int64_t e = rb; e < rb + da; ++e) {
    const int v = ci[e];
    work += rp[v + 1] - rp[v];
    if (work > t.t_wave) return BIN_BLOCK;
  }
  if (work == 0) return kNoBin;
  if (da > kCapG32 - 1 || work > t.t_g32) return BIN_WAVE;
  if (da > kCapG16 - 1 || work > t.t_g16) return BIN_G32;
  if (da > kCapG8 - 1 || work > t.t_g8) return BIN_G16;
  return BIN_G8;
}

// Pass 1: bin of every pivot and, per planning block, the count per bin.
__global__ __launch_bounds__(kBlock) void plan_bin_kernel(LaunchArgs a) {
  __shared__ uint32_t s_cnt[kNumBins];
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (threadIdx.x < kNumBins) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  if (p < a.P) {
    const int g = find_graph(a.descs, a.ngraphs, p);
    const GraphDesc d = a.descs[g];
    const int64_t u = d.row_begin + (p - d.plan_base);
    const int64_t* rp = a.rp + d.rp_off;
    const int32_t* ci = a.ci + d.ci_off;
    const int64_t rb = rp[u];
    const int bin = choose_bin(rp, ci, rb, rp[u + 1] - rb, a.tune);
    a.bufs.bin_of[p] = (uint8_t)bin;
    if (bin != kNoBin) atomicAdd(&s_cnt[bin], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kNumBins) a.bufs.blk[(int64_t)blockIdx.x * kBlkStride + threadIdx.x] = s_cnt[threadIdx.x];
}

// Pass 2 (one workgroup of 1024): per-block counts -> item offsets, ordered
// by (bin, block); bin totals into the counters.
constexpr int kScanThreads = 1024;

__global__ __launch_bounds__(kScanThreads) void plan_scan_kernel(LaunchArgs a, int64_t nblk) {
  __shared__ uint32_t s_wave[kScanThreads / 64];
  __shared__ uint32_t s_bin_total[kNumBins];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t chunk = (nblk + kScanThreads - 1) / kScanThreads;
  const int64_t b0 = min((int64_t)tid * chunk, nblk), b1 = min(b0 + chunk, nblk);
  uint32_t* blk = a.bufs.blk;
  uint32_t excl[kNumBins];
  for (int b = 0; b < kNumBins; ++b) {
    uint32_t local = 0;
    for (int64_t k = b0; k < b1; ++k) local += blk[k * kBlkStride + b];
    // inclusive scan over the 1024 threads
    uint32_t incl = local;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(incl, d, 64);
      if (lane >= d) incl += y;
    }
    if (lane == 63) s_wave[wid] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) {
      if (w < wid) before += s_wave[w];
      all += s_wave[w];
    }
    excl[b] = before + incl - local;
    if (tid == 0) s_bin_total[b] = all;
    __syncthreads();
  }
  uint32_t base = 0;
  for (int b = 0; b < kNumBins; ++b) {
    uint32_t run = base + excl[b];
    for (int64_t k = b0; k < b1; ++k) {
      const uint32_t c = blk[k * kBlkStride + b];
      blk[k * kBlkStride + b] = run;
      run += c;
    }
    base += s_bin_total[b];
  }
  if (tid < kNumBins) a.bufs.counters[kCntBinCount + tid] = s_bin_total[tid];
}

// Pass 3: stable scatter of (graph, row) items into their bin segment.
__global__ __launch_bounds__(kBlock) void plan_scatter_kernel(LaunchArgs a) {
  __shared__ uint32_t s_wc[kWaves][kNumBins];
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int bin = (p < a.P) ? (int)a.bufs.bin_of[p] : (int)kNoBin;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t rank = 0;
#pragma unroll
  for (int b = 0; b < kNumBins; ++b) {
    const uint64_t m = __ballot(bin == b);
    if (bin == b) rank = (uint32_t)__popcll(m & lt);
    if (lane == 0) s_wc[wid][b] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (bin != kNoBin) {
    uint32_t off = a.bufs.blk[(int64_t)blockIdx.x * kBlkStride + bin] + rank;
    for (int w = 0; w < wid; ++w) off += s_wc[w][bin];
    const int g = find_graph(a.descs, a.ngraphs, p);
    const GraphDesc d = a.descs[g];
    const int64_t u = d.row_begin + (p - d.plan_base);
    a.bufs.items[off] = ((uint64_t)(uint32_t)g << 32) | (uint32_t)u;
  }
}

// ---------------------------------------------------------------------------
// persistent count
// ---------------------------------------------------------------------------
struct __align__(16) CountLds {
  int32_t A[kCapBlock];          // pivot-tree slots (layout depends on the tile's bin)
  int32_t End[kBlock];           // per-lane edge end in the batch's flat index
  int64_t Base[kBlock];          // per-lane colidx base of that edge
  uint64_t red[kWaves];
  int64_t tab_cnt[1];
  int64_t cur_first, cur_seg;
};
static_assert(kCapWave * kWaves <= kCapBlock, "wave slots exceed LDS region");
static_assert(kCapG32 * (kBlock / 32) <= kCapBlock, "g32 slots exceed LDS region");
static_assert(kCapG16 * (kBlock / 16) <= kCapBlock, "g16 slots exceed LDS region");
static_assert(kCapG8 * (kBlock / 8) <= kCapBlock, "g8 slots exceed LDS region");

// Per-lane running count; in batch mode it is flushed to outs[g] whenever
// the lane moves to a pivot of another graph.
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

template <int G, int CAP>
__device__ __forceinline__ void group_tile(const LaunchArgs& a, const uint64_t* items, int64_t first,
                                           int64_t last, CountLds& L, bool batch, Acc& acc) {
  constexpr int NGRP = kBlock / G;
  const int grp = threadIdx.x / G, gl = threadIdx.x % G;
  int32_t* slot = L.A + grp * CAP;
  int32_t* sEnd = L.End + grp * G;
  int64_t* sBase = L.Base + grp * G;
  for (int64_t i = first + grp; i < last; i += NGRP) {
    const uint64_t item = items[i];
    const int g = (int)(item >> 32);
    const int u = (int)(uint32_t)item;
    const GraphDesc d = a.descs[g];
    const int64_t* rp = a.rp + d.rp_off;
    const int32_t* ci = a.ci + d.ci_off;
    const int64_t rb = rp[u];
    const int da = (int)(rp[u + 1] - rb);
    stage_pivot(slot, ci + rb, da, gl, G);
    wave_sync();
    const uint32_t h = pivot_edges<G, true>(rp, ci, slot, da, 0, G, sEnd, sBase, gl);
    acc.to_graph(g, batch, a.bufs.outs);
    acc.sum += h;
    wave_sync();
  }
}

template <bool kLdsPivot>
__device__ __forceinline__ void block_pivot(const LaunchArgs& a, uint64_t item, CountLds& L, bool batch,
                                            Acc& acc) {
  const int g = (int)(item >> 32);
  const int u = (int)(uint32_t)item;
  const GraphDesc d = a.descs[g];
  const int64_t* rp = a.rp + d.rp_off;
  const int32_t* ci = a.ci + d.ci_off;
  const int64_t rb = rp[u];
  const int da = (int)(rp[u + 1] - rb);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t h;
  if (kLdsPivot) {
    stage_pivot(L.A, ci + rb, da, threadIdx.x, kBlock);
    __syncthreads();
    h = pivot_edges<64, true>(rp, ci, L.A, da, wid * 64, kBlock, L.End + wid * 64, L.Base + wid * 64, lane);
  } else {
    h = pivot_edges<64, false>(rp, ci, ci + rb, da, wid * 64, kBlock, L.End + wid * 64, L.Base + wid * 64,
                               lane);
  }
  acc.to_graph(g, batch, a.bufs.outs);
  acc.sum += h;
}

// Minimum resident waves per SIMD the compiler must allow for (caps VGPRs:
// 8 waves = 64 VGPRs).  Latency hiding dominates this kernel's speed.
#ifndef TC_COUNT_WAVES_PER_EU
#define TC_COUNT_WAVES_PER_EU 8
#endif

// Pivots per tile of each bin (rounds = pivots per group in light bins).
__device__ __forceinline__ int tile_pivots(int bin, int rounds) {
  switch (bin) {
    case BIN_GLOBAL:
    case BIN_BLOCK: return 1;
    case BIN_WAVE: return kWaves;
    case BIN_G32: return (kBlock / 32) * rounds;
    case BIN_G16: return (kBlock / 16) * rounds;
    default: return (kBlock / 8) * rounds;
  }
}

// One persistent kernel per bin, specialised so that each keeps only its own
// working set in registers.  Blocks pull tiles from the bin's queue head.
template <int BIN>
__global__ __launch_bounds__(kBlock, TC_COUNT_WAVES_PER_EU) void count_kernel(LaunchArgs a) {
  __shared__ CountLds L;
  const bool batch = a.ngraphs > 1;
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
    const int64_t t = L.cur_first;
    if (t >= ntiles) break;
    const int64_t first = t * piv;
    const int64_t last = min(first + (int64_t)piv, cnt);
    if (BIN == BIN_GLOBAL) block_pivot<false>(a, items[first], L, batch, acc);
    else if (BIN == BIN_BLOCK) block_pivot<true>(a, items[first], L, batch, acc);
    else if (BIN == BIN_WAVE) group_tile<64, kCapWave>(a, items, first, last, L, batch, acc);
    else if (BIN == BIN_G32) group_tile<32, kCapG32>(a, items, first, last, L, batch, acc);
    else if (BIN == BIN_G16) group_tile<16, kCapG16>(a, items, first, last, L, batch, acc);
    else group_tile<8, kCapG8>(a, items, first, last, L, batch, acc);
    if (batch) {
      // Items are ordered by graph, so a tile almost always belongs to one
      // graph: then one block-reduced atomic; otherwise one per lane.
      const int tg = (int)(items[first] >> 32);
      if (!__syncthreads_or(acc.sum != 0 && acc.g != tg)) {
        const uint64_t tot = block_sum(acc.sum, L.red);
        if (threadIdx.x == 0 && tot) atomicAdd(&a.bufs.outs[tg], tot);
      } else if (acc.sum) {
        atomicAdd(&a.bufs.outs[acc.g], acc.sum);
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
// plain kernel: one wave64 per pivot, grid-stride (bs_warp_vertex.cuh:2-35)
// ---------------------------------------------------------------------------
struct __align__(16) PlainLds {
  int32_t A[kWaves][kCapWave];
  int32_t End[kBlock];
  int64_t Base[kBlock];
  uint64_t red[kWaves];
};

__global__ __launch_bounds__(kBlock, TC_COUNT_WAVES_PER_EU) void plain_kernel(const int64_t* __restrict__ rp,
                                                       const int32_t* __restrict__ ci,
                                                       int32_t row_begin, int32_t row_end,
                                                       uint64_t* out) {
  __shared__ PlainLds L;
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * kWaves;
  uint64_t acc = 0;
  for (int64_t u = row_begin + (int64_t)blockIdx.x * kWaves + wid; u < row_end; u += nw) {
    const int64_t rb = rp[u];
    const int da = (int)(rp[u + 1] - rb);
    if (da < 2) continue;
    if (da <= kCapWave - 1) {
      stage_pivot(L.A[wid], ci + rb, da, lane, 64);
      wave_sync();
      acc += pivot_edges<64, true>(rp, ci, L.A[wid], da, 0, 64, L.End + wid * 64, L.Base + wid * 64, lane);
      wave_sync();
    } else {
      acc += pivot_edges<64, false>(rp, ci, ci + rb, da, 0, 64, L.End + wid * 64, L.Base + wid * 64, lane);
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

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_zero(void* p, int64_t words32, hipStream_t s) {
  if (words32 <= 0) return hipSuccess;
  int64_t blocks = (words32 + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  zero_kernel<<<(unsigned)blocks, kBlock, 0, s>>>((uint32_t*)p, words32);
  return hipGetLastError();
}

hipError_t launch_plan(const LaunchArgs& a, hipStream_t s) {
  if (a.P <= 0) return hipSuccess;
  const int64_t blocks = (a.P + kBlock - 1) / kBlock;
  plan_bin_kernel<<<(unsigned)blocks, kBlock, 0, s>>>(a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  plan_scan_kernel<<<1, kScanThreads, 0, s>>>(a, blocks);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  plan_scatter_kernel<<<(unsigned)blocks, kBlock, 0, s>>>(a);
  return hipGetLastError();
}

template <int BIN>
static int bin_blocks_per_cu() {
  static const int n = [] {
    int k = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, count_kernel<BIN>, kBlock, 0) != hipSuccess) k = 1;
    return k > 0 ? k : 1;
  }();
  return n;
}

template <int BIN>
static void launch_bin(const LaunchArgs& a, int cus, int grid, hipStream_t s) {
  const int g = grid > 0 ? grid : bin_blocks_per_cu<BIN>() * cus;
  count_kernel<BIN><<<g, kBlock, 0, s>>>(a);
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

hipError_t launch_plain(const int64_t* rp, const int32_t* ci, int32_t row_begin, int32_t row_end,
                        uint64_t* out, int grid, hipStream_t s) {
  plain_kernel<<<grid, kBlock, 0, s>>>(rp, ci, row_begin, row_end, out);
  return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, int64_t bytes, int grid, hipStream_t s) {
  copy_kernel<<<grid, kBlock, 0, s>>>((const uint4*)src, (uint4*)dst, bytes / 16);
  return hipGetLastError();
}

int count_blocks_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, count_kernel<BIN_WAVE>, kBlock, 0) != hipSuccess)
    return 1;
  return n > 0 ? n : 1;
}

int plain_blocks_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, plain_kernel, kBlock, 0) != hipSuccess) return 1;
  return n > 0 ? n : 1;
}

}  // namespace tck
