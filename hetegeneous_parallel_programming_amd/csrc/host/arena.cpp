// First-fit extent bookkeeping with coalescing of neighbouring free extents.
// Used by the serving loop to place request CSRs inside one HBM allocation.
#include <map>
#include <mutex>

#include "host_common.hpp"

struct tch_arena {
  uint64_t capacity = 0, align = 256;
  std::map<uint64_t, uint64_t> free_;  // offset -> size, disjoint, never adjacent
  std::map<uint64_t, uint64_t> live_;  // offset -> size
  uint64_t used = 0, peak = 0;
  int64_t allocs = 0, failed = 0;
  mutable std::mutex mu;
};

using namespace tch;

extern "C" {

tc_status tch_arena_create(uint64_t capacity, uint64_t align, tch_arena** out) try {
  if (!out) return fail(TC_EINVAL, "null out");
  if (align == 0 || (align & (align - 1))) return fail(TC_EINVAL, "align must be a power of two");
  std::unique_ptr<tch_arena> a(new tch_arena);
  a->align = align;
  a->capacity = capacity / align * align;
  if (a->capacity) a->free_[0] = a->capacity;
  *out = a.release();
  return TC_OK;
} catch (...) {
  return tch::caught();
}

void tch_arena_destroy(tch_arena* a) { delete a; }

tc_status tch_arena_alloc(tch_arena* a, uint64_t bytes, uint64_t* offset) try {
  if (!a || !offset) return fail(TC_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(a->mu);
  const uint64_t need = bytes == 0 ? a->align : (bytes + a->align - 1) / a->align * a->align;
  for (auto it = a->free_.begin(); it != a->free_.end(); ++it) {
    if (it->second < need) continue;
    const uint64_t off = it->first, rest = it->second - need;
    a->free_.erase(it);
    if (rest) a->free_[off + need] = rest;
    a->live_[off] = need;
    a->used += need;
    if (a->used > a->peak) a->peak = a->used;
    ++a->allocs;
    *offset = off;
    return TC_OK;
  }
  ++a->failed;
  return fail(TC_ENOMEM, "arena has no free extent of " + std::to_string(need) + " bytes");
} catch (...) {
  return tch::caught();
}

tc_status tch_arena_free(tch_arena* a, uint64_t offset) try {
  if (!a) return fail(TC_EINVAL, "null arena");
  std::lock_guard<std::mutex> lk(a->mu);
  auto it = a->live_.find(offset);
  if (it == a->live_.end()) return fail(TC_EINVAL, "offset is not a live extent");
  uint64_t off = it->first, size = it->second;
  a->live_.erase(it);
  a->used -= size;
  auto nxt = a->free_.lower_bound(off);
  if (nxt != a->free_.end() && off + size == nxt->first) {  // merge with the following extent
    size += nxt->second;
    nxt = a->free_.erase(nxt);
  }
  if (nxt != a->free_.begin()) {  // merge with the preceding extent
    auto prv = std::prev(nxt);
    if (prv->first + prv->second == off) {
      prv->second += size;
      return TC_OK;
    }
  }
  a->free_[off] = size;
  return TC_OK;
} catch (...) {
  return tch::caught();
}

tc_status tch_arena_stats_get(const tch_arena* a, tch_arena_stats* out) try {
  if (!a || !out) return fail(TC_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(a->mu);
  out->capacity = a->capacity;
  out->used = a->used;
  out->peak_used = a->peak;
  uint64_t lf = 0;
  for (auto& kv : a->free_) lf = kv.second > lf ? kv.second : lf;
  out->largest_free = lf;
  out->live_extents = (int64_t)a->live_.size();
  out->free_extents = (int64_t)a->free_.size();
  out->allocs = a->allocs;
  out->failed_allocs = a->failed;
  return TC_OK;
} catch (...) {
  return tch::caught();
}

}  // extern "C"
