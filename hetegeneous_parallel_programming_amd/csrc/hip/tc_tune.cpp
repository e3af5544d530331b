// Autotuner of the degree-binned count, and a per-device tune file.
//
// The reference launches one fixed shape (one warp per vertex at the
// maximum-residency grid, multigpu_nvshmem_pushpull.cu:41-51).  Here the
// launch shape is a set of parameters -- the wedge-work thresholds that pick
// each pivot's group width (8/16/32 lanes, a wave64, a workgroup), the pivots
// per group in one light-bin tile, the persistent grid and the LDS layout of
// the pivot sets (hash or tree) -- and this file
// picks them by timing the real workload: coordinate descent over a fixed
// candidate list per parameter, each candidate timed as the minimum of
// `reps` calls after one untimed call (which also re-captures a batch's
// hipGraph).  Every candidate's counts must equal the first candidate's; a
// difference is reported as a device error, never tuned around.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "tc_internal.hpp"

using tcr::fail;

namespace {

struct Workload {
  tc_ctx* ctx;
  tc_graph* g;
  tc_batch* b;
  std::vector<uint64_t> want, got;
  bool have_want = false;
};

// One timed evaluation of the context's current tune: min wall ms over reps.
tc_status evaluate(Workload& w, int32_t reps, double* best_ms) {
  double best = 1e300;
  for (int32_t r = -1; r < reps; ++r) {
    tc_stats st;
    const tc_status s = w.b ? tc_batch_count(w.ctx, w.b, w.got.data(), &st)
                            : tc_count_binned(w.ctx, w.g, w.got.data(), &st);
    if (s != TC_OK) return s;
    if (!w.have_want) {
      w.want = w.got;
      w.have_want = true;
    } else if (w.got != w.want) {
      return fail(TC_EDEVICE, "autotune: counts changed between launch settings");
    }
    if (r >= 0) best = std::min(best, st.total_ms);
  }
  *best_ms = best;
  return TC_OK;
}

bool ordered(const tc_tune& t) {
  return 0 <= t.t_g8 && t.t_g8 <= t.t_g16 && t.t_g16 <= t.t_g32 && t.t_g32 <= t.t_wave;
}

// The device part of a record id: the context's name with blanks and '|'
// replaced, so that a record line splits on blanks.
std::string device_id(const tc_ctx* ctx) {
  std::string d = ctx->name;
  for (char& c : d)
    if (c == ' ' || c == '\t' || c == '|' || c == '\n' || c == '\r') c = '_';
  return d;
}

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

}  // namespace

extern "C" {

tc_status tc_autotune(tc_ctx* ctx, tc_graph* g, tc_batch* b, int32_t reps, double budget_ms,
                      tc_tune* best, tc_autotune_report* report) try {
  if (!ctx || (!g) == (!b)) return fail(TC_EINVAL, "autotune: give exactly one of a graph or a batch");
  if (reps < 1 || reps > 100) return fail(TC_EINVAL, "autotune: reps must be in [1, 100]");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  const double t_start = tcr::now_ms();
  Workload w{ctx, g, b, {}, {}, false};
  const int32_t n_out = b ? tc_batch_graphs(b) : 1;
  w.got.assign((size_t)std::max(n_out, 1), 0);

  const tc_tune start = ctx->tune;
  tc_tune cur = start;
  double cur_ms = 0;
  tc_status s = evaluate(w, reps, &cur_ms);
  if (s != TC_OK) return s;
  const double start_ms = cur_ms;
  int32_t trials = 1;

  const int cus = ctx->cus;
  struct Param { int32_t tc_tune::*field; std::vector<int32_t> cand; };
  const std::vector<Param> params = {
      {&tc_tune::t_wave, {4096, 8192, 16384, 32768, 65536}},
      {&tc_tune::t_g32, {512, 1024, 2048, 4096, 8192}},
      {&tc_tune::t_g16, {128, 256, 512, 1024, 2048}},
      {&tc_tune::t_g8, {0, 64, 128, 256, 512}},
      {&tc_tune::rounds, {1, 2, 4, 8, 16}},
      {&tc_tune::grid, {0, 2 * cus, 4 * cus, 8 * cus}},
      // pivot_set 1 (a tree for every set) is not searched: 2 (trees for
      // sub-wave groups only) matched or beat it on every workload measured,
      // and 1 roughly doubles the time of rows with long lists, which a batch
      // tuned on small graphs may later meet.
      {&tc_tune::pivot_set, {0, 2}},
  };
  bool out_of_time = false;
  for (int sweep = 0; sweep < 2 && !out_of_time; ++sweep) {
    bool improved = false;
    for (const Param& p : params) {
      for (int32_t v : p.cand) {
        if (tcr::now_ms() - t_start > budget_ms) { out_of_time = true; break; }
        if (cur.*(p.field) == v) continue;
        tc_tune t = cur;
        t.*(p.field) = v;
        if (!ordered(t)) continue;
        ctx->tune = t;
        double ms = 0;
        s = evaluate(w, reps, &ms);
        ++trials;
        if (s != TC_OK) { ctx->tune = start; return s; }
        if (ms < cur_ms * 0.98) {  // keep a change only for a clear gain
          cur = t;
          cur_ms = ms;
          improved = true;
        }
      }
      if (out_of_time) break;
    }
    if (!improved) break;
  }
  ctx->tune = cur;
  if (best) *best = cur;
  if (report) {
    report->trials = trials;
    report->start_ms = start_ms;
    report->best_ms = cur_ms;
    report->elapsed_ms = tcr::now_ms() - t_start;
  }
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

tc_status tc_tune_save(tc_ctx* ctx, const char* path, const char* key) try {
  if (!ctx || !path || !key) return fail(TC_EINVAL, "null argument");
  if (std::strpbrk(key, " \t\r\n|")) return fail(TC_EINVAL, "tune key must not hold blanks or '|'");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  const std::string id = device_id(ctx) + "|" + key;
  std::vector<std::string> keep;
  {
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
      const std::string t = trim(line);
      if (t.empty()) continue;
      if (t.compare(0, id.size() + 1, id + " ") == 0) continue;
      keep.push_back(t);
    }
  }
  const tc_tune& t = ctx->tune;
  std::ostringstream rec;
  rec << id << " " << t.t_g8 << " " << t.t_g16 << " " << t.t_g32 << " " << t.t_wave << " "
      << t.rounds << " " << t.grid << " " << t.pivot_set << " " << t.hash_kicks;
  keep.push_back(rec.str());
  const std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream out(tmp, std::ios::trunc);
    if (!out) return fail(TC_EIO, std::string("cannot write ") + tmp);
    for (const std::string& l : keep) out << l << "\n";
    if (!out) return fail(TC_EIO, std::string("cannot write ") + tmp);
  }
  if (std::rename(tmp.c_str(), path) != 0) return fail(TC_EIO, std::string("cannot replace ") + path);
  return TC_OK;
} catch (...) {
  return tcr::caught();
}

tc_status tc_tune_load(tc_ctx* ctx, const char* path, const char* key) try {
  if (!ctx || !path || !key) return fail(TC_EINVAL, "null argument");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  const std::string id = device_id(ctx) + "|" + key;
  std::ifstream in(path);
  if (!in) return fail(TC_EIO, std::string("cannot read ") + path);
  std::string line;
  while (std::getline(in, line)) {
    const std::string t = trim(line);
    if (t.compare(0, id.size() + 1, id + " ") != 0) continue;
    std::istringstream f(t.substr(id.size() + 1));
    tc_tune v = tcr::default_tune();
    if (!(f >> v.t_g8 >> v.t_g16 >> v.t_g32 >> v.t_wave >> v.rounds >> v.grid >> v.pivot_set >> v.hash_kicks))
      return fail(TC_EIO, std::string("malformed tune record in ") + path);
    tc_status s = tcr::check_tune(v);
    if (s != TC_OK) return s;
    ctx->tune = v;
    return TC_OK;
  }
  return fail(TC_EIO, std::string("no tune record for ") + id + " in " + path);
} catch (...) {
  return tcr::caught();
}

}  // extern "C"
