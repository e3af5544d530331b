// Internal helpers of the host graph substrate: owned CSR storage, per-thread
// error text, counter-based random draws and OpenMP prefix sums.
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <omp.h>

#include "tc_host.h"

namespace tch {

// Owned CSR.  Arrays are allocated uninitialised (new T[]), so building a
// 2 GB colidx does not pay for a zero fill it would overwrite anyway.
struct Csr {
  int32_t n = 0;
  int64_t nnz = 0;
  std::unique_ptr<int64_t[]> rowptr;
  std::unique_ptr<int32_t[]> colidx;
  int32_t max_degree = 0;

  void allocate(int32_t n_, int64_t nnz_) {
    n = n_;
    nnz = nnz_;
    rowptr.reset(new int64_t[(size_t)n_ + 1]);
    colidx.reset(new int32_t[nnz_ > 0 ? (size_t)nnz_ : 1]);
  }
  tc_csr_view view() const { return tc_csr_view{rowptr.get(), colidx.get(), n, nnz}; }
  void compute_max_degree();
};

// Sets this thread's error text and returns `st`.
tc_status fail(tc_status st, const std::string& msg);
void clear_error();

// Called from a catch block: turns the exception in flight into a status, so
// that no C++ exception crosses the C ABI (a failed host allocation, e.g. for
// sizes read from a file, is TC_ENOMEM).
tc_status caught() noexcept;

// Size of a file in bytes, or -1 when it cannot be examined.
int64_t file_bytes(const std::string& path);
// Reads exactly `bytes` bytes (the whole file) into dst / writes them out.
bool read_exact(const std::string& path, void* dst, size_t bytes);
bool write_exact(const std::string& path, const void* src, size_t bytes);

// splitmix64 finaliser: a bijective 64-bit mixer.
static inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// Draw number `b` of stream `a` under `seed`: independent of evaluation order.
static inline uint64_t draw(uint64_t seed, uint64_t a, uint64_t b) {
  return mix64(mix64(seed + 0x632BE59BD9B4E019ull * (a + 1)) + b);
}
// Uniform double in [0, 1) from the top 53 bits.
static inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// Exclusive prefix sum of counts[0..n) into out[0..n] (out[n] = total).
template <typename TIn, typename TOut>
void exclusive_scan(const TIn* counts, int64_t n, TOut* out) {
  int nt = omp_get_max_threads();
  if (n < (int64_t)1 << 16 || nt == 1) {
    TOut s = 0;
    for (int64_t i = 0; i < n; ++i) { out[i] = s; s += (TOut)counts[i]; }
    out[n] = s;
    return;
  }
  std::vector<TOut> part(nt + 1, 0);
#pragma omp parallel num_threads(nt)
  {
    // The team may be smaller than nt (e.g. when nested): use its real size.
    int t = omp_get_thread_num();
    int T = omp_get_num_threads();
    int64_t lo = n * t / T, hi = n * (t + 1) / T;
    TOut s = 0;
    for (int64_t i = lo; i < hi; ++i) s += (TOut)counts[i];
    part[t + 1] = s;
#pragma omp barrier
#pragma omp single
    for (int i = 1; i <= T; ++i) part[i] += part[i - 1];
    TOut base = part[t];
    for (int64_t i = lo; i < hi; ++i) { out[i] = base; base += (TOut)counts[i]; }
  }
  out[n] = out[n - 1] + (TOut)counts[n - 1];
}

// Builds the symmetric deduplicated sorted CSR of the pairs; entries < 0 skip.
// `threads` <= 1 runs single-threaded (used per graph in batch generation).
tc_status build_symmetric(int32_t n, const int32_t* src, const int32_t* dst, int64_t m,
                          int threads, Csr& out);

// Orientation into a fresh Csr (graph.cc:252-298).
tc_status orient(const tc_csr_view& sym, int threads, Csr& dag);

// Generators writing pairs (entries set to -1 are skipped downstream).
void gen_er_pairs(int32_t n, int64_t m, uint64_t seed, int32_t* src, int32_t* dst, int threads);
void gen_rmat_pairs(int32_t n, int32_t scale, int64_t m, double a, double b, double c,
                    uint64_t seed, int32_t* src, int32_t* dst, int threads);
int32_t ceil_log2(int64_t n);

}  // namespace tch

struct tch_csr {
  tch::Csr g;
};
