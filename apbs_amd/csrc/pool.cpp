// pool.cpp -- see pool.h.
#include "pool.h"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace pmgpool {
namespace {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fprintf(stderr, "apbs_pmg pool: %s failed: %s\n", what, hipGetErrorString(e));
    abort();
  }
}

struct State {
  std::mutex m;
  std::multimap<size_t, void*> freeb;      // cached device blocks by size
  std::unordered_map<void*, size_t> live;  // device blocks handed out
  size_t inuse = 0, peak = 0, cached = 0;
  std::vector<double*> pinned;
  std::vector<hipStream_t> streams;
};

// never destroyed: callers may free during static destruction
State& st() {
  static State* s = new State();
  return *s;
}

size_t round_size(size_t n) {
  if (n == 0) n = 1;
  if (n < (1u << 20)) return (n + 511) & ~size_t(511);
  return (n + (1u << 20) - 1) & ~size_t((1u << 20) - 1);
}

}  // namespace

void* dev_alloc(size_t bytes) {
  const size_t sz = round_size(bytes);
  State& S = st();
  {
    std::lock_guard<std::mutex> g(S.m);
    // reuse a cached block of at most 1.25x the size
    auto it = S.freeb.lower_bound(sz);
    if (it != S.freeb.end() && it->first <= sz + sz / 4) {
      void* p = it->second;
      const size_t have = it->first;
      S.freeb.erase(it);
      S.cached -= have;
      S.live[p] = have;
      S.inuse += have;
      if (S.inuse > S.peak) S.peak = S.inuse;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim();
    check(hipMalloc(&p, sz), "hipMalloc");
  }
  std::lock_guard<std::mutex> g(S.m);
  S.live[p] = sz;
  S.inuse += sz;
  if (S.inuse > S.peak) S.peak = S.inuse;
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  State& S = st();
  std::lock_guard<std::mutex> g(S.m);
  auto it = S.live.find(p);
  if (it == S.live.end()) {
    fprintf(stderr, "apbs_pmg pool: free of an unknown block %p\n", p);
    abort();
  }
  const size_t sz = it->second;
  S.live.erase(it);
  S.inuse -= sz;
  S.freeb.insert({sz, p});
  S.cached += sz;
}

void trim() {
  State& S = st();
  std::vector<void*> drop;
  {
    std::lock_guard<std::mutex> g(S.m);
    for (auto& kv : S.freeb) drop.push_back(kv.second);
    S.freeb.clear();
    S.cached = 0;
  }
  for (void* p : drop) check(hipFree(p), "hipFree");
}

double* pinned_scalars() {
  State& S = st();
  {
    std::lock_guard<std::mutex> g(S.m);
    if (!S.pinned.empty()) {
      double* p = S.pinned.back();
      S.pinned.pop_back();
      return p;
    }
  }
  double* p = nullptr;
  check(hipHostMalloc((void**)&p, 32 * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
  return p;
}

void pinned_release(double* p) {
  if (!p) return;
  State& S = st();
  std::lock_guard<std::mutex> g(S.m);
  S.pinned.push_back(p);
}

hipStream_t stream_acquire() {
  State& S = st();
  {
    std::lock_guard<std::mutex> g(S.m);
    if (!S.streams.empty()) {
      hipStream_t s = S.streams.back();
      S.streams.pop_back();
      return s;
    }
  }
  hipStream_t s;
  check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  return s;
}

void stream_release(hipStream_t s) {
  State& S = st();
  std::lock_guard<std::mutex> g(S.m);
  S.streams.push_back(s);
}

void stats(size_t* inuse, size_t* peak, size_t* cached) {
  State& S = st();
  std::lock_guard<std::mutex> g(S.m);
  if (inuse) *inuse = S.inuse;
  if (peak) *peak = S.peak;
  if (cached) *cached = S.cached;
}

void reset_peak() {
  State& S = st();
  std::lock_guard<std::mutex> g(S.m);
  S.peak = S.inuse;
}

}  // namespace pmgpool
