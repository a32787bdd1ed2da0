// host_util.hpp -- small host-side helpers shared by the C++ sources.
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace emst {

// Number of worker threads for host-side parallel loops.
int host_threads();
void set_host_threads(int t);

// Calls fn(begin, end, worker) on contiguous slices of [0, n).  Slices are
// a pure function of (n, workers), so results written per index do not
// depend on scheduling.
template <class F>
void parallel_for(int64_t n, F&& fn, int workers = 0)
{
  if (workers <= 0) workers = host_threads();
  if (n <= 0) return;
  if (workers > n) workers = (int)n;
  if (workers <= 1) { fn((int64_t)0, n, 0); return; }
  std::vector<std::thread> th;
  th.reserve(workers);
  for (int t = 0; t < workers; t++) {
    const int64_t b = n * t / workers, e = n * (t + 1) / workers;
    th.emplace_back([&fn, b, e, t] { fn(b, e, t); });
  }
  for (auto& x : th) x.join();
}

// SplitMix64 finaliser: a bijective 64-bit mix (Steele, Lea, Flood 2014).
static inline uint64_t mix64(uint64_t z)
{
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Counter-based stream: value i of stream (seed, tag).
static inline uint64_t stream64(uint64_t seed, uint64_t tag, uint64_t i)
{
  return mix64(mix64(seed ^ (tag * 0xd1b54a32d192ed03ull)) + i);
}

// Unbiased integer in [0, range) from a counter-based source (Lemire 2019,
// "Fast random integer generation in an interval"); rejected draws are
// replaced by re-mixing, so the result stays a pure function of x.
static inline uint64_t bounded64(uint64_t x, uint64_t range)
{
  unsigned __int128 m = (unsigned __int128)x * range;
  uint64_t l = (uint64_t)m;
  if (l < range) {
    const uint64_t t = (0 - range) % range;
    while (l < t) {
      x = mix64(x);
      m = (unsigned __int128)x * range;
      l = (uint64_t)m;
    }
  }
  return (uint64_t)(m >> 64);
}

// Uniform double in [0, 1) with 53 random bits.
static inline double unit_double(uint64_t x)
{
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace emst
