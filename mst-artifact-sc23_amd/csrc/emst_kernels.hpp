// emst_kernels.hpp -- gfx950 device code for the edge-based, data-driven
// Boruvka rounds of ECL-MST (reference: ECL-MST/ECL_MST.cu:127-251).
//
// Worklist records are 16 bytes, {x, y, weight, j}, where x and y are the
// endpoints (or their component representatives) and j is the
// upper-direction CSR position of the edge (ECL_MST.cu:176, 201).  The
// min-edge key is ((uint64)(uint32)weight << 32) | j (ECL_MST.cu:221).
//
// Everything is written for 64-lane wavefronts: worklist appends are
// wave-aggregated into a per-wave LDS staging buffer and flushed with one
// global atomicAdd per buffer, and CSR streaming assigns 64 consecutive
// edges to the 64 lanes of a wave (coalesced 256-byte loads).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <climits>

namespace emst {

typedef unsigned long long u64;

constexpr int kWave = 64;
constexpr u64 kNoKey = ~0ull;                 // ECL_MST.cu:152, 248

// Indices of the device statistics array.
enum {
  ST_LIGHT = 0,       // records written by the light initialisation
  ST_HEAVY = 1,       // records written by the heavy initialisation
  ST_ROUNDS0 = 2,     // rounds with a non-empty worklist, light phase
  ST_ROUNDS1 = 3,     // same, heavy phase
  ST_TAIL = 4,        // of those, rounds run inside the one-workgroup tail
  ST_IN_SUM = 5,      // sum of worklist sizes entering a round
  ST_OUT_SUM = 6,     // sum of surviving records
  ST_COUNT = 8
};

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Streaming accesses that are read or written once per pass are issued
// non-temporal, so they do not evict the randomly accessed per-vertex
// arrays from L2 and the memory-side cache.
typedef int v4i __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int ld_stream(const int* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_stream(int4* p, int4 r)
{
  v4i v = {r.x, r.y, r.z, r.w};
  __builtin_nontemporal_store(v, reinterpret_cast<v4i*>(p));
}

// Number of set bits of m in lanes below this one.
__device__ __forceinline__ int lanes_below(u64 m)
{
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Orders this wave's LDS accesses: LDS operations of one wave execute in
// issue order, so a compiler barrier plus wavefront-scope fences suffice.
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ u64 make_key(int w, int j)
{
  return ((u64)(unsigned)w << 32) | (u64)(unsigned)j;
}

// Root of v.  Hooking only ever points a larger id at a smaller one, so
// chains are acyclic and the root is the smallest id of the tree.  Read-only:
// path compression happens by writing representatives back into records
// (ECL_MST.cu:127-134, 218-219).
__device__ __forceinline__ int find_root(int v, const int* __restrict__ parent)
{
  int p = parent[v];
  while (p != v) {
    v = p;
    p = parent[v];
  }
  return v;
}

// Same, with every load performed at device (agent) scope so that it sees
// hooks made earlier in the same kernel by other waves.
__device__ __forceinline__ int find_root_coherent(int v, int* parent)
{
  int p = __hip_atomic_load(&parent[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (p != v) {
    v = p;
    p = __hip_atomic_load(&parent[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return v;
}

// Links the trees of a and b: the larger root is CAS-hooked under the
// smaller; when the CAS finds the larger id no longer a root it retries
// from that id's new parent (ECL_MST.cu:137-144).
__device__ __forceinline__ void hook(int a, int b, int* parent)
{
  for (;;) {
    const int hi = a > b ? a : b;
    const int lo = a > b ? b : a;
    const int seen = atomicCAS(&parent[hi], hi, lo);
    if (seen == hi) return;
    a = lo;
    b = seen;
  }
}

// Guarded 64-bit atomic min (ECL_MST.cu:222-223): skip the atomic when a
// relaxed device-scope read already shows a key no larger.  Within a round
// minv only decreases, so a stale read can only cost an extra atomic.
// Returns false when the key cannot be the minimum (a key no larger was
// already there); true otherwise.
__device__ __forceinline__ bool min_key(u64* minv, int rep, u64 key, bool guard)
{
  if (guard) {
    const u64 cur = __hip_atomic_load(&minv[rep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur <= key) return false;
  }
  atomicMin(&minv[rep], key);
  return true;
}

// ---- per-wave LDS staging of worklist appends ------------------------------
// Each workgroup of an initialisation kernel owns a private segment of the
// output list; a flush reserves its range with an atomic on a workgroup
// counter in LDS, so no two workgroups ever touch the same counter (a
// global same-address atomic serialises at about 12 ns on MI355X).
struct Stage {
  int4* buf;   // this wave's LDS slice
  int cap;     // records it holds (>= 64)
  int n;       // records currently staged (wave-uniform)
};

__device__ __forceinline__ void stage_flush(Stage& s, int4* __restrict__ seg, int* fill)
{
  if (s.n == 0) return;
  wave_lds_sync();
  int base = 0;
  if (lane_id() == 0) base = atomicAdd(fill, s.n);
  base = __shfl(base, 0);
  for (int i = lane_id(); i < s.n; i += kWave) st_stream(&seg[(long long)base + i], s.buf[i]);
  wave_lds_sync();
  s.n = 0;
}

// Wave-uniform call: lanes with keep append rec.
__device__ __forceinline__ void stage_push(Stage& s, bool keep, int4 rec, int4* __restrict__ seg, int* fill)
{
  const u64 mask = __ballot(keep);
  const int k = __popcll(mask);
  if (k == 0) return;
  if (s.n + k > s.cap) stage_flush(s, seg, fill);
  if (keep) s.buf[s.n + lanes_below(mask)] = rec;
  s.n += k;
}

// Inclusive maximum-scan across the 64 lanes of a wavefront with DPP moves
// (ALU-speed, no LDS crossbar): doubling shifts inside each row of 16
// lanes, then the last lane of rows 0 and 2 into rows 1 and 3, then lane 31
// into rows 2 and 3.  Lanes without a source take INT_MIN, the identity.
__device__ __forceinline__ int wave_max_scan(int x)
{
  x = max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, 0x111, 0xf, 0xf, false));  // row_shr:1
  x = max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, 0x112, 0xf, 0xf, false));  // row_shr:2
  x = max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, 0x114, 0xf, 0xf, false));  // row_shr:4
  x = max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, 0x118, 0xf, 0xf, false));  // row_shr:8
  x = max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
  x = max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return x;
}

// Number of lanes whose bnd is <= p, for bnd sorted ascending across the
// wave and p < bnd of lane 63 (so the answer is at most 63).  Six
// cross-lane reads (ds_bpermute) of a branch-free binary search.
__device__ __forceinline__ int count_le(int bnd, int p)
{
  int c = 0;
#pragma unroll
  for (int st = 32; st >= 1; st >>= 1) {
    const int v = __shfl(bnd, c + st - 1);
    if (v <= p) c += st;
  }
  return c;
}

}  // namespace emst
