// contract.hip -- a related-work GPU baseline on MI355X: vertex-based
// Boruvka with graph contraction, the algorithm of the reference's UMinho
// GPU code (UMinho_BOTH/src/BoruvkaUMinho_GPU.cu: kernels 10-180, host loop
// 183-455).  Per iteration:
//   1. every vertex picks its lightest edge, ties to the smaller target
//      (find_min_per_vertex, 10-41);
//   2. of two vertices that picked each other, the smaller keeps no edge and
//      becomes the root (remove_duplicates, 68-85);
//   3. the picked edges are marked in the forest and each vertex's colour
//      is its target, then pointer jumping until no colour changes
//      (mark_mst_edges 87-95, initialize_color 43-51, propagate_color 53-66);
//   4. roots become the next graph's vertices (exclusive scan of flags,
//      create_new_vertex_id 98-108);
//   5. edges between different colours are counted per new vertex, scanned
//      into offsets and copied with their new endpoints and their original
//      CSR position (count_new_edges 110-129, insert_new_edges 139-166).
// The loop ends when the contracted graph has no edges.  The marked edges
// form a minimum spanning forest: its weight equals any MSF's, but under
// equal weights the edges chosen can differ from the (w, j) order of
// gpuMST, because ties go to the smaller target vertex here.
//
// MI355X choices (not the reference's): vertices of degree above 64 are
// walked by a whole wave instead of one thread, vertices of degree above
// 16384 (a supervertex that has absorbed most of a graph) are split into
// 8192-edge chunks, one workgroup each, and the per-new-vertex counts and
// insert cursors are combined per workgroup in LDS before one global atomic
// per (workgroup, new vertex), since same-address atomics serialise at about
// 12 ns each here and the largest supervertex is the target of most of
// them.  Equal (weight, target) picks go to the smaller original CSR
// position, so the result does not depend on the order in which a
// contracted adjacency was filled.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "emst_internal.hpp"
#include "emst_kernels.hpp"
#include "host_util.hpp"

using namespace emst;

namespace {

constexpr int kCBlock = 256;
constexpr int kCoopDeg = 64;   // degree above which a wave walks the adjacency
constexpr int kBigDeg = 16384; // degree above which workgroups share the adjacency
constexpr int kChunk = 8192;   // edges per workgroup of a big adjacency
constexpr int kCTable = 512;   // LDS combining slots (twice the workgroup size)
constexpr int kMaxIterations = 64;
constexpr int kMaxColorPasses = 64;

__device__ __forceinline__ u64 edge_key(int w, int d) { return ((u64)(unsigned)w << 32) | (unsigned)d; }

// Lighter of two (key, original position, edge) candidates; equal keys keep
// the smaller original position.
__device__ __forceinline__ void take_min(u64& k, int& o, int& e, u64 k2, int o2, int e2)
{
  if (e2 >= 0 && (e < 0 || k2 < k || (k2 == k && o2 < o))) {
    k = k2;
    o = o2;
    e = e2;
  }
}

__device__ __forceinline__ int orig_of(const int* map, int j) { return map ? map[j] : j; }

// 1. lightest edge of every vertex (self-loops never count), -1 if none;
// vertices above kBigDeg are left to the chunked kernels
__global__ __launch_bounds__(kCBlock) void k_cmin(int n, const int* __restrict__ idx, const int* __restrict__ dst,
                                                  const int* __restrict__ wt, const int* __restrict__ map,
                                                  int* __restrict__ minedge)
{
  const long long vl = (long long)blockIdx.x * kCBlock + threadIdx.x;
  const int v = (int)std::min<long long>(vl, INT_MAX);
  int beg = 0, end = 0;
  if (vl < n) {
    beg = idx[v];
    end = idx[v + 1];
  }
  const bool coop = end - beg > kCoopDeg && end - beg <= kBigDeg;
  if (vl < n && end - beg <= kCoopDeg) {
    u64 k = ~0ull;
    int o = 0, e = -1;
    for (int j = beg; j < end; j++) {
      const int d = dst[j];
      if (d != v) take_min(k, o, e, edge_key(wt[j], d), orig_of(map, j), j);
    }
    minedge[v] = e;
  }
  u64 todo = __ballot(coop);
  const int lane = lane_id();
  while (todo) {
    const int who = __builtin_ctzll(todo);
    todo &= todo - 1;
    const int wv = __shfl(v, who), wb = __shfl(beg, who), we = __shfl(end, who);
    u64 k = ~0ull;
    int o = 0, e = -1;
    for (int j = wb + lane; j < we; j += kWave) {
      const int d = dst[j];
      if (d != wv) take_min(k, o, e, edge_key(wt[j], d), orig_of(map, j), j);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const u64 k2 = __shfl_xor(k, off);
      const int o2 = __shfl_xor(o, off);
      const int e2 = __shfl_xor(e, off);
      take_min(k, o, e, k2, o2, e2);
    }
    if (lane == 0) minedge[wv] = e;
  }
}

// ---- big adjacencies: 8192-edge chunks, one workgroup each --------------
// A chunk is {big-vertex slot, first edge, end edge, vertex}.

// Lists the vertices above kBigDeg as {vertex, begin, end, 0}.
__global__ __launch_bounds__(kCBlock) void k_cbig_find(int n, const int* __restrict__ idx, int4* __restrict__ big,
                                                       int* __restrict__ nbig)
{
  const long long v = (long long)blockIdx.x * kCBlock + threadIdx.x;
  if (v < n) {
    const int b = idx[v], e = idx[v + 1];
    if (e - b > kBigDeg) big[atomicAdd(nbig, 1)] = make_int4((int)v, b, e, 0);
  }
}

__device__ __forceinline__ u64 wg_min_u64(u64 x, u64* s_part)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x = min(x, (u64)__shfl_xor(x, off));
  if (lane_id() == 0) s_part[threadIdx.x / kWave] = x;
  __syncthreads();
  u64 r = s_part[0];
  for (int w = 1; w < kCBlock / kWave; w++) r = min(r, s_part[w]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ int wg_sum(int x, int* s_part)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
  if (lane_id() == 0) s_part[threadIdx.x / kWave] = x;
  __syncthreads();
  int r = 0;
  for (int w = 0; w < kCBlock / kWave; w++) r += s_part[w];
  __syncthreads();
  return r;
}

// Lightest (weight, target) key of each big vertex ...
__global__ __launch_bounds__(kCBlock) void k_cbig_key(const int4* __restrict__ chunks, const int* __restrict__ dst,
                                                      const int* __restrict__ wt, u64* bigkey)
{
  __shared__ u64 s_part[kCBlock / kWave];
  const int4 c = chunks[blockIdx.x];
  u64 k = ~0ull;
  for (int j = c.y + threadIdx.x; j < c.z; j += kCBlock) {
    const int d = dst[j];
    if (d != c.w) k = min(k, edge_key(wt[j], d));
  }
  k = wg_min_u64(k, s_part);
  if (threadIdx.x == 0 && k != ~0ull) atomicMin(&bigkey[c.x], k);
}

// ... then, among its edges with that key, the smallest original position
// packed with the edge, (position << 32) | edge
__global__ __launch_bounds__(kCBlock) void k_cbig_pick(const int4* __restrict__ chunks, const int* __restrict__ dst,
                                                       const int* __restrict__ wt, const int* __restrict__ map,
                                                       const u64* __restrict__ bigkey, u64* bigpick)
{
  __shared__ u64 s_part[kCBlock / kWave];
  const int4 c = chunks[blockIdx.x];
  const u64 want = bigkey[c.x];
  u64 p = ~0ull;
  for (int j = c.y + threadIdx.x; j < c.z; j += kCBlock) {
    const int d = dst[j];
    if (d != c.w && edge_key(wt[j], d) == want) p = min(p, ((u64)(unsigned)orig_of(map, j) << 32) | (unsigned)j);
  }
  p = wg_min_u64(p, s_part);
  if (threadIdx.x == 0 && p != ~0ull) atomicMin(&bigpick[c.x], p);
}

__global__ __launch_bounds__(kCBlock) void k_cbig_set(const int4* __restrict__ big, int nbig,
                                                      const u64* __restrict__ bigpick, int* __restrict__ minedge)
{
  const int i = blockIdx.x * kCBlock + threadIdx.x;
  if (i < nbig) minedge[big[i].x] = bigpick[i] == ~0ull ? -1 : (int)(unsigned)(bigpick[i] & 0xffffffffu);
}

// Edges of a chunk between different colours, added to the new vertex's count.
__global__ __launch_bounds__(kCBlock) void k_cbig_count(const int4* __restrict__ chunks, const int* __restrict__ dst,
                                                        const int* __restrict__ color, const int* __restrict__ newid,
                                                        int* deg)
{
  __shared__ int s_part[kCBlock / kWave];
  const int4 c = chunks[blockIdx.x];
  const int cv = color[c.w];
  int k = 0;
  for (int j = c.y + threadIdx.x; j < c.z; j += kCBlock) k += color[dst[j]] != cv;
  k = wg_sum(k, s_part);
  if (threadIdx.x == 0 && k > 0) atomicAdd(&deg[newid[cv]], k);
}

// Copies a chunk's edges between different colours into a block reserved
// from the new vertex's cursor, in order.
__global__ __launch_bounds__(kCBlock) void k_cbig_insert(const int4* __restrict__ chunks, const int* __restrict__ dst,
                                                         const int* __restrict__ wt, const int* __restrict__ map,
                                                         const int* __restrict__ color, const int* __restrict__ newid,
                                                         int* cursor, int* __restrict__ ndst, int* __restrict__ nwt,
                                                         int* __restrict__ nmap)
{
  __shared__ int s_part[kCBlock / kWave];
  __shared__ long long s_base;
  const int4 c = chunks[blockIdx.x];
  const int cv = color[c.w];
  int k = 0;
  for (int j = c.y + threadIdx.x; j < c.z; j += kCBlock) k += color[dst[j]] != cv;
  k = wg_sum(k, s_part);
  if (k == 0) return;  // uniform across the workgroup
  if (threadIdx.x == 0) s_base = atomicAdd(&cursor[newid[cv]], k);
  __syncthreads();
  long long base = s_base;
  const int wid = threadIdx.x / kWave;
  for (int j0 = c.y; j0 < c.z; j0 += kCBlock) {
    const int j = j0 + threadIdx.x;
    int dc = cv;
    if (j < c.z) dc = color[dst[j]];
    const bool keep = dc != cv;
    const u64 m = __ballot(keep);
    if (lane_id() == 0) s_part[wid] = __popcll(m);
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < kCBlock / kWave; w++) {
      before += w < wid ? s_part[w] : 0;
      total += s_part[w];
    }
    if (keep) {
      const long long p = base + before + lanes_below(m);
      ndst[p] = newid[dc];
      nwt[p] = wt[j];
      nmap[p] = orig_of(map, j);
    }
    base += total;
    __syncthreads();
  }
}

// 2 + 3. mutual picks keep one edge; mark the picked edges; colour = target
__global__ __launch_bounds__(kCBlock) void k_cpick(int n, const int* __restrict__ dst, const int* __restrict__ map,
                                                   const int* __restrict__ minedge, int* __restrict__ color,
                                                   uint8_t* __restrict__ inMST)
{
  const long long v = (long long)blockIdx.x * kCBlock + threadIdx.x;
  if (v >= n) return;
  int e = minedge[v];
  int c = (int)v;
  if (e >= 0) {
    const int d = dst[e];
    const int e2 = minedge[d];
    if (e2 >= 0 && dst[e2] == (int)v && (int)v < d) e = -1;  // the smaller of a mutual pair is the root
    if (e >= 0) {
      c = d;
      inMST[map ? map[e] : e] = 1;
    }
  }
  color[v] = c;
}

// 3. pointer jumping: colour = colour of colour
__global__ __launch_bounds__(kCBlock) void k_cjump(int n, int* color, int* __restrict__ changed)
{
  const long long v = (long long)blockIdx.x * kCBlock + threadIdx.x;
  bool ch = false;
  if (v < n) {
    const int c = __hip_atomic_load(&color[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int cc = __hip_atomic_load(&color[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (c != cc) {
      __hip_atomic_store(&color[v], cc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ch = true;
    }
  }
  if (__ballot(ch) && lane_id() == 0) *changed = 1;
}

// 4. every root becomes a vertex of the next graph.  (The reference drops
// roots without out-edges, create_new_vertex_id 98-108; with an asymmetric
// CSR such a root can still be the target of a kept edge, so roots are
// kept, and an isolated one costs a slot per iteration.)  flag[n] = 0 closes
// the scan.
__global__ __launch_bounds__(kCBlock) void k_croots(int n, const int* __restrict__ color, int* __restrict__ flag)
{
  const long long v = (long long)blockIdx.x * kCBlock + threadIdx.x;
  if (v < n) flag[v] = color[v] == (int)v;
  if (v == n) flag[v] = 0;
}

// Workgroup-combined reservation: every thread brings (key, count); the
// counts of one key are summed in LDS and reserved from cursor[key] with one
// global atomic per key, and each thread gets the start of its own part.
// Every thread of the workgroup must call it.  At most kCBlock distinct keys
// occupy the kCTable slots, so linear probing always finds one.
__device__ long long wg_reserve(int key, int count, int* cursor, int* s_key, int* s_cnt, int* s_base)
{
  for (int i = threadIdx.x; i < kCTable; i += kCBlock) {
    s_key[i] = -1;
    s_cnt[i] = 0;
  }
  __syncthreads();
  int slot = -1, off = 0;
  if (count > 0) {
    int h = (int)(((unsigned)key * 0x9E3779B1u) >> 23);  // 9 bits: kCTable slots
    for (;;) {
      int k = s_key[h];
      if (k == -1) {
        k = atomicCAS(&s_key[h], -1, key);
        if (k == -1) k = key;
      }
      if (k == key) break;
      h = (h + 1) & (kCTable - 1);
    }
    slot = h;
    off = atomicAdd(&s_cnt[slot], count);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kCTable; i += kCBlock)
    if (s_key[i] >= 0) s_base[i] = atomicAdd(&cursor[s_key[i]], s_cnt[i]);
  __syncthreads();
  return slot >= 0 ? (long long)s_base[slot] + off : 0;
}

// 5a. edges between different colours, counted per new vertex
__global__ __launch_bounds__(kCBlock) void k_ccount(int n, const int* __restrict__ idx, const int* __restrict__ dst,
                                                    const int* __restrict__ color, const int* __restrict__ newid,
                                                    int* outdeg)
{
  __shared__ int s_key[kCTable], s_cnt[kCTable], s_base[kCTable];
  const long long vl = (long long)blockIdx.x * kCBlock + threadIdx.x;
  const int v = (int)std::min<long long>(vl, INT_MAX);
  int beg = 0, end = 0, c = -1;
  if (vl < n) {
    beg = idx[v];
    end = idx[v + 1];
    c = color[v];
  }
  const bool coop = end - beg > kCoopDeg && end - beg <= kBigDeg;
  int cnt = 0;
  if (vl < n && end - beg <= kCoopDeg)
    for (int j = beg; j < end; j++) cnt += color[dst[j]] != c;
  (void)wg_reserve(cnt > 0 ? newid[c] : 0, cnt, outdeg, s_key, s_cnt, s_base);
  u64 todo = __ballot(coop);
  const int lane = lane_id();
  while (todo) {
    const int who = __builtin_ctzll(todo);
    todo &= todo - 1;
    const int wb = __shfl(beg, who), we = __shfl(end, who), wc = __shfl(c, who);
    int k = 0;
    for (int j = wb + lane; j < we; j += kWave) k += color[dst[j]] != wc;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) k += __shfl_xor(k, off);
    if (lane == 0 && k > 0) atomicAdd(&outdeg[newid[wc]], k);
  }
}

// 5b. copy of every edge between different colours into the next graph:
// target = new id of the target's colour, weight, original CSR position
__global__ __launch_bounds__(kCBlock) void k_cinsert(int n, const int* __restrict__ idx, const int* __restrict__ dst,
                                                     const int* __restrict__ wt, const int* __restrict__ map,
                                                     const int* __restrict__ color, const int* __restrict__ newid,
                                                     int* cursor, int* __restrict__ ndst, int* __restrict__ nwt,
                                                     int* __restrict__ nmap)
{
  __shared__ int s_key[kCTable], s_cnt[kCTable], s_base[kCTable];
  const long long vl = (long long)blockIdx.x * kCBlock + threadIdx.x;
  const int v = (int)std::min<long long>(vl, INT_MAX);
  int beg = 0, end = 0, c = -1;
  if (vl < n) {
    beg = idx[v];
    end = idx[v + 1];
    c = color[v];
  }
  const bool coop = end - beg > kCoopDeg && end - beg <= kBigDeg;
  int cnt = 0;
  if (vl < n && end - beg <= kCoopDeg)
    for (int j = beg; j < end; j++) cnt += color[dst[j]] != c;
  long long pos = wg_reserve(cnt > 0 ? newid[c] : 0, cnt, cursor, s_key, s_cnt, s_base);
  if (cnt > 0)
    for (int j = beg; j < end; j++) {
      const int dc = color[dst[j]];
      if (dc != c) {
        ndst[pos] = newid[dc];
        nwt[pos] = wt[j];
        nmap[pos] = orig_of(map, j);
        pos++;
      }
    }
  u64 todo = __ballot(coop);
  const int lane = lane_id();
  while (todo) {
    const int who = __builtin_ctzll(todo);
    todo &= todo - 1;
    const int wb = __shfl(beg, who), we = __shfl(end, who), wc = __shfl(c, who);
    int k = 0;
    for (int j = wb + lane; j < we; j += kWave) k += color[dst[j]] != wc;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) k += __shfl_xor(k, off);
    if (k == 0) continue;
    long long base = 0;
    if (lane == 0) base = atomicAdd(&cursor[newid[wc]], k);
    base = __shfl(base, 0);
    for (int j0 = wb; j0 < we; j0 += kWave) {
      const int j = j0 + lane;
      int dc = wc;
      if (j < we) dc = color[dst[j]];
      const bool keep = dc != wc;
      const u64 m = __ballot(keep);
      if (keep) {
        const long long p = base + lanes_below(m);
        ndst[p] = newid[dc];
        nwt[p] = wt[j];
        nmap[p] = orig_of(map, j);
      }
      base += __popcll(m);
    }
  }
}

inline unsigned cblocks(long long n) { return (unsigned)std::max<long long>(1, (n + kCBlock - 1) / kCBlock); }

struct Buf {
  void* p = nullptr;
  ~Buf()
  {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

// True when every directed edge (v, u, w) has a reverse (u, v, w), with the
// same multiplicity.  Each list is copied as (neighbour, weight) keys and
// sorted; then, per run of equal keys in v's list, the run of (v, w) in u's
// list must have the same length.  O(m log d) host work, outside the timed
// loop.
static bool csr_symmetric(const emst_graph* g)
{
  const int64_t n = g->nodes;
  std::vector<uint64_t> key((size_t)g->edges);
  emst::parallel_for(n, [&](int64_t b, int64_t e, int) {
    for (int64_t v = b; v < e; v++) {
      const int32_t lo = g->nindex[v], hi = g->nindex[v + 1];
      for (int32_t j = lo; j < hi; j++)
        key[j] = ((uint64_t)(uint32_t)g->nlist[j] << 32) | (uint32_t)g->eweight[j];
      std::sort(key.begin() + lo, key.begin() + hi);
    }
  });
  std::vector<char> bad((size_t)emst::host_threads(), 0);
  emst::parallel_for(n, [&](int64_t b, int64_t e, int t) {
    for (int64_t v = b; v < e && !bad[t]; v++) {
      const int32_t lo = g->nindex[v], hi = g->nindex[v + 1];
      for (int32_t j = lo; j < hi;) {
        const uint64_t k = key[j];
        int32_t r = j + 1;
        while (r < hi && key[r] == k) r++;
        const int64_t u = (int64_t)(k >> 32);
        const uint64_t back = ((uint64_t)(uint32_t)v << 32) | (uint32_t)k;
        const auto rb = key.begin() + g->nindex[u], re = key.begin() + g->nindex[u + 1];
        const auto range = std::equal_range(rb, re, back);
        if (range.second - range.first != r - j) {
          bad[t] = 1;
          break;
        }
        j = r;
      }
    }
  }, (int)bad.size());
  for (char c : bad)
    if (c) return false;
  return true;
}

extern "C" int emst_contract_mst(emst_ctx* ctx, const emst_graph* g, uint8_t* inMST, emst_contract_stats* stats)
{
  if (!ctx || !g || !inMST) return EMST_E_ARG;
  int rc = emst_load(ctx, g);
  if (rc) return rc;
  // one-way edges can close a cycle of picks: such inputs are refused
  if (!csr_symmetric(g)) return ctx_fail(ctx, EMST_E_GRAPH, "the contraction baseline needs a symmetric CSR");
  const int n0 = g->nodes, m0 = g->edges;
  hipStream_t s = ctx->stream;
  auto fail = [&](hipError_t e, const char* what) {
    char b[256];
    std::snprintf(b, sizeof(b), "HIP error %d in %s: %s", (int)e, what, hipGetErrorString(e));
    return ctx_fail(ctx, EMST_E_HIP, b);
  };
#define CCHK(call)                             \
  do {                                         \
    hipError_t e_ = (call);                    \
    if (e_ != hipSuccess) return fail(e_, #call); \
  } while (0)
  // two graph buffers (ping-pong) and the per-vertex arrays
  const size_t nv = (size_t)n0 + 1, ne = (size_t)std::max(m0, 1);
  Buf gidx[2], gdst[2], gwt[2], gmap[2], b_min, b_color, b_flag, b_newid, b_deg, b_off, b_cur, b_changed, b_tmp;
  for (int i = 0; i < 2; i++)
    if (hipMalloc(&gidx[i].p, nv * 4) || hipMalloc(&gdst[i].p, ne * 4) || hipMalloc(&gwt[i].p, ne * 4) ||
        hipMalloc(&gmap[i].p, ne * 4))
      return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of a contracted graph failed");
  if (hipMalloc(&b_min.p, nv * 4) || hipMalloc(&b_color.p, nv * 4) || hipMalloc(&b_flag.p, nv * 4) ||
      hipMalloc(&b_newid.p, nv * 4) || hipMalloc(&b_deg.p, nv * 4) || hipMalloc(&b_off.p, nv * 4) ||
      hipMalloc(&b_cur.p, nv * 4) || hipMalloc(&b_changed.p, 16))
    return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of the vertex arrays failed");
  // big-vertex list and chunk table: a big vertex has more than kBigDeg edges
  const size_t nbmax = (size_t)m0 / kBigDeg + 1, chmax = (size_t)m0 / kChunk + nbmax + 1;
  Buf b_big, b_nbig, b_chunks, b_bigkey, b_bigpick;
  if (hipMalloc(&b_big.p, nbmax * sizeof(int4)) || hipMalloc(&b_nbig.p, 16) ||
      hipMalloc(&b_chunks.p, chmax * sizeof(int4)) || hipMalloc(&b_bigkey.p, nbmax * 8) ||
      hipMalloc(&b_bigpick.p, nbmax * 8))
    return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of the big-vertex tables failed");
  size_t tmp_bytes = 0;
  CCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, (int*)b_flag.p, (int*)b_newid.p, (int)nv, s));
  if (hipMalloc(&b_tmp.p, std::max<size_t>(tmp_bytes, 16))) return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of scan storage failed");
  int* minedge = (int*)b_min.p;
  int* color = (int*)b_color.p;
  int* flag = (int*)b_flag.p;
  int* newid = (int*)b_newid.p;
  int* deg = (int*)b_deg.p;
  int* cursor = (int*)b_cur.p;
  int* changed = (int*)b_changed.p;
  if (m0 > 0) CCHK(hipMemsetAsync(ctx->d_inMST, 0, (size_t)m0, s));
  CCHK(hipStreamSynchronize(s));

  CCHK(hipEventRecord(ctx->ev[0], s));
  // iteration 0 reads the loaded CSR in place; its map is the identity
  const int* idx = ctx->d_nindex;
  const int* dst = ctx->d_nlist;
  const int* wt = ctx->d_eweight;
  const int* map = nullptr;
  int n = n0;
  long long m = m0;
  int it = 0, passes = 0, cur = 0;
  while (m > 0) {
    if (++it > kMaxIterations) return ctx_fail(ctx, EMST_E_LIMIT, "contraction did not converge");
    // big adjacencies: listed, cut into chunks on the host, minimum by chunks
    CCHK(hipMemsetAsync(b_nbig.p, 0, 4, s));
    hipLaunchKernelGGL(k_cbig_find, dim3(cblocks(n)), dim3(kCBlock), 0, s, n, idx, (int4*)b_big.p, (int*)b_nbig.p);
    int nb = 0;
    CCHK(hipMemcpyAsync(&nb, b_nbig.p, 4, hipMemcpyDeviceToHost, s));
    CCHK(hipStreamSynchronize(s));
    int nchunks = 0;
    if (nb > 0) {
      if ((size_t)nb > nbmax) return ctx_fail(ctx, EMST_E_LIMIT, "big-vertex list overflow");
      std::vector<int4> hb((size_t)nb), hc;
      CCHK(hipMemcpy(hb.data(), b_big.p, (size_t)nb * sizeof(int4), hipMemcpyDeviceToHost));
      for (int i = 0; i < nb; i++)
        for (int c = hb[i].y; c < hb[i].z; c += std::min(kChunk, hb[i].z - c))
          hc.push_back(make_int4(i, c, (int)std::min<long long>((long long)c + kChunk, hb[i].z), hb[i].x));
      if (hc.size() > chmax) return ctx_fail(ctx, EMST_E_LIMIT, "chunk table overflow");
      nchunks = (int)hc.size();
      // on the solve's stream (non-blocking, so not ordered with the null
      // stream); hc must outlive the copy, hence the synchronisation
      CCHK(hipMemcpyAsync(b_chunks.p, hc.data(), hc.size() * sizeof(int4), hipMemcpyHostToDevice, s));
      CCHK(hipStreamSynchronize(s));
      CCHK(hipMemsetAsync(b_bigkey.p, 0xff, (size_t)nb * 8, s));
      CCHK(hipMemsetAsync(b_bigpick.p, 0xff, (size_t)nb * 8, s));
      hipLaunchKernelGGL(k_cbig_key, dim3(nchunks), dim3(kCBlock), 0, s, (const int4*)b_chunks.p, dst, wt,
                         (u64*)b_bigkey.p);
      hipLaunchKernelGGL(k_cbig_pick, dim3(nchunks), dim3(kCBlock), 0, s, (const int4*)b_chunks.p, dst, wt, map,
                         (const u64*)b_bigkey.p, (u64*)b_bigpick.p);
      hipLaunchKernelGGL(k_cbig_set, dim3(cblocks(nb)), dim3(kCBlock), 0, s, (const int4*)b_big.p, nb,
                         (const u64*)b_bigpick.p, minedge);
    }
    hipLaunchKernelGGL(k_cmin, dim3(cblocks(n)), dim3(kCBlock), 0, s, n, idx, dst, wt, map, minedge);
    hipLaunchKernelGGL(k_cpick, dim3(cblocks(n)), dim3(kCBlock), 0, s, n, dst, map, minedge, color, ctx->d_inMST);
    for (;;) {
      if (++passes > kMaxColorPasses * kMaxIterations) return ctx_fail(ctx, EMST_E_LIMIT, "colours did not converge");
      CCHK(hipMemsetAsync(changed, 0, 4, s));
      hipLaunchKernelGGL(k_cjump, dim3(cblocks(n)), dim3(kCBlock), 0, s, n, color, changed);
      int h = 0;
      CCHK(hipMemcpyAsync(&h, changed, 4, hipMemcpyDeviceToHost, s));
      CCHK(hipStreamSynchronize(s));
      if (!h) break;
    }
    // new vertex ids: exclusive scan of the root flags (n + 1 entries)
    hipLaunchKernelGGL(k_croots, dim3(cblocks((long long)n + 1)), dim3(kCBlock), 0, s, n, color, flag);
    CCHK(hipcub::DeviceScan::ExclusiveSum(b_tmp.p, tmp_bytes, flag, newid, n + 1, s));
    int nn = 0;
    CCHK(hipMemcpyAsync(&nn, newid + n, 4, hipMemcpyDeviceToHost, s));
    CCHK(hipStreamSynchronize(s));
    // edge counts per new vertex, offsets, then the copy
    const int nxt = cur ^ (it > 1 ? 1 : 0);
    int* nidx = (int*)gidx[nxt].p;
    CCHK(hipMemsetAsync(deg, 0, ((size_t)nn + 1) * 4, s));
    hipLaunchKernelGGL(k_ccount, dim3(cblocks(n)), dim3(kCBlock), 0, s, n, idx, dst, color, newid, deg);
    if (nchunks)
      hipLaunchKernelGGL(k_cbig_count, dim3(nchunks), dim3(kCBlock), 0, s, (const int4*)b_chunks.p, dst, color, newid,
                         deg);
    CCHK(hipcub::DeviceScan::ExclusiveSum(b_tmp.p, tmp_bytes, deg, nidx, nn + 1, s));
    int mm = 0;
    CCHK(hipMemcpyAsync(&mm, nidx + nn, 4, hipMemcpyDeviceToHost, s));
    CCHK(hipMemcpyAsync(cursor, nidx, ((size_t)nn + 1) * 4, hipMemcpyDeviceToDevice, s));
    CCHK(hipStreamSynchronize(s));
    if (mm > 0) {
      hipLaunchKernelGGL(k_cinsert, dim3(cblocks(n)), dim3(kCBlock), 0, s, n, idx, dst, wt, map, color, newid, cursor,
                         (int*)gdst[nxt].p, (int*)gwt[nxt].p, (int*)gmap[nxt].p);
      if (nchunks)
        hipLaunchKernelGGL(k_cbig_insert, dim3(nchunks), dim3(kCBlock), 0, s, (const int4*)b_chunks.p, dst, wt, map,
                           color, newid, cursor, (int*)gdst[nxt].p, (int*)gwt[nxt].p, (int*)gmap[nxt].p);
    }
    CCHK(hipGetLastError());
    idx = nidx;
    dst = (const int*)gdst[nxt].p;
    wt = (const int*)gwt[nxt].p;
    map = (const int*)gmap[nxt].p;
    cur = nxt;
    n = nn;
    m = mm;
  }
  CCHK(hipEventRecord(ctx->ev[1], s));
  CCHK(hipEventSynchronize(ctx->ev[1]));
  float ms = 0;
  CCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  if (m0 > 0) CCHK(hipMemcpy(inMST, ctx->d_inMST, (size_t)m0, hipMemcpyDeviceToHost));
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->ms = ms;
    stats->iterations = it;
    stats->color_passes = passes;
    for (int32_t j = 0; j < m0; j++)
      if (inMST[j]) {
        stats->mst_edges++;
        stats->weight += (uint64_t)(uint32_t)g->eweight[j];
      }
  }
#undef CCHK
  return EMST_OK;
}
