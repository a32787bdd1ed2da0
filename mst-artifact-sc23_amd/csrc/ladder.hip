// ladder.hip -- the reference's kernel sequence for gpuMST with each of its
// optimizations switchable: the de-optimization ladder of
// Deoptimizations/ECL_MST_No*.cu, where each file removes one more
// optimization from ECL_MST.cu.
//
// The sequence (ECL_MST.cu:304-340): initPM; per phase initWL, then rounds
// of kernel1 (find roots, drop edges inside one tree, offer keys), kernel2
// (join the minimum edges) and kernel3 (reset the offered minima), with the
// worklist size read back by the host after every kernel1.  The switches:
//   guard       guarded atomicMin (ECL_MST.cu:222-223)
//   warp        wave-cooperative adjacency walks at degree >= 4 in initWL
//               (ECL_MST.cu:167, 180-205); off: one thread per vertex
//   implicit PC kernel1 writes the roots it found into the records
//               (ECL_MST.cu:218-219); off: records keep their endpoints and
//               kernel2/kernel3 find the roots again, with path splitting
//               (NoImplicitPathCompression.cu:121-131, 198-199, 213-216)
//   one dir.    only u > v enters the worklist (ECL_MST.cu:170, 195)
//   tuples      16-byte {x, y, w, j} records; off: edge ids and a per-edge
//               source array (NoTuples.cu:153-187)
//   data-driven each round compacts the crossing records; off: every round
//               revisits the phase's whole edge set and a flag ends the loop
//               (NoDataDriven.cu:166-176, 262-270)
//   edge-based  one thread per record; off: one thread per vertex walks its
//               adjacency in every kernel, no worklist (NoEdgeBased.cu:152-197)
// The filter (two phases at a light-edge threshold) is the caller's filter
// mode, as for the fused schedule.
//
// Every combination joins only edges that were a component's minimum at
// the start of the round (a round-start minimum stays a safe edge whatever
// else joins), and the key order is strict, so every combination gives the
// same forest, byte for byte, as the fused schedule.
//
// MI355X choices that are not rungs of the ladder: appends are aggregated
// per 64-lane wave (one counter atomic per wave instead of one per record:
// same-address atomics cost ~12 ns each here), and the cooperative walks use
// the 64-lane wave where the reference uses a 32-thread warp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>

#include "emst_internal.hpp"
#include "emst_kernels.hpp"

using namespace emst;

namespace {

constexpr int kLadBlock = 512;  // ThreadsPerBlock of the reference (ECL_MST.cu:48)

struct LadderMode {
  bool guard, warp, implicit_pc, one_dir, tuples, data_driven, edge_based;
};

// Relaxed device-scope accesses to parent for kernels in which other waves
// hook or split concurrently.
__device__ __forceinline__ int ld_parent(int* parent, int v)
{
  return __hip_atomic_load(&parent[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of v with path splitting: every vertex on the path is pointed at its
// grandparent.  Concurrent splitting and hooking are safe: a split only
// points a non-root at one of its ancestors, and roots change only by the
// CAS in hook.
__device__ __forceinline__ int find_split(int v, int* parent)
{
  int cur = ld_parent(parent, v);
  if (cur == v) return v;
  int prev = v;
  for (;;) {
    const int next = ld_parent(parent, cur);
    if (next == cur) return cur;
    __hip_atomic_store(&parent[prev], next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    prev = cur;
    cur = next;
  }
}

__device__ __forceinline__ int lad_find(int v, int* parent, bool split)
{
  return split ? find_split(v, parent) : find_root_coherent(v, parent);
}

// Appends rec for the active lanes with keep: one counter atomic per wave,
// issued by the lowest active lane.  Callable from divergent code (the
// per-thread adjacency loops have different trip counts per lane).
template <class Rec>
__device__ __forceinline__ void wave_append(bool keep, const Rec& rec, Rec* __restrict__ out, int* cnt)
{
  const u64 active = __ballot(1);
  const u64 m = __ballot(keep);
  if (m == 0) return;
  const int leader = __builtin_ctzll(active);
  int base = 0;
  if (lane_id() == leader) base = atomicAdd(cnt, __popcll(m));
  base = __shfl(base, leader);
  if (keep) out[base + lanes_below(m)] = rec;
}

// Sets *flag when any active lane has set: one store per wave.
__device__ __forceinline__ void wave_flag(bool set, int* flag)
{
  const u64 active = __ballot(1);
  if (__ballot(set) && lane_id() == __builtin_ctzll(active)) *flag = 1;
}

// Whether CSR position j (source v, target n, weight w) belongs to the phase:
// light w <= T, heavy w > T (ECL_MST.cu:172, 197), and n > v in one-direction
// mode (ECL_MST.cu:170, 195).
__device__ __forceinline__ bool in_phase(int v, int n, int w, int T, bool heavy, bool one_dir)
{
  if (one_dir && n <= v) return false;
  return heavy ? w > T : w <= T;
}

// initPM (ECL_MST.cu:147-154).
__global__ __launch_bounds__(kLadBlock) void k_lad_setup(int nodes, int* __restrict__ parent, u64* __restrict__ minv)
{
  const long long v = (long long)blockIdx.x * kLadBlock + threadIdx.x;
  if (v < nodes) {
    parent[v] = (int)v;
    minv[v] = kNoKey;
  }
}

// initWL (ECL_MST.cu:157-206): the phase's edges with their roots (the light
// phase starts from singletons, so its roots are the endpoints; the heavy
// phase keeps edges whose roots differ).  With slist, the source of every
// position is recorded (edge-id worklists); with list == nullptr nothing is
// appended (rounds that revisit every edge need no list).
template <bool TUPLES>
__device__ __forceinline__ void init_edge(int v, int a, int j, const int* __restrict__ nlist,
                                          const int* __restrict__ eweight, int* parent, int T, bool heavy,
                                          bool one_dir, int* __restrict__ slist, void* list, int* cnt)
{
  const int n = nlist[j];
  if (slist) slist[j] = v;
  const int w = eweight[j];
  bool keep = in_phase(v, n, w, T, heavy, one_dir);
  int b = n;
  if (keep && heavy) {
    b = find_root(n, parent);
    keep = a != b;
  }
  if (!list) return;
  if (TUPLES)
    wave_append(keep, make_int4(a, b, w, j), static_cast<int4*>(list), cnt);
  else
    wave_append(keep, j, static_cast<int*>(list), cnt);
}

template <bool TUPLES>
__global__ __launch_bounds__(kLadBlock) void k_lad_init(int nodes, const int* __restrict__ nindex,
                                                        const int* __restrict__ nlist, const int* __restrict__ eweight,
                                                        int* parent, int T, int heavy, int one_dir, int warp,
                                                        int* __restrict__ slist, void* list, int* cnt)
{
  const long long vl = (long long)blockIdx.x * kLadBlock + threadIdx.x;
  const int v = (int)std::min<long long>(vl, INT_MAX);
  int beg = 0, end = 0;
  if (vl < nodes) {
    beg = nindex[v];
    end = nindex[v + 1];
  }
  const bool coop = warp && end - beg >= 4;
  if (vl < nodes && !coop) {
    const int a = heavy ? find_root(v, parent) : v;
    for (int j = beg; j < end; j++) init_edge<TUPLES>(v, a, j, nlist, eweight, parent, T, heavy, one_dir, slist, list, cnt);
  }
  if (!warp) return;
  // the wave walks each high-degree adjacency of its lanes together
  u64 todo = __ballot(coop);
  const int lane = lane_id();
  while (todo) {
    const int who = __builtin_ctzll(todo);
    todo &= todo - 1;
    const int wv = __shfl(v, who), wb = __shfl(beg, who), we = __shfl(end, who);
    const int a = heavy ? find_root(wv, parent) : wv;
    for (int j = wb + lane; j < we; j += kWave)
      init_edge<TUPLES>(wv, a, j, nlist, eweight, parent, T, heavy, one_dir, slist, list, cnt);
  }
}

// The edge a worklist entry (or, with no list, CSR position i) stands for.
struct EdgeRef {
  int x, y, w, j;
  bool live;  // false: position outside the phase (no-list mode only)
};

template <bool TUPLES>
__device__ __forceinline__ EdgeRef load_edge(const void* list, long long i, const int* __restrict__ slist,
                                             const int* __restrict__ nlist, const int* __restrict__ eweight, int T,
                                             bool heavy, bool one_dir)
{
  EdgeRef e;
  if (!list) {  // every CSR position, filtered by the phase
    e.j = (int)i;
    e.x = slist[i];
    e.y = nlist[i];
    e.w = eweight[i];
    e.live = in_phase(e.x, e.y, e.w, T, heavy, one_dir);
  } else if (TUPLES) {
    const int4 r = static_cast<const int4*>(list)[i];
    e.x = r.x;
    e.y = r.y;
    e.w = r.z;
    e.j = r.w;
    e.live = true;
  } else {
    e.j = static_cast<const int*>(list)[i];
    e.x = slist[e.j];
    e.y = nlist[e.j];
    e.w = eweight[e.j];
    e.live = true;
  }
  return e;
}

// kernel1 (ECL_MST.cu:210-226).  Data-driven: crossing entries are appended
// to out (with their roots when implicit PC is on).  Otherwise the entries
// stay where they are (tuples take their new roots in place when implicit
// PC is on) and *go records whether any entry crosses.
template <bool TUPLES>
__global__ __launch_bounds__(kLadBlock) void k_lad_select(void* list, long long size, void* out, int* cnt_out,
                                                          int* go, int* parent, u64* __restrict__ minv,
                                                          const int* __restrict__ slist, const int* __restrict__ nlist,
                                                          const int* __restrict__ eweight, int T, int heavy,
                                                          int one_dir, int guard, int implicit_pc)
{
  const long long i = (long long)blockIdx.x * kLadBlock + threadIdx.x;
  if (i >= size) return;
  const EdgeRef e = load_edge<TUPLES>(list, i, slist, nlist, eweight, T, heavy, one_dir);
  bool cross = false;
  int a = e.x, b = e.y;
  if (e.live) {
    a = implicit_pc ? find_root(e.x, parent) : find_split(e.x, parent);
    b = implicit_pc ? find_root(e.y, parent) : find_split(e.y, parent);
    cross = a != b;
  }
  if (out) {
    if (TUPLES)
      wave_append(cross, implicit_pc ? make_int4(a, b, e.w, e.j) : make_int4(e.x, e.y, e.w, e.j),
                  static_cast<int4*>(out), cnt_out);
    else
      wave_append(cross, e.j, static_cast<int*>(out), cnt_out);
  } else {
    if (TUPLES && list && implicit_pc && (a != e.x || b != e.y))
      static_cast<int4*>(list)[i] = make_int4(a, b, e.w, e.j);
    wave_flag(cross, go);
  }
  if (cross) {
    const u64 key = make_key(e.w, e.j);
    min_key(minv, a, key, guard);
    min_key(minv, b, key, guard);
  }
}

// kernel2 (ECL_MST.cu:229-240): an entry whose key is the minimum of one of
// its roots joins them.  refind: the entry's endpoints may be stale (no
// implicit PC, edge ids, or a list that kept non-crossing entries), so the
// roots are found again; split selects path splitting for those finds.
template <bool TUPLES>
__global__ __launch_bounds__(kLadBlock) void k_lad_join(const void* list, long long size, int* parent,
                                                        const u64* __restrict__ minv, uint8_t* __restrict__ inMST,
                                                        const int* __restrict__ slist, const int* __restrict__ nlist,
                                                        const int* __restrict__ eweight, int T, int heavy, int one_dir,
                                                        int refind, int split)
{
  const long long i = (long long)blockIdx.x * kLadBlock + threadIdx.x;
  if (i >= size) return;
  const EdgeRef e = load_edge<TUPLES>(list, i, slist, nlist, eweight, T, heavy, one_dir);
  if (!e.live) return;
  const int a = refind ? lad_find(e.x, parent, split) : e.x;
  const int b = refind ? lad_find(e.y, parent, split) : e.y;
  if (a == b) return;
  const u64 key = make_key(e.w, e.j);
  if (key == minv[a] || key == minv[b]) {
    hook(a, b, parent);
    inMST[e.j] = 1;
  }
}

// kernel3 (ECL_MST.cu:243-251): every root that was offered a key is an
// endpoint root of some entry, so resetting the entries' roots clears every
// offer before the next round (and before the heavy phase).
template <bool TUPLES>
__global__ __launch_bounds__(kLadBlock) void k_lad_reset(const void* list, long long size, int* parent,
                                                         u64* __restrict__ minv, const int* __restrict__ slist,
                                                         const int* __restrict__ nlist, const int* __restrict__ eweight,
                                                         int T, int heavy, int one_dir, int refind, int split)
{
  const long long i = (long long)blockIdx.x * kLadBlock + threadIdx.x;
  if (i >= size) return;
  const EdgeRef e = load_edge<TUPLES>(list, i, slist, nlist, eweight, T, heavy, one_dir);
  if (!e.live) return;
  const int a = refind ? lad_find(e.x, parent, split) : e.x;
  const int b = refind ? lad_find(e.y, parent, split) : e.y;
  minv[a] = kNoKey;
  minv[b] = kNoKey;
}

// Vertex-based rounds (NoEdgeBased.cu:152-197): one thread per vertex walks
// its adjacency in every kernel; the phase's edges are picked by weight (and
// direction) on every visit.
__global__ __launch_bounds__(kLadBlock) void k_lad_vselect(int nodes, const int* __restrict__ nindex,
                                                           const int* __restrict__ nlist,
                                                           const int* __restrict__ eweight, int* parent,
                                                           u64* __restrict__ minv, int T, int heavy, int one_dir,
                                                           int guard, int split, int* go)
{
  const long long vl = (long long)blockIdx.x * kLadBlock + threadIdx.x;
  const int v = (int)std::min<long long>(vl, INT_MAX);
  bool any = false;
  if (vl < nodes) {
    const int beg = nindex[v], end = nindex[v + 1];
    for (int j = beg; j < end; j++) {
      const int n = nlist[j];
      const int w = eweight[j];
      if (!in_phase(v, n, w, T, heavy, one_dir)) continue;
      const int a = split ? find_split(v, parent) : find_root(v, parent);
      const int b = split ? find_split(n, parent) : find_root(n, parent);
      if (a != b) {
        any = true;
        const u64 key = make_key(w, j);
        min_key(minv, a, key, guard);
        min_key(minv, b, key, guard);
      }
    }
  }
  wave_flag(any, go);
}

__global__ __launch_bounds__(kLadBlock) void k_lad_vjoin(int nodes, const int* __restrict__ nindex,
                                                         const int* __restrict__ nlist, const int* __restrict__ eweight,
                                                         int* parent, const u64* __restrict__ minv,
                                                         uint8_t* __restrict__ inMST, int T, int heavy, int one_dir,
                                                         int split)
{
  const long long vl = (long long)blockIdx.x * kLadBlock + threadIdx.x;
  if (vl >= nodes) return;
  const int v = (int)vl;
  const int beg = nindex[v], end = nindex[v + 1];
  for (int j = beg; j < end; j++) {
    const int n = nlist[j];
    const int w = eweight[j];
    if (!in_phase(v, n, w, T, heavy, one_dir)) continue;
    const int a = lad_find(v, parent, split);
    const int b = lad_find(n, parent, split);
    if (a == b) continue;
    const u64 key = make_key(w, j);
    if (key == minv[a] || key == minv[b]) {
      hook(a, b, parent);
      inMST[j] = 1;
    }
  }
}

__global__ __launch_bounds__(kLadBlock) void k_lad_vreset(int nodes, u64* __restrict__ minv)
{
  const long long v = (long long)blockIdx.x * kLadBlock + threadIdx.x;
  if (v < nodes) minv[v] = kNoKey;
}

int hip_err(emst_ctx* ctx, hipError_t e, const char* what)
{
  char buf[256];
  std::snprintf(buf, sizeof(buf), "HIP error %d in %s: %s", (int)e, what, hipGetErrorString(e));
  return ctx_fail(ctx, EMST_E_HIP, buf);
}

#define LADCHK(call)                                      \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return hip_err(ctx, e_, #call); \
  } while (0)

int grow(emst_ctx* ctx, void** p, size_t* cap, size_t bytes, const char* what)
{
  if (*p && *cap >= bytes) return EMST_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) {
    char buf[128];
    std::snprintf(buf, sizeof(buf), "hipMalloc of %zu bytes for the %s failed", bytes, what);
    return ctx_fail(ctx, EMST_E_NOMEM, buf);
  }
  *cap = bytes;
  return EMST_OK;
}

inline unsigned blocks_for(long long n) { return (unsigned)std::max<long long>(1, (n + kLadBlock - 1) / kLadBlock); }

template <bool TUPLES>
int ladder_phases(emst_ctx* ctx, const LadderMode& md, int32_t nodes, int32_t edges, const int* nindex,
                  const int* nlist, const int* eweight, int nphases, int T, uint8_t* inMST, hipStream_t s,
                  long long stat[ST_COUNT])
{
  int* d_cnt = ctx->d_lad_cnt;  // [list size, flag]
  void* lists[2] = {ctx->d_lad[0], ctx->d_lad[1]};
  int* slist = md.edge_based && !TUPLES ? ctx->d_slist : nullptr;
  const bool no_list = md.edge_based && !md.data_driven && !TUPLES;  // revisit every CSR position
  // entries hold round-start roots only with implicit PC on 16-byte records
  const bool refind = !md.implicit_pc || !TUPLES;
  const bool split = !md.implicit_pc;
  for (int phase = 0; phase < nphases; phase++) {
    const bool heavy = phase == 1;
    const int rounds_slot = heavy ? ST_ROUNDS1 : ST_ROUNDS0;
    if (!md.edge_based) {
      // vertex-based rounds: no worklist
      for (;;) {
        LADCHK(hipMemsetAsync(&d_cnt[1], 0, 4, s));
        hipLaunchKernelGGL(k_lad_vselect, dim3(blocks_for(nodes)), dim3(kLadBlock), 0, s, nodes, nindex, nlist,
                           eweight, ctx->d_parent, ctx->d_minv, T, (int)heavy, (int)md.one_dir, (int)md.guard,
                           (int)split, &d_cnt[1]);
        int go = 0;
        LADCHK(hipMemcpyAsync(&go, &d_cnt[1], 4, hipMemcpyDeviceToHost, s));
        LADCHK(hipStreamSynchronize(s));
        stat[rounds_slot]++;
        stat[ST_IN_SUM] += edges;
        if (!go) break;
        hipLaunchKernelGGL(k_lad_vjoin, dim3(blocks_for(nodes)), dim3(kLadBlock), 0, s, nodes, nindex, nlist, eweight,
                           ctx->d_parent, ctx->d_minv, inMST, T, (int)heavy, (int)md.one_dir, (int)split);
        hipLaunchKernelGGL(k_lad_vreset, dim3(blocks_for(nodes)), dim3(kLadBlock), 0, s, nodes, ctx->d_minv);
        LADCHK(hipGetLastError());
      }
      continue;
    }
    // initWL; the light phase also records every position's source when
    // the worklist holds edge ids (with no list, that is all it does)
    if (!(no_list && heavy)) {
      LADCHK(hipMemsetAsync(&d_cnt[0], 0, 4, s));
      hipLaunchKernelGGL(k_lad_init<TUPLES>, dim3(blocks_for(nodes)), dim3(kLadBlock), 0, s, nodes, nindex, nlist,
                         eweight, ctx->d_parent, T, (int)heavy, (int)md.one_dir, (int)md.warp,
                         heavy ? (int*)nullptr : slist, no_list ? nullptr : lists[0], &d_cnt[0]);
      LADCHK(hipGetLastError());
    }
    long long size = edges;
    if (!no_list) {
      int sz = 0;
      LADCHK(hipMemcpyAsync(&sz, &d_cnt[0], 4, hipMemcpyDeviceToHost, s));
      LADCHK(hipStreamSynchronize(s));
      size = sz;
      stat[heavy ? ST_HEAVY : ST_LIGHT] = sz;
    }
    int cur = 0;
    while (size > 0) {
      void* in = no_list ? nullptr : lists[cur];
      void* out = md.data_driven ? lists[cur ^ 1] : nullptr;
      LADCHK(hipMemsetAsync(md.data_driven ? &d_cnt[0] : &d_cnt[1], 0, 4, s));
      hipLaunchKernelGGL(k_lad_select<TUPLES>, dim3(blocks_for(size)), dim3(kLadBlock), 0, s, in, size, out,
                         &d_cnt[0], &d_cnt[1], ctx->d_parent, ctx->d_minv, slist, nlist, eweight, T, (int)heavy,
                         (int)md.one_dir, (int)md.guard, (int)md.implicit_pc);
      int next = 0;
      LADCHK(hipMemcpyAsync(&next, md.data_driven ? &d_cnt[0] : &d_cnt[1], 4, hipMemcpyDeviceToHost, s));
      LADCHK(hipStreamSynchronize(s));
      stat[rounds_slot]++;
      stat[ST_IN_SUM] += size;
      if (md.data_driven) {
        stat[ST_OUT_SUM] += next;
        cur ^= 1;
        size = next;
        in = lists[cur];
      }
      if (next == 0) break;
      hipLaunchKernelGGL(k_lad_join<TUPLES>, dim3(blocks_for(size)), dim3(kLadBlock), 0, s, in, size, ctx->d_parent,
                         ctx->d_minv, inMST, slist, nlist, eweight, T, (int)heavy, (int)md.one_dir, (int)refind,
                         (int)split);
      hipLaunchKernelGGL(k_lad_reset<TUPLES>, dim3(blocks_for(size)), dim3(kLadBlock), 0, s, in, size, ctx->d_parent,
                         ctx->d_minv, slist, nlist, eweight, T, (int)heavy, (int)md.one_dir, (int)refind, (int)split);
      LADCHK(hipGetLastError());
    }
  }
  return EMST_OK;
}

}  // namespace

namespace emst {

int run_ladder(emst_ctx* ctx, int32_t nodes, int32_t edges, const int* d_nindex, const int* d_nlist,
               const int* d_eweight, int use_filter, int32_t threshold, uint8_t* d_inMST, hipStream_t s,
               long long wl_cap, emst_stats* stats)
{
  const int f = ctx->cfg.flags;
  LadderMode md;
  md.guard = !(f & EMST_F_NO_GUARD);
  md.warp = !(f & EMST_F_NO_WARP);
  md.implicit_pc = !(f & EMST_F_NO_IMPLICIT_PC);
  md.one_dir = !(f & EMST_F_NO_ONE_DIRECTION);
  md.tuples = !(f & EMST_F_NO_TUPLES);
  md.data_driven = !(f & EMST_F_NO_DATA_DRIVEN);
  md.edge_based = !(f & EMST_F_NO_EDGE_BASED);
  const bool filter = use_filter == EMST_FILTER_PLAIN;
  const int T = filter ? threshold : INT_MAX;
  const int nphases = filter ? 2 : 1;

  // buffers: two worklists (records or ids), the per-position sources, and
  // the size/flag words
  const long long recs = md.one_dir ? std::min<long long>(std::max<long long>(wl_cap, 0), edges) : edges;
  const size_t rec_bytes = md.tuples ? sizeof(int4) : sizeof(int);
  int rc = ensure_buffers(ctx, nodes, edges, 1);
  if (rc) return rc;
  if (md.edge_based && !(md.tuples == false && md.data_driven == false))
    for (int i = 0; i < 2; i++)
      if ((rc = grow(ctx, &ctx->d_lad[i], &ctx->cap_lad[i], (size_t)recs * rec_bytes, "ladder worklist"))) return rc;
  if (md.edge_based && !md.tuples)
    if ((rc = grow(ctx, (void**)&ctx->d_slist, &ctx->cap_slist, (size_t)edges * 4, "per-edge source array"))) return rc;
  if ((rc = grow(ctx, (void**)&ctx->d_lad_cnt, &ctx->cap_lad_cnt, 16, "ladder counters"))) return rc;

  LADCHK(hipEventRecord(ctx->ev[0], s));
  hipLaunchKernelGGL(k_lad_setup, dim3(blocks_for(nodes)), dim3(kLadBlock), 0, s, nodes, ctx->d_parent, ctx->d_minv);
  if (edges > 0) LADCHK(hipMemsetAsync(d_inMST, 0, (size_t)edges, s));
  long long stat[ST_COUNT] = {};
  rc = md.tuples ? ladder_phases<true>(ctx, md, nodes, edges, d_nindex, d_nlist, d_eweight, nphases, T, d_inMST, s, stat)
                 : ladder_phases<false>(ctx, md, nodes, edges, d_nindex, d_nlist, d_eweight, nphases, T, d_inMST, s, stat);
  if (rc) return rc;
  LADCHK(hipEventRecord(ctx->ev[1], s));
  LADCHK(hipEventSynchronize(ctx->ev[1]));
  LADCHK(hipGetLastError());
  if (stats) {
    float t = 0;
    LADCHK(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]));
    std::memset(stats, 0, sizeof(*stats));
    stats->ms = t;
    stats->threshold_used = T;
    stats->filter_used = filter ? 1 : 0;
    stats->learned_q = -1.f;
    stats->light = stat[ST_LIGHT];
    stats->heavy = stat[ST_HEAVY];
    stats->light_rounds = (int32_t)stat[ST_ROUNDS0];
    stats->heavy_rounds = (int32_t)stat[ST_ROUNDS1];
    stats->round_in_sum = stat[ST_IN_SUM];
    stats->round_out_sum = stat[ST_OUT_SUM];
  }
  return EMST_OK;
}

}  // namespace emst
