// emst_solver.hip -- MI355X (gfx950) minimum spanning forest solver with the
// semantics of ECL-MST's gpuMST path (ECL-MST/ECL_MST.cu:127-362).
//
// Device flow of one solve (the reference's timed region, ECL_MST.cu:301-345):
//   k_setup            parent[v]=v, minv[v]=max, per-task start vertices
//                      (initPM, ECL_MST.cu:147-154); inMST cleared by k_zero
//   k_init_wl<light>   records {v, u, w, j} of upper-direction edges with w<=T
//                      (initWL<true>, ECL_MST.cu:157-207)
//   k_select<first>    the phase's first kernel1 round: every endpoint is a
//                      root and every record crosses, so it only offers keys
//   rounds             k_join: the record holding a root's minimum key
//                      clears that minimum, hooks, and marks inMST[j]
//                      (kernel2 + kernel3, ECL_MST.cu:229-251, fused)
//                      k_select: drop records inside one component, write
//                      back roots, offer keys (kernel1, ECL_MST.cu:210-226)
//   k_pick_big, k_flatten  parent[v] = root(v); class map (big component, roots)
//   k_init_wl<heavy>   heavy upper-direction edges whose roots differ
//                      (initWL<false>, ECL_MST.cu:157-207, launch :326)
//   rounds             as above
//
// The host does not read the worklist size back every round (the reference
// does, ECL_MST.cu:318).  Round kernels run as fixed grid-stride grids that
// read sizes from device memory; k_join posts the count it joins into
// host-mapped memory and the host stops enqueueing rounds when it reads 0,
// a couple of rounds behind the device.  When a list fits cfg.tail_max,
// workgroup 0 of k_join finishes the whole phase with workgroup barriers.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/emst.h"
#include "emst_internal.hpp"
#include "emst_kernels.hpp"
#include "host_util.hpp"
#include "learned.hpp"

using namespace emst;

namespace {

constexpr int kInitWaves = 4;     // waves per workgroup in k_init_wl
constexpr int kMarkSpan = 256;   // longest vertex range of a task resolved by marks
constexpr int kStartEdges = 256;  // granularity of the task start table (the smallest task)
constexpr int kSparseSub = 4;     // 256-edge tasks in the sparse heavy initialisation
constexpr int kHflagSlots = 512;  // host-visible round results per solve
constexpr int kSelItems = 4;      // records per thread per k_select tile
constexpr int kTbl = 1024;        // LDS min-combining table entries
constexpr int kPair = 1024;       // LDS parallel-record table entries (k_select after the first round)
constexpr int kReportWords = ST_COUNT + 5;  // counters, threshold, filter flag, quantile bits, sparse heavy, light offers

// Dynamic LDS of an initialisation workgroup: per-wave staging and marks,
// then the workgroup's fill count (16 bytes keep the size a multiple of 16).
size_t init_lds_bytes(int stage_recs, int task_edges)
{
  return (size_t)kInitWaves * stage_recs * sizeof(int4) + (size_t)kInitWaves * task_edges * sizeof(int) + 16;
}

// ---------------------------------------------------------------------------
// k_setup: initPM (ECL_MST.cu:147-154) plus the task start table: for every
// kStartEdges-edge block, the vertex v whose adjacency range holds the
// block's first edge writes it (a task of k blocks reads every k-th entry).
// Zero-degree vertices own no edge and write nothing.  A vertex that starts
// more than 16 blocks (a hub) has its entries written by its whole wave.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_setup(int nodes, const int* __restrict__ nindex,
                                               int* __restrict__ parent, u64* __restrict__ minv,
                                               int* __restrict__ gsrc, int* __restrict__ cnt,
                                               long long* __restrict__ stats, unsigned char* __restrict__ hasmem)
{
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v == 0) {
    cnt[0] = 0;
    cnt[1] = 0;
    cnt[2] = 0;
    cnt[3] = 0;  // the sparse heavy decision, set by k_pick_big
    cnt[4] = 0;  // the light initialisation's choice, set by k_pick_light
    for (int i = 0; i < ST_COUNT; i++) stats[i] = 0;
  }
  int g0 = 0, g1 = 0;
  if (v < nodes) {
    parent[v] = (int)v;
    minv[v] = kNoKey;
    if (hasmem) hasmem[v] = 0;
    const int beg = nindex[v], end = nindex[v + 1];
    if (end > beg) {
      g0 = (int)(((long long)beg + kStartEdges - 1) / kStartEdges);
      g1 = (int)(((long long)end + kStartEdges - 1) / kStartEdges);
    }
  }
  const bool hub = g1 - g0 > 16;
  if (!hub)
    for (int g = g0; g < g1; g++) gsrc[g] = (int)v;
  for (u64 m = __ballot(hub); m; m &= m - 1) {
    const int l = __builtin_ctzll(m);
    const int a = __shfl(g0, l), b = __shfl(g1, l), hv = __shfl((int)v, l);
    for (int g = a + lane_id(); g < b; g += kWave) gsrc[g] = hv;
  }
}

// Clears `bytes` bytes: an unaligned head and tail byte by byte, the
// 16-byte-aligned middle with 16-byte stores (graph-capture safe, unlike a
// memset node).
__global__ __launch_bounds__(256) void k_zero(uint8_t* __restrict__ p, long long bytes)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long head = (long long)((16 - ((uintptr_t)p & 15)) & 15);
  if (head > bytes) head = bytes;
  const long long n16 = (bytes - head) >> 4;
  const long long tail0 = head + (n16 << 4);
  if (t < head) p[t] = 0;
  if (t < bytes - tail0) p[tail0 + t] = 0;
  int4* q = reinterpret_cast<int4*>(p + head);
  for (long long i = t; i < n16; i += stride) st_stream(&q[i], make_int4(0, 0, 0, 0));
}

// ---------------------------------------------------------------------------
// k_init_wl: streams the CSR and emits worklist records.
//   HEAVY=false: {v, u, w, j} for u > v and w <= T      (ECL_MST.cu:170-176)
//   HEAVY=true : {root(v), root(u), w, j} for u > v with different roots
//                (ECL_MST.cu:166, 173-174); parent is flat here.  The
//                reference also tests w > T (ECL_MST.cu:197); after the
//                light phase every light edge lies inside one component, so
//                "roots differ" already implies w > T and the weight is
//                read only for crossing edges.  The vertex class map (two
//                bits per vertex, 4 MB for 2^24 vertices, mostly L2
//                resident) marks the members of one large component and the
//                roots, whose root is known without touching parent.
// One wave task covers SUB*64 consecutive CSR positions.  Lane i takes
// positions base+i, base+64+i, ... so every load is one coalesced 256-byte
// request, and all SUB loads of a task are issued before the first use,
// one task ahead of their use.
// The source vertex of a position, src(p) = max{v : nindex[v] <= p}, comes
// from marks: every vertex whose adjacency starts inside the task writes
// its id at that offset in a per-wave LDS array, and a DPP max-scan over the
// offsets fills the rest.  Tasks whose vertex range exceeds kMarkSpan (runs
// of zero-degree vertices) search windows of 64 boundaries instead.  This
// edge-centric split replaces the reference's per-vertex thread/warp split
// at degree 4 (ECL_MST.cu:167, 185): every wave streams whole 256-byte
// lines whatever the degrees, and hub vertices spread over many waves.
// T comes from the argument, from dthr (decided on the device), or per
// source vertex from gthr[vgid[v]] when several graphs are solved together.
// Output: workgroup b appends to its own segment wl[b*seg_cap ...], reserving
// ranges with an LDS counter, and finally writes its record count to
// segcnt[b] and adds it to cnt[0].  seg_cap bounds the upper-direction edges
// of the workgroup's tasks, so a segment never overflows.
// ---------------------------------------------------------------------------
// Vertex boundaries of a task, loaded before the next task's stream loads
// are issued (so waiting for them does not wait for those): up to kMarkSpan
// boundaries one per lane and row when the task's vertex range is short,
// else a window of 64 for the search below.
struct TaskBounds {
  int stv[kMarkSpan / kWave];
  int bnd;
  bool by_marks;
};

__device__ __forceinline__ void task_bounds_load(TaskBounds& tb, const int* __restrict__ nindex, int nodes, int base,
                                                 int vend, int lane)
{
  tb.by_marks = vend - base <= kMarkSpan;
  tb.bnd = 0;
  if (tb.by_marks) {
#pragma unroll
    for (int q = 0; q < kMarkSpan / kWave; q++) {
      const long long v = (long long)base + 1 + q * kWave + lane;
      tb.stv[q] = (q * kWave < vend - base && v <= vend) ? nindex[v] : INT_MAX;
    }
  } else {
    const long long bi = (long long)base + 1 + lane;
    tb.bnd = bi <= nodes ? nindex[bi] : INT_MAX;
  }
}

// Source vertex of every position of the task starting at gbeg whose first
// position belongs to vertex base: src(p) = max{v : nindex[v] <= p}.
template <int SUB>
__device__ __forceinline__ void task_sources(int (&src)[SUB], TaskBounds& tb, long long gbeg, int base, int edges,
                                             int nodes, const int* __restrict__ nindex, int* marks, int lane)
{
  constexpr int TE = SUB * kWave;
  if (tb.by_marks) {
    // Every vertex v in (base, vend] whose adjacency starts inside the
    // task writes v at its start offset (the largest v wins where
    // zero-degree vertices share a start); a prefix maximum over the
    // task's offsets then gives the source of every position.
#pragma unroll
    for (int u = 0; u < SUB; u++) marks[u * kWave + lane] = -1;
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < kMarkSpan / kWave; q++) {
      const long long off = (long long)tb.stv[q] - gbeg;
      if (off > 0 && off < TE) atomicMax(&marks[off], base + 1 + q * kWave + lane);
    }
    wave_lds_sync();
    // prefix maximum in position order: a wave max-scan per row of 64
    // positions, carried from row to row
    int carry = base;
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      const int x = max(wave_max_scan(marks[u * kWave + lane]), carry);
      src[u] = x;
      carry = __builtin_amdgcn_readlane(x, kWave - 1);
    }
    wave_lds_sync();  // the next task rewrites the marks
  } else {
    // long vertex ranges (runs of zero-degree vertices): search windows
    // of 64 vertex boundaries held one per lane
    int bnd = tb.bnd;
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      const long long pl = gbeg + u * kWave + lane;
      const int p = pl < edges ? (int)pl : INT_MAX - 1;
      bool done = pl >= edges;
      int sv = 0;
      for (;;) {
        const int c = count_le(bnd, p);
        const int b63 = __shfl(bnd, kWave - 1);
        if (!done && p < b63) {
          sv = base + c;
          done = true;
        }
        if (__all(done)) break;
        base += kWave;
        const long long bi = (long long)base + 1 + lane;
        bnd = bi <= nodes ? nindex[bi] : INT_MAX;
      }
      src[u] = sv;
    }
  }
}

// Ends an initialisation workgroup: flushes the wave's staging, then posts
// the segment's record count and adds it to the list size and the counter.
__device__ __forceinline__ void init_finish(Stage& st, int4* seg, int& s_fill, int* segcnt, int* cnt,
                                            long long* stats, int stat)
{
  stage_flush(st, seg, &s_fill);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int f = s_fill;
    segcnt[blockIdx.x] = f;
    if (f) {
      atomicAdd(cnt, f);
      atomicAdd((unsigned long long*)&stats[stat], (unsigned long long)f);
    }
  }
}

template <bool HEAVY, int SUB, bool UNION>
__global__ __launch_bounds__(kInitWaves * kWave) void k_init_wl(
    const int* __restrict__ nindex, const int* __restrict__ nlist, const int* __restrict__ eweight,
    const int* __restrict__ gsrc, int nodes, int edges, int T, const int* __restrict__ label,
    const unsigned* __restrict__ vclass, int4* __restrict__ wl, long long seg_cap, int* __restrict__ segcnt,
    int* __restrict__ cnt, int stage_cap, long long* __restrict__ stats, const int* __restrict__ dthr,
    const int* __restrict__ vgid, const int* __restrict__ gthr, const int* __restrict__ bigroot_p,
    const int* __restrict__ alt_ran)
{
  constexpr int TE = SUB * kWave;
  // dynamic LDS: per-wave staging, per-wave marks, the workgroup's fill count
  extern __shared__ int4 lds_stage[];
  int& s_fill = *(reinterpret_cast<int*>(lds_stage + kInitWaves * stage_cap) + kInitWaves * TE);
  if (dthr) {
    if (HEAVY && dthr[1] == 0) return;  // no heavy phase when unfiltered
    T = dthr[0];
  }
  // the sparse pass was enqueued too and the marked component chose it
  if (alt_ran && *alt_ran) return;
  if (threadIdx.x == 0) s_fill = 0;
  __syncthreads();
  int4* const seg = wl + (long long)blockIdx.x * seg_cap;
  const int bigroot = HEAVY ? *bigroot_p : -1;
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  Stage st{lds_stage + wid * stage_cap, stage_cap, 0};
  int* marks = reinterpret_cast<int*>(lds_stage + kInitWaves * stage_cap) + wid * TE;
  const long long ngroups = ((long long)edges + TE - 1) / TE;
  const long long nwaves = (long long)gridDim.x * kInitWaves;
  constexpr int R = TE / kStartEdges;  // start table entries per task
  // The stream loads of a wave's next task are issued before the current
  // task is processed (after the current task's boundary window, so waiting
  // for that window does not wait for them).
  long long g = (long long)blockIdx.x * kInitWaves + wid;
  int nl[SUB], w[SUB], base = 0, vend = 0;
  if (g < ngroups) {
    base = gsrc[g * R];
    vend = g + 1 < ngroups ? gsrc[(g + 1) * R] : nodes - 1;
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      const long long p = g * TE + u * kWave + lane;
      nl[u] = p < edges ? ld_stream(&nlist[p]) : -1;
      if (!HEAVY) w[u] = p < edges ? ld_stream(&eweight[p]) : 0;
    }
  }
  for (; g < ngroups; g += nwaves) {
    const long long gbeg = g * TE;
    TaskBounds tb;
    task_bounds_load(tb, nindex, nodes, base, vend, lane);
    const long long gn = g + nwaves;
    int nl_n[SUB], w_n[SUB], base_n = 0, vend_n = 0;
    if (gn < ngroups) {
      base_n = gsrc[gn * R];
      vend_n = gn + 1 < ngroups ? gsrc[(gn + 1) * R] : nodes - 1;
#pragma unroll
      for (int u = 0; u < SUB; u++) {
        const long long p = gn * TE + u * kWave + lane;
        nl_n[u] = p < edges ? ld_stream(&nlist[p]) : -1;
        if (!HEAVY) w_n[u] = p < edges ? ld_stream(&eweight[p]) : 0;
      }
    }
    // A union's graphs are packed by vertex id, so a task whose first and
    // last sources lie in one graph has that graph's threshold (scalar
    // loads).  In a task that spans two graphs the light pass gathers it per
    // position, and the heavy pass tests every upper-direction position for
    // crossing (a light edge never crosses, so only the gathers are lost).
    int tT = T;
    bool span = false;
    if (UNION) {
      const int g0 = vgid[__builtin_amdgcn_readfirstlane(base)];
      span = g0 != vgid[__builtin_amdgcn_readfirstlane(vend)];
      tT = HEAVY && span ? INT_MIN : gthr[g0];
    }
    int src[SUB];
    task_sources<SUB>(src, tb, gbeg, base, edges, nodes, nindex, marks, lane);
    if (!HEAVY) {
#pragma unroll
      for (int u = 0; u < SUB; u++) {
        const int j = (int)(gbeg + u * kWave + lane);
        const bool up = nl[u] > src[u];
        const int t = (UNION && span && up) ? gthr[vgid[src[u]]] : tT;
        const bool keep = up && w[u] <= t;
        stage_push(st, keep, make_int4(src[u], nl[u], w[u], j), seg, &s_fill);
      }
    } else {
      // Every light edge (w <= T; a union's graph's own T) lies inside one
      // component after the light phase, so only the upper-direction
      // positions with w > T (the reference's own test, ECL_MST.cu:197)
      // need roots.  Their
      // weights are loaded with the source classes (positions of a task are
      // consecutive, so the loads are coalesced), and the class and root
      // gathers are skipped for the light ones.  The class map resolves
      // most roots without touching parent: a member of the marked
      // component has root bigroot, a root is its own root; only the
      // remaining endpoints (inside other non-trivial components) are
      // looked up.
      bool hv[SUB];
      int a[SUB], b[SUB], wt[SUB];
      unsigned cv[SUB];
#pragma unroll
      for (int u = 0; u < SUB; u++) {
        const long long p = gbeg + u * kWave + lane;
        const bool up = nl[u] > src[u];
        wt[u] = up ? eweight[p] : 0;
        cv[u] = up ? (vclass[(unsigned)src[u] >> 4] >> ((src[u] & 15) * 2)) & 3u : 0u;
      }
#pragma unroll
      for (int u = 0; u < SUB; u++) {
        hv[u] = nl[u] > src[u] && wt[u] > tT;
        a[u] = bigroot;
        b[u] = bigroot;
        if (hv[u]) {
          const unsigned cu = (vclass[(unsigned)nl[u] >> 4] >> ((nl[u] & 15) * 2)) & 3u;
          a[u] = (cv[u] & 1u) ? bigroot : (cv[u] & 2u) ? src[u] : -1;
          b[u] = (cu & 1u) ? bigroot : (cu & 2u) ? nl[u] : -1;
        }
      }
#pragma unroll
      for (int u = 0; u < SUB; u++) {
        if (a[u] < 0) a[u] = label[src[u]];
        if (b[u] < 0) b[u] = label[nl[u]];
      }
#pragma unroll
      for (int u = 0; u < SUB; u++) {
        const long long p = gbeg + u * kWave + lane;
        const bool keep = hv[u] && a[u] != b[u];
        stage_push(st, keep, make_int4(a[u], b[u], wt[u], (int)p), seg, &s_fill);
      }
    }
    base = base_n;
    vend = vend_n;
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      nl[u] = nl_n[u];
      if (!HEAVY) w[u] = w_n[u];
    }
  }
  init_finish(st, seg, s_fill, segcnt, cnt, stats, HEAVY ? ST_HEAVY : ST_LIGHT);
}

// ---------------------------------------------------------------------------
// Sparse heavy initialisation (the records of k_init_wl<true>, for sorted,
// symmetric lists with reverse positions), fused with the heavy phase's
// first round of offers.  A crossing edge has at least one endpoint outside
// the marked component M, so only the positions of sources outside M are
// examined, and a task whose source vertices all lie in M is skipped without
// streaming it (skip[g], from k_task_skip).  For a source v outside M and a
// position p of v with neighbour y:
//   y > v (p is the upper copy): emitted when root(v) != root(y), as in
//       k_init_wl<true>;
//   y < v, y in M (p is the lower copy of an edge crossing out of M): the
//       edge's upper copy is rev[p] in y's list, so the record is
//       {root(y), root(v), eweight[rev[p]], rev[p]};
//   y < v, y outside M: not emitted; y's own walk emits that edge's upper
//       copy.
// Every crossing upper-direction edge is thus emitted exactly once, the
// reference's initWL<false> set (ECL_MST.cu:157-207, launch :326).  With
// mirrored weights eweight[p] equals eweight[rev[p]] and is streamed with
// the task.
// The first round's offers (kernel1's atomicMin, ECL_MST.cu:222-223) are
// made here, so the offer-only select of that round is skipped: v's list
// holds every edge incident to v (one copy of each, either direction), so
// the minimum key over v's crossing positions, keyed by their upper copy, is
// v's offer.  Per task it is reduced in LDS slots (one per source vertex);
// the position holding the slot's minimum writes it to minv[v] with a plain
// store when v is a component of its own (a root without members,
// hasmem[v] == 0) whose list lies inside the task, else with a guarded
// atomicMin to minv[root(v)].  M's offer is the minimum over the records
// that touch M, reduced per wave and per workgroup, then one atomicMin per
// workgroup.  A record carries candidate flags in bit 31 of x and y for the
// first join: the source side is a candidate only if its key equals the
// task's minimum for that source (it cannot be the component's minimum
// otherwise); the other side is always a candidate.
// The wave's tasks are g0, g0 + nwaves, ...; their skip flags are read 64 at
// a time (one per lane) and only the tasks to do are visited, the next one's
// stream loads issued before the current one is processed.
// ---------------------------------------------------------------------------
constexpr int kCandBit = INT_MIN;  // bit 31 of a record endpoint: candidate of the first join

// Dynamic LDS of the sparse heavy initialisation: the initialisation layout,
// then per-wave minimum slots for the sources of a task.
size_t sparse_lds_bytes(int stage_recs)
{
  return init_lds_bytes(stage_recs, kSparseSub * kWave) + (size_t)kInitWaves * (kMarkSpan + 1) * sizeof(u64);
}

template <int SUB>
__global__ __launch_bounds__(kInitWaves * kWave) __attribute__((amdgpu_waves_per_eu(5))) void k_init_heavy_sparse(
    const int* __restrict__ nindex, const int* __restrict__ nlist, const int* __restrict__ eweight,
    const int* __restrict__ gsrc, int nodes, int edges, const int* __restrict__ label,
    const unsigned* __restrict__ vclass, const unsigned char* __restrict__ skip, const int* __restrict__ rev,
    int mirrored, int kidmode, const unsigned char* __restrict__ hasmem, u64* __restrict__ minv, int4* __restrict__ wl,
    long long seg_cap, int* __restrict__ segcnt, int* __restrict__ cnt, int stage_cap, long long* __restrict__ stats,
    const int* __restrict__ dthr, const int* __restrict__ bigroot_p)
{
  constexpr int TE = SUB * kWave;
  constexpr int NSLOT = kMarkSpan + 1;
  extern __shared__ int4 lds_stage[];
  int* const tail = reinterpret_cast<int*>(lds_stage + kInitWaves * stage_cap) + kInitWaves * TE;
  int& s_fill = tail[0];
  u64& s_bigmin = *reinterpret_cast<u64*>(tail + 2);
  if (dthr && dthr[1] == 0) return;  // no heavy phase when unfiltered
  if (!bigroot_p[1]) return;         // the full pass runs instead
  if (threadIdx.x == 0) {
    s_fill = 0;
    s_bigmin = kNoKey;
  }
  __syncthreads();
  int4* const seg = wl + (long long)blockIdx.x * seg_cap;
  const int bigroot = *bigroot_p;
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  Stage st{lds_stage + wid * stage_cap, stage_cap, 0};
  int* marks = reinterpret_cast<int*>(lds_stage + kInitWaves * stage_cap) + wid * TE;
  u64* slot = reinterpret_cast<u64*>(tail + 4) + wid * NSLOT;
  const long long ngroups = ((long long)edges + TE - 1) / TE;
  const long long nwaves = (long long)gridDim.x * kInitWaves;
  const long long g0 = (long long)blockIdx.x * kInitWaves + wid;
  // the wave's tasks to do, 64 at a time: bit l of todo is task
  // g0 + (64 * (chunk - 1) + l) * nwaves
  long long chunk = 0;
  u64 todo = 0;
  auto next_task = [&]() -> long long {
    for (;;) {
      if (todo) {
        const int l = __builtin_ctzll(todo);
        todo &= todo - 1;
        return g0 + (64 * (chunk - 1) + l) * nwaves;
      }
      if (g0 + 64 * chunk * nwaves >= ngroups) return -1;
      const long long gl = g0 + (64 * chunk + lane) * nwaves;
      todo = __ballot(gl < ngroups && !skip[gl]);
      chunk++;
    }
  };
  auto load_task = [&](long long t, int (&nl)[SUB], int& base, int& vend) {
    static_assert(TE == kStartEdges, "sparse tasks are start-table blocks");
    base = gsrc[t];
    vend = t + 1 < ngroups ? gsrc[t + 1] : nodes - 1;
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      const long long p = t * TE + u * kWave + lane;
      nl[u] = p < edges ? ld_stream(&nlist[p]) : -1;
    }
  };
  long long g = next_task();
  int nl[SUB], base = 0, vend = 0;
  if (g >= 0) load_task(g, nl, base, vend);
  while (g >= 0) {
    const long long gbeg = g * TE;
    TaskBounds tb;
    task_bounds_load(tb, nindex, nodes, base, vend, lane);
    const long long gn = next_task();
    int nl_n[SUB], base_n = 0, vend_n = 0;
    if (gn >= 0) load_task(gn, nl_n, base_n, vend_n);
    // mirrored weights are streamed with the task (used at the keys)
    int w[SUB];
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      const long long p = gbeg + u * kWave + lane;
      w[u] = (mirrored && p < edges) ? ld_stream(&eweight[p]) : 0;
    }
    int src[SUB];
    task_sources<SUB>(src, tb, gbeg, base, edges, nodes, nindex, marks, lane);
    const int pbeg = (int)gbeg;  // positions fit in int (edges < 2^31)
    // roots of both endpoints of the positions whose source lies outside M;
    // a position crosses when they differ
    int a[SUB], b[SUB];
    bool cross[SUB], tom[SUB], lower[SUB], alone[SUB];
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      const int p = pbeg + u * kWave + lane;
      lower[u] = nl[u] < src[u];
      const unsigned cv = (vclass[(unsigned)src[u] >> 4] >> ((src[u] & 15) * 2)) & 3u;
      const bool act = p < edges && !(cv & 1u) && nl[u] != src[u];
      const unsigned cu = act ? (vclass[(unsigned)nl[u] >> 4] >> ((nl[u] & 15) * 2)) & 3u : 0u;
      tom[u] = cu & 1u;
      a[u] = (cv & 2u) ? src[u] : -1;
      b[u] = (cu & 1u) ? bigroot : (cu & 2u) ? nl[u] : -1;
      // a root source without other members is a component of its own
      alone[u] = act && a[u] >= 0 && !hasmem[src[u]];
      if (act && a[u] < 0) a[u] = label[src[u]];
      if (act && b[u] < 0) b[u] = label[nl[u]];
      cross[u] = act && a[u] != b[u];
    }
    // keys ((w, j) of the edge's upper-direction copy)
    int j[SUB], wv[SUB];
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      const int p = pbeg + u * kWave + lane;
      j[u] = (cross[u] && (lower[u] || kidmode)) ? rev[p] : p;
      wv[u] = (cross[u] && !mirrored) ? eweight[j[u]] : w[u];
    }
    // each source's minimum over this task: LDS slots by vertex (vertex
    // ranges up to kMarkSpan), else straight to minv
    const bool slots = tb.by_marks;
    if (slots) {
      for (int i = lane; i <= vend - base; i += kWave) slot[i] = kNoKey;
      wave_lds_sync();
#pragma unroll
      for (int u = 0; u < SUB; u++)
        if (cross[u]) atomicMin(&slot[src[u] - base], make_key(wv[u], j[u]));
      wave_lds_sync();
    } else {
#pragma unroll
      for (int u = 0; u < SUB; u++)
        if (cross[u]) min_key(minv, a[u], make_key(wv[u], j[u]), true);
    }
    // M's minimum: over the crossing positions whose neighbour is in M
    {
      u64 bigk = kNoKey;
#pragma unroll
      for (int u = 0; u < SUB; u++) {
        const u64 key = make_key(wv[u], j[u]);
        if (cross[u] && tom[u] && key < bigk) bigk = key;
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const u64 o = __shfl_xor(bigk, off);
        bigk = o < bigk ? o : bigk;
      }
      if (lane == 0 && bigk != kNoKey) atomicMin(&s_bigmin, bigk);
    }
    // records, with the first join's candidate flags.  The position that
    // holds its source's minimum over the task (one per source: keys are
    // distinct) writes it: a plain store for a component of its own whose
    // list lies inside the task (the only contributor), else a guarded
    // atomicMin to the source's root.
    const int first_src = __builtin_amdgcn_readfirstlane(src[0]);
    const int last_src = __builtin_amdgcn_readlane(src[SUB - 1], kWave - 1);
    const bool whole = pbeg + TE <= edges;
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      const bool keep = cross[u] && (!lower[u] || tom[u]);
      const u64 key = make_key(wv[u], j[u]);
      const bool cand = !slots || key == slot[src[u] - base];
      if (slots && cross[u] && cand) {
        if (alone[u] && whole && src[u] != first_src && src[u] != last_src) minv[src[u]] = key;
        else min_key(minv, a[u], key, true);
      }
      const int xa = cand ? (a[u] | kCandBit) : a[u];
      // records list the lower endpoint's root first, as k_init_wl<true>
      const int4 rec = lower[u] ? make_int4(b[u] | kCandBit, xa, wv[u], j[u]) : make_int4(xa, b[u] | kCandBit, wv[u], j[u]);
      stage_push(st, keep, rec, seg, &s_fill);
    }
    if (slots) wave_lds_sync();  // the next task resets the slots
    g = gn;
    base = base_n;
    vend = vend_n;
#pragma unroll
    for (int u = 0; u < SUB; u++) nl[u] = nl_n[u];
  }
  init_finish(st, seg, s_fill, segcnt, cnt, stats, ST_HEAVY);
  if (threadIdx.x == 0 && s_bigmin != kNoKey) min_key(minv, bigroot, s_bigmin, true);
}

// ---------------------------------------------------------------------------
// Light initialisation fused with the light phase's first round of offers,
// for sorted, symmetric lists with reverse positions and mirrored weights.
// The records are k_init_wl<false>'s: {v, u, w, p} for every upper-direction
// position p of v (u = nlist[p] > v) with w <= T (ECL_MST.cu:170-176).  When
// the light phase starts every vertex is a component of its own, and v's
// list holds every edge incident to v, so v's first offer (kernel1's
// atomicMin, ECL_MST.cu:222-223) is the minimum key over v's light
// positions, each keyed by its upper-direction copy: (w, p) for an upper
// position, (w, rev[p]) for a lower one (mirrored weights: both copies carry
// w, so the light test needs no reverse position).  The minimum is reduced
// per task in LDS slots, one per source vertex, and the position holding it
// writes minv[v]: a plain store when v's whole list lies inside the task,
// else a guarded atomicMin.
// Records carry the first join's candidate flags in bit 31 of their
// endpoints, as in the sparse heavy pass: v's side only if the key is the
// task's minimum for v, u's side always.  Tasks are kStartEdges positions,
// wave w of workgroup b takes tasks b*kInitWaves + w + k*nwaves, and the
// next task's stream loads (lists, weights and reverse positions: light
// edges are dense where this pass runs) are issued before the current one is
// processed.
// Runs when *decide is set (k_pick_light); k_init_wl<false> runs otherwise.
// ---------------------------------------------------------------------------
template <int SUB, bool UNION>
__global__ __launch_bounds__(kInitWaves * kWave) __attribute__((amdgpu_waves_per_eu(5))) void k_init_light_offers(
    const int* __restrict__ nindex, const int* __restrict__ nlist, const int* __restrict__ eweight,
    const int* __restrict__ gsrc, int nodes, int edges, int T, const int* __restrict__ rev, int kidmode,
    u64* __restrict__ minv,
    int4* __restrict__ wl, long long seg_cap, int* __restrict__ segcnt, int* __restrict__ cnt, int stage_cap,
    long long* __restrict__ stats, const int* __restrict__ dthr, const int* __restrict__ decide,
    const int* __restrict__ vgid, const int* __restrict__ gthr)
{
  constexpr int TE = SUB * kWave;
  static_assert(TE == kStartEdges, "light offer tasks are start-table blocks");
  constexpr int NSLOT = kMarkSpan + 1;
  extern __shared__ int4 lds_stage[];
  int* const tail = reinterpret_cast<int*>(lds_stage + kInitWaves * stage_cap) + kInitWaves * TE;
  int& s_fill = tail[0];
  if (!*decide) return;
  if (dthr) T = dthr[0];
  if (threadIdx.x == 0) s_fill = 0;
  __syncthreads();
  int4* const seg = wl + (long long)blockIdx.x * seg_cap;
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  Stage st{lds_stage + wid * stage_cap, stage_cap, 0};
  int* marks = reinterpret_cast<int*>(lds_stage + kInitWaves * stage_cap) + wid * TE;
  u64* slot = reinterpret_cast<u64*>(tail + 4) + wid * NSLOT;
  const long long ngroups = ((long long)edges + TE - 1) / TE;
  const long long nwaves = (long long)gridDim.x * kInitWaves;
  auto load_task = [&](long long t, int (&nl_t)[SUB], int (&w_t)[SUB], int (&r_t)[SUB], int& b, int& e) {
    b = gsrc[t];
    e = t + 1 < ngroups ? gsrc[t + 1] : nodes - 1;
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      const long long p = t * TE + u * kWave + lane;
      nl_t[u] = p < edges ? ld_stream(&nlist[p]) : -1;
      w_t[u] = p < edges ? ld_stream(&eweight[p]) : 0;
      r_t[u] = p < edges ? ld_stream(&rev[p]) : 0;
    }
  };
  long long g = (long long)blockIdx.x * kInitWaves + wid;
  int nl[SUB], w[SUB], rv[SUB], base = 0, vend = 0;
  if (g < ngroups) load_task(g, nl, w, rv, base, vend);
  for (; g < ngroups; g += nwaves) {
    const long long gbeg = g * TE;
    TaskBounds tb;
    task_bounds_load(tb, nindex, nodes, base, vend, lane);
    const long long gn = g + nwaves;
    int nl_n[SUB], w_n[SUB], rv_n[SUB], base_n = 0, vend_n = 0;
    if (gn < ngroups) load_task(gn, nl_n, w_n, rv_n, base_n, vend_n);
    // a union's threshold: the task's graph's (scalar loads), per position
    // in a task that spans two graphs (its graphs are packed by vertex id)
    int tT = T;
    bool span = false;
    if (UNION) {
      const int g0 = vgid[__builtin_amdgcn_readfirstlane(base)];
      span = g0 != vgid[__builtin_amdgcn_readfirstlane(vend)];
      tT = gthr[g0];
    }
    int src[SUB];
    task_sources<SUB>(src, tb, gbeg, base, edges, nodes, nindex, marks, lane);
    const int pbeg = (int)gbeg;  // positions fit in int (edges < 2^31)
    // light positions and their keys
    bool light[SUB], lower[SUB];
    int j[SUB];
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      const int p = pbeg + u * kWave + lane;
      lower[u] = nl[u] < src[u];
      const int t = (UNION && span && p < edges) ? gthr[vgid[src[u]]] : tT;
      light[u] = p < edges && nl[u] != src[u] && w[u] <= t;
      j[u] = (light[u] && (lower[u] || kidmode)) ? rv[u] : p;
    }
    // each source's minimum over this task: LDS slots by vertex (vertex
    // ranges up to kMarkSpan), else straight to minv
    const bool slots = tb.by_marks;
    if (slots) {
      for (int i = lane; i <= vend - base; i += kWave) slot[i] = kNoKey;
      wave_lds_sync();
#pragma unroll
      for (int u = 0; u < SUB; u++)
        if (light[u]) atomicMin(&slot[src[u] - base], make_key(w[u], j[u]));
      wave_lds_sync();
    } else {
#pragma unroll
      for (int u = 0; u < SUB; u++)
        if (light[u]) min_key(minv, src[u], make_key(w[u], j[u]), true);
    }
    // records with candidate flags; the position holding its source's
    // minimum (one per source: keys are distinct) writes it
    const int first_src = __builtin_amdgcn_readfirstlane(src[0]);
    const int last_src = __builtin_amdgcn_readlane(src[SUB - 1], kWave - 1);
    const bool whole = pbeg + TE <= edges;
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      const u64 key = make_key(w[u], j[u]);
      const bool cand = !slots || key == slot[src[u] - base];
      if (slots && light[u] && cand) {
        if (whole && src[u] != first_src && src[u] != last_src) minv[src[u]] = key;
        else min_key(minv, src[u], key, true);
      }
      const int x = cand ? (src[u] | kCandBit) : src[u];
      stage_push(st, light[u] && !lower[u], make_int4(x, nl[u] | kCandBit, w[u], j[u]), seg, &s_fill);
    }
    if (slots) wave_lds_sync();  // the next task resets the slots
    base = base_n;
    vend = vend_n;
#pragma unroll
    for (int u = 0; u < SUB; u++) {
      nl[u] = nl_n[u];
      w[u] = w_n[u];
      rv[u] = rv_n[u];
    }
  }
  init_finish(st, seg, s_fill, segcnt, cnt, stats, ST_LIGHT);
}

// Chooses between the two light initialisations of a graph with reverse
// positions: the pass that also makes the first offers spends more per
// position and saves the first select, which costs per light record, so it
// pays when light edges are a large share of the adjacency.  *out = 1 when
// at least a tenth of 1024 evenly spaced positions are light (w <= T).
__global__ __launch_bounds__(64) void k_pick_light(int edges, const int* __restrict__ eweight, int T,
                                                   const int* __restrict__ dthr, int* __restrict__ out, int force)
{
  const int lane = lane_id();
  if (edges <= 0 || force) {
    if (lane == 0) *out = edges > 0;
    return;
  }
  // (every sampled position is below edges, so the loads need no branch)
  int wv[16];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const long long i = k * kWave + lane;
    wv[k] = eweight[i * edges / 1024 + edges / 2048];
  }
  if (dthr) T = dthr[0];
  int c = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) c += wv[k] <= T;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
  if (lane == 0) *out = 10 * c >= 1024;
}

// Skip flags of the sparse heavy initialisation: task g (positions
// g*kStartEdges ...) is skipped when every vertex in [gsrc[g], gsrc[g+1]]
// (a superset of its sources) lies in the marked component (bit 0 of its
// class).  Ranges longer than 4096 vertices are not examined and not
// skipped.
__global__ __launch_bounds__(256) void k_task_skip(long long ngroups, int nodes, const int* __restrict__ gsrc,
                                                   const unsigned* __restrict__ vclass, unsigned char* __restrict__ skip,
                                                   const int* __restrict__ dthr, const int* __restrict__ sparse_on)
{
  if ((dthr && dthr[1] == 0) || !*sparse_on) return;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const int lo = gsrc[g];
  const int hi = g + 1 < ngroups ? gsrc[g + 1] : nodes - 1;
  bool all = hi - lo <= 4096;
  for (int v = lo; all && v <= hi; v++) all = (vclass[(unsigned)v >> 4] >> ((v & 15) * 2)) & 1u;
  skip[g] = all ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Reverse positions, built once per loaded graph (emst_load): for a
// lower-direction position p of vertex x (y = nlist[p] < x), rev[p] is the
// position of the same edge's upper-direction copy in y's list: the k-th copy
// of y in x's list pairs with the k-th copy of x in y's list, found by a
// binary search (lists must be sorted).  bad[0] counts positions out of
// order, bad[1] lower copies without such a partner, bad[2] partners of
// another weight.  With no position out of order, every lower copy
// partnered and as many lower as upper copies, the pairing is a bijection
// and the sparse heavy initialisation is exact.  starts[t] is the source
// of position t*kRevTask, which narrows the search for a position's source.
// ---------------------------------------------------------------------------
constexpr int kRevTask = 256;

__global__ __launch_bounds__(256) void k_rev_starts(int nodes, const int* __restrict__ nindex, int* __restrict__ starts)
{
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nodes) return;
  const int beg = nindex[v], end = nindex[v + 1];
  for (long long t = ((long long)beg + kRevTask - 1) / kRevTask; t < ((long long)end + kRevTask - 1) / kRevTask; t++)
    starts[t] = (int)v;
}

__global__ __launch_bounds__(256) void k_rev(int nodes, int edges, const int* __restrict__ nindex,
                                             const int* __restrict__ nlist, const int* __restrict__ eweight,
                                             const int* __restrict__ starts, int* __restrict__ rev,
                                             unsigned long long* __restrict__ bad)
{
  const long long ntasks = ((long long)edges + kRevTask - 1) / kRevTask;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < edges; p += stride) {
    const long long t = p / kRevTask;
    int lo = starts[t], hi = t + 1 < ntasks ? starts[t + 1] : nodes - 1;
    while (lo < hi) {  // the source: the largest v with nindex[v] <= p
      const int mid = lo + (hi - lo + 1) / 2;
      if (nindex[mid] <= p) lo = mid;
      else hi = mid - 1;
    }
    const int x = lo, y = nlist[p], first = nindex[x];
    if (p > first && nlist[p - 1] > y) atomicAdd(&bad[0], 1ull);
    if (y >= x) continue;
    int k = 0;
    for (long long q = p - 1; q >= first && nlist[q] == y; q--) k++;
    int a = nindex[y], b = nindex[y + 1];
    const int yend = b;
    while (a < b) {  // first position of x in y's list
      const int mid = a + (b - a) / 2;
      if (nlist[mid] < x) a = mid + 1;
      else b = mid;
    }
    const long long r = (long long)a + k;
    if (r < yend && nlist[r] == x) {
      rev[p] = (int)r;
      if (eweight[r] != eweight[p]) atomicAdd(&bad[2], 1ull);
    } else {
      rev[p] = -1;
      atomicAdd(&bad[1], 1ull);
    }
  }
}

// Picks the component that the heavy pass' class map marks: the most frequent
// root among the targets of 64 evenly spaced positions (one wave), so the
// vote weighs components by their adjacency (vertices when there are no
// edges).  Any choice is correct; one holding most positions saves the most
// root lookups, and lets the sparse heavy initialisation skip the most.
// out[0] = its root; out[1] = 1 when it holds at least half of the sampled
// positions, which selects the sparse heavy initialisation where it is
// available (it skips the component's own adjacency, so it pays only when
// that adjacency dominates).
__global__ __launch_bounds__(64) void k_pick_big(int nodes, int edges, const int* __restrict__ nlist,
                                                 const int* __restrict__ parent, int* __restrict__ out,
                                                 const int* __restrict__ dthr)
{
  if (dthr && dthr[1] == 0) return;
  const int lane = lane_id();
  int v;
  if (edges > 0) {
    const long long p = ((long long)lane * edges) / kWave + (long long)edges / (2 * kWave);
    v = nlist[p < edges ? p : edges - 1];
  } else {
    v = (int)(((long long)lane * nodes) / kWave + ((long long)nodes / (2 * kWave)));
  }
  const int r = find_root(v < nodes ? v : nodes - 1, parent);
  int c = 0;
  for (int k = 0; k < kWave; k++) c += __shfl(r, k) == r;
  int best = (c << 8) | (kWave - 1 - lane);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) best = max(best, __shfl_xor(best, off));
  const int winner = kWave - 1 - (best & 0xff);
  const int big = __shfl(r, winner);
  if (lane == 0) {
    out[0] = big;
    out[1] = edges > 0 && 2 * (best >> 8) >= kWave;
  }
}

// Spreads the 16 low bits of x to the even bit positions of a word.
__device__ __forceinline__ unsigned spread16(unsigned x)
{
  x &= 0xffffu;
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  return (x | (x << 1)) & 0x55555555u;
}

// parent[v] = root(v), and the vertex class map: two bits per vertex,
// sixteen vertices per word, bit 0 = (root(v) == *big), bit 1 = (v is a
// root).  Concurrent writers only shorten chains toward the same root, so
// readers stay correct.  With hasmem (cleared by k_setup) and the sparse
// heavy pass chosen (big[1]), its only reader, every root outside the marked
// component that has other members gets hasmem[root] = 1.
__global__ __launch_bounds__(256) void k_flatten(int nodes, int* __restrict__ parent, int* __restrict__ cnt,
                                                 const int* __restrict__ big, unsigned* __restrict__ vclass,
                                                 const int* __restrict__ dthr, unsigned char* __restrict__ hasmem)
{
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v == 0) {
    cnt[0] = 0;
    cnt[1] = 0;
  }
  if (dthr && dthr[1] == 0) return;  // unfiltered: no heavy phase
  const int bigroot = big[0];
  // member flags only for the sparse heavy pass (big[1], set by k_pick_big)
  unsigned char* const hm = big[1] ? hasmem : nullptr;
  int r = -1;
  if (v < nodes) {
    r = find_root((int)v, parent);
    parent[v] = r;
    if (hm && r != (int)v && r != bigroot) hm[r] = 1;
  }
  const u64 mb = __ballot(r == bigroot && r >= 0);
  const u64 mr = __ballot(r >= 0 && r == (int)v);
  const int lane = lane_id();
  const long long v0 = v - lane;
  const long long words = ((long long)nodes + 15) / 16;
  const long long wi = (v0 >> 4) + lane;
  if (lane < 4 && v0 < nodes && wi < words)
    vclass[wi] = spread16((unsigned)(mb >> (16 * lane))) | (spread16((unsigned)(mr >> (16 * lane))) << 1);
}

// Offers key to rep: min-combined in the workgroup's LDS table when the rep
// holds (or claims) a slot within 4 probes, else a guarded global atomicMin.
// With CAND, returns whether the offer can still be rep's minimum: the
// workgroup table held no smaller key, or the guarded global read showed
// none.  The winner of every rep is among these candidates (keys are
// distinct).
template <bool CAND>
__device__ __forceinline__ bool offer_key(int* tags, u64* keys, u64* minv, int rep, u64 key, bool guard,
                                          bool combine)
{
  if (combine) {
    const unsigned h = ((unsigned)rep * 0x9E3779B1u) >> (32 - 10);
#pragma unroll
    for (int pr = 0; pr < 4; pr++) {
      const unsigned sl = (h + pr) & (kTbl - 1);
      int t = tags[sl];
      if (t == -1) {
        t = atomicCAS(&tags[sl], -1, rep);
        if (t == -1) t = rep;
      }
      if (t == rep) {
        if (!CAND) {
          atomicMin(&keys[sl], key);
          return true;
        }
        // a read before the atomic decides candidacy: a stale (larger)
        // value only makes a record a candidate it need not be, and the
        // non-returning LDS atomic keeps hot slots fast
        const bool c = keys[sl] > key;
        if (c) atomicMin(&keys[sl], key);
        return c;
      }
    }
  }
  return min_key(minv, rep, key, guard);
}

// ---------------------------------------------------------------------------
// k_select (kernel1, ECL_MST.cu:210-226): for each record find both roots;
// records inside one component are dropped, the rest are written back with
// their roots (implicit path compression) and offer their key to both roots.
//
// Two MI355X-specific changes, both invisible in the results:
//  * Survivors of a WAVES*256-record tile are compacted with a workgroup
//    scan and reserved with one atomicAdd per tile (a per-record or per-wave
//    counter atomic serialises on one address: ~12 ns each on MI355X).
//  * Keys are first min-combined per workgroup in an LDS table keyed by
//    root (4-probe open addressing, claimed slots are never released).  A hub
//    component that appears in most records is claimed early and then
//    costs one global atomicMin per workgroup instead of one per record.
// FIRST: the phase's first round; every endpoint is a root and every record
// crosses (both guaranteed by the initialisation), so finds and the copy
// are skipped and the list is joined in place.
// ---------------------------------------------------------------------------
// Parallel records: after the first round, two records whose endpoints
// have the same pair of roots connect the same two components, and of the
// two only the lighter can be in the forest (it closes a cycle with the
// heavier, which is that cycle's maximum).  Each tile therefore keeps, per
// root pair, only its lightest record; records whose pair finds no LDS slot
// within 8 probes are kept.  The joins of every round are unchanged (each
// component's minimum record survives), so rounds and results are those of
// the reference schedule; only the worklist sizes shrink.  With dedup = 0
// every crossing record is kept and the sizes are the reference's.
//
// Segmented input: a phase's first select and join, and the select after
// that join, read the initialisation's output, which is nseg segments of
// seg_cap records each (segment s holds segcnt[s] records at its start).
// Positions are physical (s*seg_cap + offset); seg_cap is a multiple of the
// select tile, so a tile never straddles two segments, and the candidate and
// joined-record bitmasks are indexed by physical position as well.  With
// segcnt == nullptr the list is dense: cnt_in records from position 0.
struct ListView {
  const int* segcnt;
  long long seg_cap;
  int nseg;
};

// Walks the tiles of a list that belong to this workgroup and calls
// fn(base, lim) for each: base is the tile's first physical position, lim
// the end of its segment's records.  A segmented list hands whole segments
// to workgroups (one count load per segment and no division per tile; with
// more workgroups than segments, a segment's tiles are split among several);
// a dense list is the one-segment case, walked grid-stride.  The walk is
// uniform across the workgroup.
template <class F>
__device__ __forceinline__ void walk_tiles(const ListView& lv, long long n_dense, long long tile, F&& fn)
{
  const int P = gridDim.x, w = blockIdx.x;
  const int nseg = lv.segcnt ? lv.nseg : 1;
  int s0, sstep;
  long long j0, jstep;
  if (P <= nseg) {
    s0 = w;
    sstep = P;
    j0 = 0;
    jstep = 1;
  } else {
    const int k = P / nseg;
    s0 = w < k * nseg ? w % nseg : nseg;
    sstep = nseg;
    j0 = w / nseg;
    jstep = k;
  }
  for (int sg = s0; sg < nseg; sg += sstep) {
    const long long c = lv.segcnt ? (long long)lv.segcnt[sg] : n_dense;
    const long long sb = (long long)sg * lv.seg_cap;
    for (long long j = j0; j * tile < c; j += jstep) fn(sb + j * tile, sb + c);
  }
}

// Physical positions a list spans (the size of its bitmasks).
__device__ __forceinline__ long long list_span(const ListView& lv, long long n_dense)
{
  return lv.segcnt ? (long long)lv.nseg * lv.seg_cap : n_dense;
}

__device__ __forceinline__ int pair_slot(u64* tags, u64* keys, u64 pair, u64 key)
{
  const unsigned h = ((unsigned)(pair ^ (pair >> 29)) * 0x9E3779B1u) >> (32 - 10);
#pragma unroll
  for (int pr = 0; pr < 8; pr++) {
    const unsigned sl = (h + pr) & (kPair - 1);
    u64 t = tags[sl];
    if (t == kNoKey) {
      t = atomicCAS(&tags[sl], kNoKey, pair);
      if (t == kNoKey) t = pair;
    }
    if (t == pair) {
      atomicMin(&keys[sl], key);
      return (int)sl;
    }
  }
  return -1;
}

template <bool FIRST, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void k_select(
    const int4* __restrict__ wl_in, const int* __restrict__ cnt_in, int4* __restrict__ wl_out,
    int* __restrict__ cnt_out, const int* __restrict__ parent, u64* __restrict__ minv, int guard, int combine,
    int dedup, long long* __restrict__ stats, int rounds_slot, u64* __restrict__ cand, const u64* __restrict__ dead,
    ListView lv, const int* __restrict__ fused, const int* __restrict__ kid)
{
  constexpr int BT = WAVES * kWave;
  constexpr int TILE = BT * kSelItems;
  __shared__ int s_tags[kTbl];
  __shared__ u64 s_keys[kTbl];
  __shared__ int s_wtot[WAVES];
  __shared__ int s_woff[WAVES];
  __shared__ int s_base;
  __shared__ u64 s_ptag[FIRST ? 1 : kPair];
  __shared__ u64 s_pkey[FIRST ? 1 : kPair];
  if (FIRST && fused && *fused) return;  // the initialisation made this round's offers
  const int n_in = *cnt_in;
  if (n_in == 0) return;
  const long long span = list_span(lv, n_in);
  const long long cand_words = (span + kWave - 1) / kWave;
  const int tid = threadIdx.x;
  if (blockIdx.x == 0 && tid == 0) {
    stats[rounds_slot] += 1;
    stats[ST_IN_SUM] += n_in;
  }
  for (int i = tid; i < kTbl; i += BT) {
    s_tags[i] = -1;
    s_keys[i] = kNoKey;
  }
  if (!FIRST && dedup)
    for (int i = tid; i < kPair; i += BT) {
      s_ptag[i] = kNoKey;
      s_pkey[i] = kNoKey;
    }
  __syncthreads();
  const int lane = lane_id();
  const int wid = tid / kWave;
  walk_tiles(lv, n_in, TILE, [&](const long long base, const long long lim) {
    int4 r[kSelItems];
    bool live[kSelItems];
#pragma unroll
    for (int k = 0; k < kSelItems; k++) {
      const long long i = base + k * BT + tid;
      // records joined by the first round are known to lie inside one tree
      live[k] = i < lim && !(dead && ((dead[(base + k * BT) / kWave + wid] >> lane) & 1ull));
      r[k] = live[k] ? wl_in[i] : make_int4(0, 0, 0, 0);
    }
    if (FIRST && kid) {
      // a relabelled layout's initialisation wrote positions: the records
      // take their key ids here (the list is joined in place)
#pragma unroll
      for (int k = 0; k < kSelItems; k++)
        if (live[k]) r[k].w = kid[r[k].w];
#pragma unroll
      for (int k = 0; k < kSelItems; k++)
        if (live[k]) reinterpret_cast<int*>(const_cast<int4*>(wl_in) + base + k * BT + tid)[3] = r[k].w;
    }
    bool keep[kSelItems];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < kSelItems; k++) {
      const long long i = base + k * BT + tid;
      keep[k] = false;
      bool cx = false, cy = false;
      if (live[k]) {
        // (bit 31 of an endpoint is a first-join candidate flag)
        const int a = FIRST ? r[k].x : find_root(r[k].x & INT_MAX, parent);
        const int c = FIRST ? r[k].y : find_root(r[k].y & INT_MAX, parent);
        if (a != c) {
          keep[k] = true;
          const u64 key = make_key(r[k].z, r[k].w);
          cx = offer_key<FIRST>(s_tags, s_keys, minv, a, key, guard, combine);
          cy = offer_key<FIRST>(s_tags, s_keys, minv, c, key, guard, combine);
          r[k].x = a;
          r[k].y = c;
        }
      }
      if (FIRST) {
        // the list is joined in place: candidate bits go to two bitmasks,
        // one 64-bit word per wavefront row of records
        const u64 mx = __ballot(cx), my = __ballot(cy);
        const long long word = (base + k * BT) / kWave + wid;
        if (lane == 0 && word < cand_words) {
          cand[word] = mx;
          cand[cand_words + word] = my;
        }
      }
      mine += keep[k];
    }
    if (FIRST) return;
    if (dedup) {
      // keep the lightest record of each root pair in this tile
      int slot[kSelItems];
#pragma unroll
      for (int k = 0; k < kSelItems; k++) {
        slot[k] = -1;
        if (keep[k]) {
          const unsigned lo = (unsigned)min(r[k].x, r[k].y), hi = (unsigned)max(r[k].x, r[k].y);
          slot[k] = pair_slot(s_ptag, s_pkey, ((u64)lo << 32) | hi, make_key(r[k].z, r[k].w));
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kSelItems; k++)
        if (slot[k] >= 0 && s_pkey[slot[k]] != make_key(r[k].z, r[k].w)) {
          keep[k] = false;
          mine--;
        }
      __syncthreads();
      // clear the slots this tile used (the scan below has a barrier
      // before the next tile's inserts)
#pragma unroll
      for (int k = 0; k < kSelItems; k++)
        if (slot[k] >= 0) {
          s_ptag[slot[k]] = kNoKey;
          s_pkey[slot[k]] = kNoKey;
        }
    }
    // workgroup exclusive scan of the per-thread survivor counts
    int x = mine;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == kWave - 1) s_wtot[wid] = x;
    __syncthreads();
    if (tid == 0) {
      int acc = 0;
      for (int w = 0; w < WAVES; w++) {
        s_woff[w] = acc;
        acc += s_wtot[w];
      }
      s_base = acc ? atomicAdd(cnt_out, acc) : 0;
    }
    __syncthreads();
    long long pos = (long long)s_base + s_woff[wid] + (x - mine);
#pragma unroll
    for (int k = 0; k < kSelItems; k++)
      if (keep[k]) wl_out[pos++] = r[k];
    __syncthreads();
  });
  // publish the workgroup's combined minima (after every offer of every
  // wave: the first-round pass has no per-tile barrier)
  __syncthreads();
  for (int i = tid; i < kTbl; i += BT) {
    const int rep = s_tags[i];
    if (rep >= 0) min_key(minv, rep, s_keys[i], guard);
  }
}

// One workgroup finishes a phase: join the current list, then repeat
// select/join until no record crosses components.  All shared state is
// accessed at device scope; workgroup barriers order the steps.
__device__ void tail_rounds(int4* cur, int n, int4* nxt, int* parent, u64* minv, uint8_t* inMST,
                            long long* stats, int rounds_slot)
{
  __shared__ int s_cnt;
  int rounds = 0;
  long long in_sum = 0, out_sum = 0;
  for (;;) {
    // join (kernel2 + kernel3)
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const int4 r = cur[i];
      const int x = r.x, y = r.y;
      const u64 key = make_key(r.z, r.w);
      const u64 kx = __hip_atomic_load(&minv[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const u64 ky = __hip_atomic_load(&minv[y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const bool wx = key == kx, wy = key == ky;
      if (wx) __hip_atomic_store(&minv[x], kNoKey, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (wy) __hip_atomic_store(&minv[y], kNoKey, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (wx || wy) {
        hook(x, y, parent);
        inMST[r.w] = 1;
      }
    }
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    // select (kernel1)
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const int4 r = cur[i];
      const int a = find_root_coherent(r.x, parent);
      const int c = find_root_coherent(r.y, parent);
      if (a != c) {
        const u64 key = make_key(r.z, r.w);
        const int pos = atomicAdd(&s_cnt, 1);
        nxt[pos] = make_int4(a, c, r.z, r.w);
        min_key(minv, a, key, true);
        min_key(minv, c, key, true);
      }
    }
    __syncthreads();
    const int s = s_cnt;
    rounds++;
    in_sum += n;
    out_sum += s;
    __syncthreads();
    if (s == 0) break;
    int4* t = cur;
    cur = nxt;
    nxt = t;
    n = s;
  }
  if (threadIdx.x == 0) {
    stats[rounds_slot] += rounds;
    stats[ST_TAIL] += rounds;
    stats[ST_IN_SUM] += in_sum;
    stats[ST_OUT_SUM] += out_sum;
  }
}

// ---------------------------------------------------------------------------
// k_join (kernel2 + kernel3, ECL_MST.cu:229-251): the record whose key
// equals minv of one of its roots is that root's minimum edge: it resets
// that minv (each root has exactly one such record, so the reset fuses
// kernel3 away), hooks the two roots and marks inMST[j].  A reader racing
// with a reset sees either the winner's key or the sentinel, neither equal
// to its own key, so the outcome is the same as with a separate kernel3.
// The phase's first join checks only candidate endpoints: those of the
// candidate bitmasks written by the offer-only select, or, when *fused is
// set (the initialisation made the offers), those flagged in bit 31 of the
// record's endpoints; it then also counts the round that the skipped select
// would have counted.
// ---------------------------------------------------------------------------
template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void k_join(
    int4* __restrict__ wl, int* __restrict__ cnt, int4* __restrict__ wl_other, int* __restrict__ cnt_other,
    int* __restrict__ parent, u64* __restrict__ minv, uint8_t* __restrict__ inMST, int tail_max,
    int* __restrict__ hflag, long long* __restrict__ stats, int rounds_slot, const u64* cand, u64* joined_mask,
    ListView lv, const int* __restrict__ fused)
{
  constexpr int BT = WAVES * kWave;
  const int n = *cnt;
  const bool flags = fused && *fused;
  if (!lv.segcnt && !flags && n > 0 && n <= tail_max) {
    if (blockIdx.x != 0) return;
    if (threadIdx.x == 0) {
      stats[ST_OUT_SUM] += n;
      __hip_atomic_store(hflag, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    tail_rounds(wl, n, wl_other, parent, minv, inMST, stats, rounds_slot);
    if (threadIdx.x == 0) {
      *cnt = 0;
      *cnt_other = 0;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *cnt_other = 0;
    stats[ST_OUT_SUM] += n;
    if (flags && n > 0) {
      stats[rounds_slot] += 1;
      stats[ST_IN_SUM] += n;
    }
    __hip_atomic_store(hflag, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  if (n == 0) return;  // (segment counts are only valid for a non-empty list)
  const long long span = list_span(lv, n);
  const long long cand_words = (span + kWave - 1) / kWave;
  walk_tiles(lv, n, BT, [&](const long long base, const long long lim) {
    const long long i = base + threadIdx.x;
    const bool live = i < lim;
    // after the offer-only first round, only a candidate endpoint can have
    // this record as its minimum; later rounds check both endpoints
    bool cx = live, cy = live;
    if (cand && !flags && live) {
      cx = (cand[i / kWave] >> (i % kWave)) & 1ull;
      cy = (cand[cand_words + i / kWave] >> (i % kWave)) & 1ull;
    }
    bool joined = false;
    if (cx || cy) {
      const int4 r = wl[i];
      if (flags) {
        cx = r.x < 0;
        cy = r.y < 0;
      }
      const int x = r.x & INT_MAX, y = r.y & INT_MAX;
      const u64 key = make_key(r.z, r.w);
      const bool wx = cx && minv[x] == key;
      const bool wy = cy && minv[y] == key;
      if (wx) minv[x] = kNoKey;
      if (wy) minv[y] = kNoKey;
      if (wx || wy) {
        hook(x, y, parent);
        __builtin_nontemporal_store((uint8_t)1, &inMST[r.w]);  // a scattered single-use byte
        joined = true;
      }
    }
    // A joined record lies inside one tree from now on: its bit tells the
    // next select to drop it without finds.  In the first round the mask
    // overwrites the candidate word this wavefront alone has just read.
    // Rows that start at or past the limit hold no record and write no word.
    const u64 jm = __ballot(joined);
    if (lane_id() == 0 && i < lim) joined_mask[i / kWave] = jm;
  });
}

// Copies the solve's counters, threshold and quantile into host-mapped
// memory with system-scope stores, so the host reads them after the final
// event without further copies.
__global__ __launch_bounds__(64) void k_report(const long long* __restrict__ stats, const int* __restrict__ thr,
                                               const float* __restrict__ q, const int* __restrict__ sparse,
                                               const int* __restrict__ light, long long* __restrict__ out)
{
  const int t = threadIdx.x;
  long long v = 0;
  if (t < ST_COUNT) v = stats[t];
  else if (t < ST_COUNT + 2) v = thr ? thr[t - ST_COUNT] : 0;
  else if (t == ST_COUNT + 2) v = q ? (long long)__float_as_uint(q[0]) : 0;
  else if (t == ST_COUNT + 3) v = sparse ? *sparse : 0;
  else if (t == ST_COUNT + 4) v = light ? *light : 0;
  if (t < kReportWords) __hip_atomic_store(&out[t], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---- runtime selection of template instances -------------------------------
typedef void (*InitFn)(const int*, const int*, const int*, const int*, int, int, int, const int*,
                       const unsigned*, int4*, long long, int*, int*, int, long long*, const int*, const int*,
                       const int*, const int*, const int*);
typedef void (*SparseFn)(const int*, const int*, const int*, const int*, int, int, const int*, const unsigned*,
                         const unsigned char*, const int*, int, int, const unsigned char*, u64*, int4*, long long,
                         int*, int*, int, long long*, const int*, const int*);
typedef void (*LightFn)(const int*, const int*, const int*, const int*, int, int, int, const int*, int, u64*, int4*,
                        long long, int*, int*, int, long long*, const int*, const int*, const int*, const int*);
typedef void (*SelFn)(const int4*, const int*, int4*, int*, const int*, u64*, int, int, int, long long*, int, u64*,
                      const u64*, ListView, const int*, const int*);
typedef void (*JoinFn)(int4*, int*, int4*, int*, int*, u64*, uint8_t*, int, int*, long long*, int, const u64*, u64*,
                       ListView, const int*);

// uni: a union of graphs with a threshold per graph (gthr[vgid[v]])
template <bool UNION>
InitFn init_fn_t(bool heavy, int sub)
{
  if (heavy) {
    if (sub == 4) return k_init_wl<true, 4, UNION>;
    if (sub == 16) return k_init_wl<true, 16, UNION>;
    return k_init_wl<true, 8, UNION>;
  }
  if (sub == 4) return k_init_wl<false, 4, UNION>;
  if (sub == 16) return k_init_wl<false, 16, UNION>;
  return k_init_wl<false, 8, UNION>;
}

InitFn init_fn(bool heavy, int sub, bool uni)
{
  return uni ? init_fn_t<true>(heavy, sub) : init_fn_t<false>(heavy, sub);
}

SparseFn sparse_fn()
{
  return k_init_heavy_sparse<kSparseSub>;
}

// uni: a union of graphs with a threshold per graph
LightFn light_offers_fn(bool uni)
{
  return uni ? k_init_light_offers<kSparseSub, true> : k_init_light_offers<kSparseSub, false>;
}

SelFn sel_fn(bool first, int waves)
{
  if (first) {
    if (waves == 4) return k_select<true, 4>;
    if (waves == 16) return k_select<true, 16>;
    return k_select<true, 8>;
  }
  if (waves == 4) return k_select<false, 4>;
  if (waves == 16) return k_select<false, 16>;
  return k_select<false, 8>;
}

JoinFn join_fn(int waves)
{
  if (waves == 4) return k_join<4>;
  if (waves == 16) return k_join<16>;
  return k_join<8>;
}

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================
namespace emst {

int ctx_fail(emst_ctx* c, int code, const std::string& what)
{
  if (c) c->err = what;
  return code;
}

}  // namespace emst

namespace {

int hip_fail(emst_ctx* c, hipError_t e, const char* what)
{
  char buf[256];
  std::snprintf(buf, sizeof(buf), "HIP error %d in %s: %s", (int)e, what, hipGetErrorString(e));
  return ctx_fail(c, EMST_E_HIP, buf);
}

#define HIPCHK(call)                                       \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
  } while (0)

emst_config default_config()
{
  emst_config c{};
  c.block = 512;
  c.grid_per_cu = 4;
  c.tile_edges = 512;
  c.stage_recs = 256;
  c.tail_max = 4096;
  c.flags = 0;
  c.init_grid_per_cu = 5;
  return c;
}

}  // namespace

namespace emst {

// Element counts of one solve's work buffers (see WorkSizes).  The
// initialisation grid is at most one workgroup per four tasks; workgroup b
// runs tasks b*kInitWaves + wave + k*kInitWaves*grid, so its segment needs
// room for kInitWaves*ceil(tasks / (kInitWaves*grid)) tasks of tile_edges
// positions, rounded up to a whole select tile.
WorkSizes work_sizes(int cus, const emst_config& cfg, int32_t nodes, int32_t edges, long long wl_cap)
{
  WorkSizes z;
  z.nodes = std::max<long long>(nodes, 1);
  z.vclass_words = (z.nodes + 63) / 16;
  z.gsrc = ((long long)edges + 255) / 256 + 1;  // room for the smallest task size
  const long long te = cfg.tile_edges;
  const long long tasks = ((long long)edges + te - 1) / te;
  const long long grid_max = std::max(1, cus * std::max(1, cfg.init_grid_per_cu));
  const long long grid = std::max<long long>(1, std::min(grid_max, (tasks + kInitWaves - 1) / kInitWaves));
  const long long per = kInitWaves * ((tasks + kInitWaves * grid - 1) / (kInitWaves * grid));
  const long long tile = (long long)cfg.block * kSelItems;
  z.init_grid = (int)grid;
  z.seg_cap = std::max<long long>((per * te + tile - 1) / tile * tile, tile);
  z.segs = grid;
  z.wl[1] = std::max<long long>(wl_cap, 1);
  z.wl[0] = std::max<long long>(z.wl[1], grid * z.seg_cap);
  z.cand_words = (z.wl[0] + kWave - 1) / kWave;
  return z;
}

// The context's work buffers are managed in one function so their
// capacities stay in step; any reallocation invalidates captured schedules.
int ensure_buffers(emst_ctx* ctx, int32_t nodes, int32_t edges, long long wl_cap)
{
  const WorkSizes z = work_sizes(ctx->cus, ctx->cfg, nodes, edges, wl_cap);
  auto grow = [&](void** p, size_t* cap, size_t want, size_t elem, const char* what) -> int {
    if (*p && *cap >= want) return EMST_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(p, want * elem) != hipSuccess) {
      char buf[160];
      std::snprintf(buf, sizeof(buf), "hipMalloc of %zu bytes (%s) failed", want * elem, what);
      return ctx_fail(ctx, EMST_E_NOMEM, buf);
    }
    *cap = want;
    ctx->buf_gen++;
    return EMST_OK;
  };
  if ((size_t)z.nodes > ctx->cap_nodes || !ctx->d_parent) {
    for (void** p : {(void**)&ctx->d_parent, (void**)&ctx->d_minv, (void**)&ctx->d_vclass}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
    ctx->cap_nodes = 0;
    if (hipMalloc((void**)&ctx->d_parent, z.nodes * 4) != hipSuccess ||
        hipMalloc((void**)&ctx->d_minv, z.nodes * 8) != hipSuccess ||
        hipMalloc((void**)&ctx->d_vclass, z.vclass_words * 4) != hipSuccess)
      return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of parent/minv failed");
    ctx->cap_nodes = z.nodes;
    ctx->buf_gen++;
  }
  int rc;
  if ((rc = grow((void**)&ctx->d_gsrc, &ctx->cap_gsrc, z.gsrc, 4, "task start vertices"))) return rc;
  if ((rc = grow((void**)&ctx->d_wl[0], &ctx->cap_wl[0], z.wl[0], sizeof(int4), "worklist 0"))) return rc;
  if ((rc = grow((void**)&ctx->d_wl[1], &ctx->cap_wl[1], z.wl[1], sizeof(int4), "worklist 1"))) return rc;
  if ((rc = grow((void**)&ctx->d_cand, &ctx->cap_cand, 2 * z.cand_words, 8, "candidate bitmasks"))) return rc;
  if ((rc = grow((void**)&ctx->d_segcnt, &ctx->cap_seg, z.segs, 4, "segment counts"))) return rc;
  if ((rc = grow((void**)&ctx->d_skip, &ctx->cap_skip, z.gsrc, 1, "task skip flags"))) return rc;
  if ((rc = grow((void**)&ctx->d_hasmem, &ctx->cap_hasmem, z.nodes, 1, "root member flags"))) return rc;
  return EMST_OK;
}

int build_rev(emst_ctx* ctx, int32_t nodes, int32_t edges, const int* d_nindex, const int* d_nlist,
              const int* d_eweight, int* rev, hipStream_t s, bool* ok, bool* mirrored)
{
  *ok = false;
  *mirrored = false;
  if (edges <= 0) return EMST_OK;
  const long long ntasks = ((long long)edges + kRevTask - 1) / kRevTask;
  int* starts = nullptr;
  unsigned long long* bad = nullptr;
  if (hipMalloc((void**)&starts, (size_t)ntasks * 4) != hipSuccess) return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc failed");
  if (hipMalloc((void**)&bad, 3 * 8) != hipSuccess) {
    (void)hipFree(starts);
    return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc failed");
  }
  unsigned long long hb[3] = {1, 1, 1};
  bool fine = hipMemsetAsync(bad, 0, 3 * 8, s) == hipSuccess;
  if (fine) {
    hipLaunchKernelGGL(k_rev_starts, dim3((nodes + 255) / 256), dim3(256), 0, s, nodes, d_nindex, starts);
    const int grid = (int)std::min<long long>(((long long)edges + 255) / 256, (long long)ctx->cus * 32);
    hipLaunchKernelGGL(k_rev, dim3(grid), dim3(256), 0, s, nodes, edges, d_nindex, d_nlist, d_eweight, starts, rev,
                       bad);
    fine = hipGetLastError() == hipSuccess &&
           hipMemcpyAsync(hb, bad, sizeof(hb), hipMemcpyDeviceToHost, s) == hipSuccess &&
           hipStreamSynchronize(s) == hipSuccess;
  }
  (void)hipFree(starts);
  (void)hipFree(bad);
  if (!fine) return ctx_fail(ctx, EMST_E_HIP, "building reverse positions failed");
  *ok = hb[0] == 0 && hb[1] == 0;
  *mirrored = hb[2] == 0;
  return EMST_OK;
}

}  // namespace emst

namespace emst {

int run_solve(emst_ctx* ctx, int32_t nodes, int32_t edges, const int* d_nindex, const int* d_nlist,
              const int* d_eweight, int use_filter, int32_t threshold, uint8_t* d_inMST, hipStream_t s,
              long long wl_cap, const VertexThresholds* vt, emst_stats* stats, FixedSchedule* fixed,
              const WorkBufs* wbp, const RevIndex* rix, bool learned_cached)
{
  if (nodes < 1 || edges < 0) return ctx_fail(ctx, EMST_E_ARG, "node or edge count too low");
  if (use_filter < 0 || use_filter > EMST_FILTER_LEARNED) return ctx_fail(ctx, EMST_E_ARG, "bad filter mode");
  const emst_config& cfg = ctx->cfg;
  if (cfg.flags & kLadderFlags) {
    if (use_filter == EMST_FILTER_LEARNED || vt || fixed || wbp)
      return ctx_fail(ctx, EMST_E_ARG, "the ablation ladder takes one graph and a given threshold");
    const int rc = run_ladder(ctx, nodes, edges, d_nindex, d_nlist, d_eweight, use_filter, threshold, d_inMST, s,
                              wl_cap, stats);
    if (rc == EMST_OK) {
      ctx->last_nodes = nodes;
      ctx->last_parent = ctx->d_parent;
      ctx->last_relabelled = false;
    }
    return rc;
  }
  const int task_edges = cfg.tile_edges;
  const int sub = task_edges / kWave;
  const int waves = cfg.block / kWave;
  WorkBufs wb;
  if (wbp) {
    wb = *wbp;
  } else {
    int rc = ensure_buffers(ctx, nodes, edges, wl_cap);
    if (rc) return rc;
    wb.parent = ctx->d_parent;
    wb.minv = ctx->d_minv;
    wb.vclass = ctx->d_vclass;
    wb.gsrc = ctx->d_gsrc;
    wb.wl[0] = ctx->d_wl[0];
    wb.wl[1] = ctx->d_wl[1];
    wb.cand = ctx->d_cand;
    wb.cnt = ctx->d_cnt;
    wb.stats = ctx->d_stats;
    wb.segcnt = ctx->d_segcnt;
    wb.skip = ctx->d_skip;
    wb.hasmem = ctx->d_hasmem;
  }
  // the sparse heavy initialisation needs reverse positions, skip flags and
  // the member flags of the roots
  const bool sparse = rix && rix->rev && !vt && wb.skip && wb.hasmem && !(cfg.flags & EMST_F_NO_SPARSE_HEAVY) &&
                      sparse_lds_bytes(std::max(64, cfg.stage_recs)) <= ctx->lds_per_block;
  unsigned char* const hasmem = sparse ? wb.hasmem : nullptr;
  // the light initialisation that also makes the first offers needs every
  // vertex's list to hold all its edges with their upper copies' weights;
  // with it both light passes are enqueued and k_pick_light picks one
  const bool light_offers = rix && rix->rev && rix->mirrored && !(cfg.flags & EMST_F_NO_LIGHT_OFFERS) &&
                            sparse_lds_bytes(std::max(64, cfg.stage_recs)) <= ctx->lds_per_block;
  // a relabelled layout keys every position through rev (its key ids)
  const int kidmode = rix && rix->rev && rix->kid ? 1 : 0;
  if (kidmode && (vt || !rix->mirrored)) return ctx_fail(ctx, EMST_E_ARG, "key ids need one graph with mirrored weights");
  const WorkSizes z = work_sizes(ctx->cus, cfg, nodes, edges, wl_cap);
  const bool learned = use_filter == EMST_FILTER_LEARNED && !vt;
  // a learned threshold decided before the solve (cached per loaded graph)
  // runs as a plain solve at that threshold: the kernels read no device
  // decision
  const bool known = learned && learned_cached;
  const int T = use_filter == EMST_FILTER_PLAIN ? threshold
                                                : (known && ctx->learned_filter ? ctx->learned_T : INT_MAX);
  const int* dthr = learned && !known ? ctx->d_thr : nullptr;
  // Host-sync ablation (the reference reads the worklist size back after
  // every kernel1, ECL_MST.cu:318): the host waits for every round's
  // posted size before it enqueues the next round, and no round runs in the
  // one-workgroup tail.
  const bool host_sync = (cfg.flags & EMST_F_HOST_SYNC) != 0;
  if (host_sync && (learned || vt || fixed))
    return ctx_fail(ctx, EMST_E_ARG, "host-sync mode takes one graph and a given threshold");
  const int guard = (cfg.flags & EMST_F_NO_GUARD) ? 0 : 1;
  const int combine = (cfg.flags & EMST_F_NO_COMBINE) ? 0 : 1;
  const int dedup = (cfg.flags & EMST_F_NO_PAIR_DEDUP) ? 0 : 1;
  const int tail_max = host_sync || cfg.tail_max < 0 ? 0 : cfg.tail_max;
  const int round_grid = std::max(1, ctx->cus * std::max(1, cfg.grid_per_cu));
  const int stage_cap = std::max(64, cfg.stage_recs);
  const size_t init_lds = init_lds_bytes(stage_cap, task_edges);
  if (init_lds > ctx->lds_per_block) return ctx_fail(ctx, EMST_E_ARG, "stage_recs exceeds the LDS of a workgroup");
  const int nphases =
      (vt ? vt->any_filter : known ? ctx->learned_filter != 0 : use_filter != EMST_FILTER_OFF) ? 2 : 1;
  const int* vgid = vt ? vt->vgid : nullptr;
  if (vt && (!vt->vgid || !vt->gthr)) return ctx_fail(ctx, EMST_E_ARG, "a union needs its vertex-graph map and thresholds");
  const int* gthr = vt ? vt->gthr : nullptr;
  int* d_cnt = wb.cnt;
  long long* d_stats = wb.stats;
  int4** wl = wb.wl;
  ctx->last_nodes = nodes;
  ctx->last_parent = wb.parent;
  ctx->last_relabelled = kidmode != 0;
  const SelFn sel_first = sel_fn(true, waves), sel = sel_fn(false, waves);
  const JoinFn join = join_fn(waves);
  const ListView dense{nullptr, 0, 0};
  const ListView segs{wb.segcnt, z.seg_cap, z.init_grid};

  if (!fixed) HIPCHK(hipEventRecord(ctx->ev[0], s));
  if (learned && !learned_cached) {
    // features and inference decide the threshold on the device; the filter
    // decision is also posted to host-mapped memory for the phase loop
    ctx->learned_cached = false;  // d_thr now holds this graph's decision
    __atomic_store_n(ctx->h_decide, -1, __ATOMIC_RELEASE);
    launch_learned_threshold(s, nodes, edges, d_nindex, d_eweight, ctx->d_mlp16, ctx->have_model, ctx->d_thr, ctx->d_q,
                             ctx->d_feat, ctx->d_samples, ctx->d_decide);
    HIPCHK(hipGetLastError());
  }
  {
    const int blocks = (int)(((long long)nodes + 255) / 256);
    hipLaunchKernelGGL(k_setup, dim3(blocks), dim3(256), 0, s, nodes, d_nindex, wb.parent, wb.minv, wb.gsrc, d_cnt,
                       d_stats, hasmem);
    const long long z16 = ((long long)edges + 15) / 16;
    const int zb = (int)std::min<long long>(std::max<long long>((z16 + 255) / 256, 1), (long long)ctx->cus * 16);
    hipLaunchKernelGGL(k_zero, dim3(zb), dim3(256), 0, s, d_inMST, (long long)edges);
  }
  int init_launches = 0;
  int slot = 0;
  for (int phase = 0; phase < nphases; phase++) {
    if (phase == 1 && learned && learned_cached) {
      if (!ctx->learned_filter) break;  // decided before the solve: unfiltered
    } else if (phase == 1 && learned) {
      // The device decided the filter; an unfiltered decision needs no
      // heavy phase, so none is enqueued.
      int v;
      while ((v = __atomic_load_n(ctx->h_decide, __ATOMIC_ACQUIRE)) < 0) {
        const hipError_t q = hipStreamQuery(s);
        if (q == hipSuccess) {
          v = __atomic_load_n(ctx->h_decide, __ATOMIC_ACQUIRE);
          break;
        }
        if (q != hipErrorNotReady) return hip_fail(ctx, q, "threshold inference");
      }
      if (v < 0) return ctx_fail(ctx, EMST_E_HIP, "threshold inference posted no decision");
      if (v == 0) break;
    }
    if (phase == 1) {
      const int blocks = (int)(((long long)nodes + 255) / 256);
      hipLaunchKernelGGL(k_pick_big, dim3(1), dim3(kWave), 0, s, nodes, edges, d_nlist, wb.parent, &d_cnt[2], dthr);
      hipLaunchKernelGGL(k_flatten, dim3(blocks), dim3(256), 0, s, nodes, wb.parent, d_cnt, &d_cnt[2], wb.vclass,
                         dthr, hasmem);
      if (sparse && edges > 0) {
        const long long ngroups = ((long long)edges + kStartEdges - 1) / kStartEdges;
        hipLaunchKernelGGL(k_task_skip, dim3((unsigned)((ngroups + 255) / 256)), dim3(256), 0, s, ngroups, nodes,
                           wb.gsrc, wb.vclass, wb.skip, dthr, &d_cnt[3]);
      }
    }
    if (edges > 0) {
      if (!fixed) HIPCHK(hipEventRecord(ctx->ev[2 + 2 * phase], s));
      // With reverse positions both heavy passes are enqueued; the marked
      // component's share of the adjacency (k_pick_big) lets one of them run
      // and the other return at once.  The sparse pass also makes the first
      // round's offers, and then the offer-only select returns at once too.
      if (phase == 0 && light_offers) {
        // a union always takes the offers pass (its light share is per graph)
        hipLaunchKernelGGL(k_pick_light, dim3(1), dim3(kWave), 0, s, edges, d_eweight, T, dthr, &d_cnt[4],
                           vgid != nullptr ? 1 : 0);
        hipLaunchKernelGGL(light_offers_fn(vgid != nullptr), dim3(z.init_grid), dim3(kInitWaves * kWave),
                           sparse_lds_bytes(stage_cap), s, d_nindex, d_nlist, d_eweight, wb.gsrc, nodes, edges, T,
                           rix->rev, kidmode, wb.minv, wl[0], z.seg_cap, wb.segcnt, &d_cnt[0], stage_cap, d_stats,
                           dthr, &d_cnt[4], vgid, gthr);
      }
      if (phase == 1 && sparse) {
        const SparseFn f = sparse_fn();
        hipLaunchKernelGGL(f, dim3(z.init_grid), dim3(kInitWaves * kWave), sparse_lds_bytes(stage_cap), s, d_nindex,
                           d_nlist, d_eweight, wb.gsrc, nodes, edges, wb.parent, wb.vclass, wb.skip, rix->rev,
                           rix->mirrored ? 1 : 0, kidmode, hasmem, wb.minv, wl[0], z.seg_cap, wb.segcnt, &d_cnt[0],
                           stage_cap, d_stats, dthr, &d_cnt[2]);
      }
      const InitFn f = init_fn(phase == 1, sub, vgid != nullptr);
      hipLaunchKernelGGL(f, dim3(z.init_grid), dim3(kInitWaves * kWave), init_lds, s, d_nindex, d_nlist, d_eweight,
                         wb.gsrc, nodes, edges, T, wb.parent, wb.vclass, wl[0], z.seg_cap, wb.segcnt, &d_cnt[0],
                         stage_cap, d_stats, dthr, vgid, gthr,
                         &d_cnt[2], phase == 1 ? (sparse ? &d_cnt[3] : nullptr) : (light_offers ? &d_cnt[4] : nullptr));
      if (!fixed) HIPCHK(hipEventRecord(ctx->ev[3 + 2 * phase], s));
      HIPCHK(hipGetLastError());
      init_launches++;
    }
    const int rounds_slot = phase == 0 ? ST_ROUNDS0 : ST_ROUNDS1;
    // set on the device when the phase's initialisation made its first
    // offers (the light offers pass or the sparse heavy pass ran)
    const int* const fused = phase == 0 ? (light_offers ? &d_cnt[4] : nullptr) : (sparse ? &d_cnt[3] : nullptr);
    // Device-driven rounds with a bounded look-ahead (none in host-sync
    // mode).  The phase's first select and join and the select after them
    // read the initialisation's segments; later lists are dense.
    const int lag = host_sync ? 0 : 2;
    const int first_slot = slot;
    for (int r = 0;; r++) {
      if (slot >= kHflagSlots) return ctx_fail(ctx, EMST_E_LIMIT, "round limit exceeded");
      // list `cur` holds the survivors of the last select; the phase's
      // first select is the offer-only pass over the initial list
      const int cur = r & 1, nxt = cur ^ 1;
      const ListView lv = r == 0 ? segs : dense;
      // one workgroup per segment over the segmented list (segments are
      // filled alike, so this balances the work)
      const int grid = r == 0 ? z.init_grid : round_grid;
      ctx->h_flag[slot] = -1;
      if (r == 0)
        hipLaunchKernelGGL(sel_first, dim3(grid), dim3(waves * kWave), 0, s, wl[0], &d_cnt[0], wl[1], &d_cnt[1],
                           wb.parent, wb.minv, guard, combine, dedup, d_stats, rounds_slot, wb.cand,
                           (const u64*)nullptr, lv, fused, kidmode ? rix->rev : nullptr);
      // the first join reads the candidate bitmasks; every join marks what it joined
      hipLaunchKernelGGL(join, dim3(grid), dim3(waves * kWave), 0, s, wl[cur], &d_cnt[cur], wl[nxt], &d_cnt[nxt],
                         wb.parent, wb.minv, d_inMST, tail_max, ctx->d_flag + slot, d_stats, rounds_slot,
                         (const u64*)(r == 0 ? wb.cand : nullptr), wb.cand, lv, r == 0 ? fused : nullptr);
      // each select drops the records the join before it joined
      hipLaunchKernelGGL(sel, dim3(grid), dim3(waves * kWave), 0, s, wl[cur], &d_cnt[cur], wl[nxt], &d_cnt[nxt],
                         wb.parent, wb.minv, guard, combine, dedup, d_stats, rounds_slot, (u64*)nullptr,
                         (const u64*)wb.cand, lv, (const int*)nullptr, (const int*)nullptr);
      HIPCHK(hipGetLastError());
      slot++;
      if (fixed) {
        if (r + 1 >= fixed->rounds[phase]) {
          fixed->flag_slot[phase] = slot - 1;
          break;
        }
        continue;
      }
      if (r >= lag) {
        int* f = ctx->h_flag + first_slot + (r - lag);
        int v;
        while ((v = __atomic_load_n(f, __ATOMIC_ACQUIRE)) < 0) {
          const hipError_t q = hipStreamQuery(s);
          if (q == hipSuccess) {  // stream drained: the flag is final now
            v = __atomic_load_n(f, __ATOMIC_ACQUIRE);
            break;
          }
          if (q != hipErrorNotReady) return hip_fail(ctx, q, "round kernels");
        }
        if (v <= 0) break;
      }
    }
  }
  if (fixed) return EMST_OK;  // captured: the caller times and reads back
  if (stats) {
    hipLaunchKernelGGL(k_report, dim3(1), dim3(64), 0, s, d_stats, learned ? ctx->d_thr : nullptr,
                       learned ? ctx->d_q : nullptr, sparse ? &d_cnt[3] : nullptr, light_offers ? &d_cnt[4] : nullptr,
                       ctx->d_report);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(ctx->ev[1], s));
  HIPCHK(hipEventSynchronize(ctx->ev[1]));
  HIPCHK(hipGetLastError());
  if (stats) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]));
    float tph[2] = {0, 0};
    double init_ms = 0;
    for (int ph = 0; ph < init_launches; ph++) {
      HIPCHK(hipEventElapsedTime(&tph[ph], ctx->ev[2 + 2 * ph], ctx->ev[3 + 2 * ph]));
      init_ms += tph[ph];
    }
    long long hs[kReportWords];
    for (int i = 0; i < kReportWords; i++) hs[i] = __atomic_load_n(&ctx->h_report[i], __ATOMIC_ACQUIRE);
    std::memset(stats, 0, sizeof(*stats));
    if (learned) {
      stats->threshold_used = (int32_t)hs[ST_COUNT];
      stats->filter_used = (int32_t)hs[ST_COUNT + 1];
      const unsigned qb = (unsigned)hs[ST_COUNT + 2];
      std::memcpy(&stats->learned_q, &qb, 4);
    } else {
      stats->threshold_used = T;
      stats->filter_used = nphases == 2 ? 1 : 0;
      stats->learned_q = -1.f;
    }
    stats->ms = t;
    stats->init_ms = init_ms;
    stats->init_light_ms = tph[0];
    stats->init_heavy_ms = tph[1];
    stats->init_launches = init_launches;
    stats->light = hs[ST_LIGHT];
    stats->heavy = hs[ST_HEAVY];
    stats->light_rounds = (int32_t)hs[ST_ROUNDS0];
    stats->heavy_rounds = (int32_t)hs[ST_ROUNDS1];
    stats->tail_rounds = (int32_t)hs[ST_TAIL];
    stats->round_in_sum = hs[ST_IN_SUM];
    stats->round_out_sum = hs[ST_OUT_SUM];
    stats->sparse_heavy = (int32_t)hs[ST_COUNT + 3];
    stats->light_offers = (int32_t)hs[ST_COUNT + 4];
    stats->relabelled = kidmode ? ctx->rl.mode : 0;
  }
  return EMST_OK;
}

}  // namespace emst

extern "C" {

int emst_create(int device, emst_ctx** out)
{
  if (!out) return EMST_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return EMST_E_HIP;
  if (device < 0 || device >= ndev) return EMST_E_ARG;
  emst_ctx* ctx = new emst_ctx();
  ctx->device = device;
  ctx->cfg = default_config();
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete ctx;
    return EMST_E_HIP;
  }
  ctx->cus = prop.multiProcessorCount;
  ctx->lds_per_block = prop.sharedMemPerBlock;
  bool ok = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && noalloc_setup(device) == EMST_OK;
  ok = ok && hipMalloc((void**)&ctx->d_cnt, kCntWords * sizeof(int)) == hipSuccess;
  ok = ok && hipMalloc((void**)&ctx->d_stats, ST_COUNT * sizeof(long long)) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&ctx->h_flag, kHflagSlots * sizeof(int), hipHostMallocMapped) == hipSuccess;
  ok = ok && hipHostGetDevicePointer((void**)&ctx->d_flag, ctx->h_flag, 0) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&ctx->h_report, kReportWords * sizeof(long long), hipHostMallocMapped) == hipSuccess;
  ok = ok && hipHostGetDevicePointer((void**)&ctx->d_report, ctx->h_report, 0) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&ctx->h_decide, sizeof(int), hipHostMallocMapped) == hipSuccess;
  ok = ok && hipHostGetDevicePointer((void**)&ctx->d_decide, ctx->h_decide, 0) == hipSuccess;
  for (int i = 0; ok && i < 8; i++) ok = hipEventCreate(&ctx->ev[i]) == hipSuccess;
  ok = ok && hipMalloc((void**)&ctx->d_mlp16, kParams * 2) == hipSuccess &&
       hipMalloc((void**)&ctx->d_thr, 8) == hipSuccess && hipMalloc((void**)&ctx->d_q, 4) == hipSuccess &&
       hipMalloc((void**)&ctx->d_feat, kIn * 4) == hipSuccess &&
       hipMalloc((void**)&ctx->d_samples, kWeightSamples * 4) == hipSuccess;
  if (ok) {
    const int dyn = (int)ctx->lds_per_block;
    for (int sub : {4, 8, 16}) {
      for (int h = 0; h < 4 && ok; h++)
        ok = ok && hipFuncSetAttribute((const void*)init_fn(h & 1, sub, h >= 2),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, dyn) == hipSuccess;
    }
    ok = ok && hipFuncSetAttribute((const void*)sparse_fn(), hipFuncAttributeMaxDynamicSharedMemorySize, dyn) ==
                   hipSuccess;
    ok = ok && hipFuncSetAttribute((const void*)light_offers_fn(true), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   dyn) == hipSuccess;
    ok = ok && hipFuncSetAttribute((const void*)light_offers_fn(false), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   dyn) == hipSuccess;
  }
  if (!ok) {
    emst_destroy(ctx);
    return EMST_E_HIP;
  }
  *out = ctx;
  return EMST_OK;
}

void emst_destroy(emst_ctx* ctx)
{
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  void* ptrs[] = {ctx->d_cand,   ctx->d_nindex, ctx->d_nlist, ctx->d_eweight, ctx->d_inMST, ctx->d_parent, ctx->d_segcnt,
                  ctx->d_skip,   ctx->d_hasmem, ctx->d_rev,
                  ctx->d_minv,   ctx->d_wl[0], ctx->d_wl[1],   ctx->d_gsrc,  ctx->d_cnt,
                  ctx->d_stats,  ctx->d_vclass, ctx->d_mlp16,  ctx->d_thr,   ctx->d_q,
                  ctx->d_feat,   ctx->d_samples, ctx->d_lad[0], ctx->d_lad[1], ctx->d_slist,
                  ctx->d_lad_cnt};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  free_relabel(ctx->rl);
  if (ctx->h_flag) (void)hipHostFree(ctx->h_flag);
  if (ctx->h_report) (void)hipHostFree(ctx->h_report);
  if (ctx->h_decide) (void)hipHostFree(ctx->h_decide);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* emst_last_error(const emst_ctx* ctx)
{
  if (!ctx) return "no context";
  return ctx->err.c_str();
}

int emst_set_config(emst_ctx* ctx, const emst_config* cfg)
{
  if (!ctx || !cfg) return EMST_E_ARG;
  emst_config c = default_config();
  if (cfg->block) c.block = cfg->block;
  if (cfg->grid_per_cu) c.grid_per_cu = cfg->grid_per_cu;
  if (cfg->tile_edges) c.tile_edges = cfg->tile_edges;
  if (cfg->stage_recs) c.stage_recs = cfg->stage_recs;
  if (cfg->tail_max) c.tail_max = cfg->tail_max;
  if (cfg->init_grid_per_cu) c.init_grid_per_cu = cfg->init_grid_per_cu;
  c.flags = cfg->flags;
  if (c.flags & ~kAllFlags) return ctx_fail(ctx, EMST_E_ARG, "unknown flag bits");
  if (c.grid_per_cu < 1 || c.grid_per_cu > 64 || c.init_grid_per_cu < 1 || c.init_grid_per_cu > 64)
    return ctx_fail(ctx, EMST_E_ARG, "grid_per_cu out of range");
  if (c.stage_recs < 64 || init_lds_bytes(c.stage_recs, c.tile_edges) > ctx->lds_per_block)
    return ctx_fail(ctx, EMST_E_ARG, "stage_recs out of range");
  if (c.tail_max > (1 << 20)) return ctx_fail(ctx, EMST_E_ARG, "tail_max out of range");
  if (c.tile_edges != 256 && c.tile_edges != 512 && c.tile_edges != 1024)
    return ctx_fail(ctx, EMST_E_ARG, "tile_edges must be 256, 512 or 1024");
  if (c.block != 256 && c.block != 512 && c.block != 1024) return ctx_fail(ctx, EMST_E_ARG, "block must be 256, 512 or 1024");
  ctx->cfg = c;
  return EMST_OK;
}

int emst_get_config(const emst_ctx* ctx, emst_config* cfg)
{
  if (!ctx || !cfg) return EMST_E_ARG;
  *cfg = ctx->cfg;
  return EMST_OK;
}

int emst_load(emst_ctx* ctx, const emst_graph* g)
{
  if (!ctx || !g) return EMST_E_ARG;
  if (g->edges > 0 && !g->eweight) return ctx_fail(ctx, EMST_E_NOWEIGHT, "graph has no weights");
  int rc = emst_validate(g);
  if (rc) return ctx_fail(ctx, rc, "CSR failed validation");
  HIPCHK(hipSetDevice(ctx->device));
  const size_t ni = (size_t)g->nodes + 1, ne = (size_t)g->edges;
  if (ni > ctx->cap_nindex || !ctx->d_nindex) {
    if (ctx->d_nindex) (void)hipFree(ctx->d_nindex);
    ctx->d_nindex = nullptr;
    if (hipMalloc((void**)&ctx->d_nindex, ni * 4) != hipSuccess) return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc nindex");
    ctx->cap_nindex = ni;
  }
  if (ne > ctx->cap_edges || !ctx->d_nlist) {
    for (void** p : {(void**)&ctx->d_nlist, (void**)&ctx->d_eweight, (void**)&ctx->d_inMST}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
    ctx->cap_edges = 0;
    const size_t want = std::max<size_t>(ne, 1);
    if (hipMalloc((void**)&ctx->d_nlist, want * 4) != hipSuccess ||
        hipMalloc((void**)&ctx->d_eweight, want * 4) != hipSuccess ||
        hipMalloc((void**)&ctx->d_inMST, want + 16) != hipSuccess)
      return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of the CSR failed");
    ctx->cap_edges = want;
  }
  HIPCHK(hipMemcpyAsync(ctx->d_nindex, g->nindex, ni * 4, hipMemcpyHostToDevice, ctx->stream));
  if (ne) {
    HIPCHK(hipMemcpyAsync(ctx->d_nlist, g->nlist, ne * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_eweight, g->eweight, ne * 4, hipMemcpyHostToDevice, ctx->stream));
  }
  // Worklist capacity: the number of upper-direction edges (the reference
  // assumes a symmetric input and allocates edges/2, ECL_MST.cu:279, 285).
  // Lower-direction positions are counted too: reverse positions need as
  // many of them as upper-direction ones.
  const int nt = emst::host_threads();
  std::vector<long long> part((size_t)nt, 0), part_lo((size_t)nt, 0), part_max((size_t)nt, 0),
      part_near((size_t)nt, 0);
  emst::parallel_for(g->nodes, [&](int64_t b, int64_t e, int t) {
    long long c = 0, lo = 0, mx = 0, near = 0;
    for (int64_t v = b; v < e; v++) {
      for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) {
        c += g->nlist[j] > v;
        lo += g->nlist[j] < v;
        near += g->nlist[j] != v && std::llabs((long long)g->nlist[j] - v) <= 64;
      }
      mx = std::max<long long>(mx, g->nindex[v + 1] - g->nindex[v]);
    }
    part[t] = c;
    part_lo[t] = lo;
    part_max[t] = mx;
    part_near[t] = near;
  }, nt);
  long long up = 0, low = 0, maxdeg = 0, near = 0;
  for (int t = 0; t < nt; t++) {
    up += part[t];
    low += part_lo[t];
    maxdeg = std::max(maxdeg, part_max[t]);
    near += part_near[t];
  }
  ctx->loaded = false;
  ctx->learned_cached = false;
  ctx->last_parent = nullptr;  // buffers may move; a new graph has no solve yet
  ctx->last_relabelled = false;
  ctx->rev_ok = false;
  ctx->rev_mirrored = false;
  if (up == low && ne > 0) {
    // reverse positions for the sparse heavy initialisation (a sorted,
    // symmetric CSR: the generators' output and the reference's inputs)
    if (ne > ctx->cap_rev || !ctx->d_rev) {
      if (ctx->d_rev) (void)hipFree(ctx->d_rev);
      ctx->d_rev = nullptr;
      ctx->cap_rev = 0;
      if (hipMalloc((void**)&ctx->d_rev, ne * 4) != hipSuccess)
        return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of the reverse positions failed");
      ctx->cap_rev = ne;
    }
    int rc2 = build_rev(ctx, g->nodes, g->edges, ctx->d_nindex, ctx->d_nlist, ctx->d_eweight, ctx->d_rev, ctx->stream,
                        &ctx->rev_ok, &ctx->rev_mirrored);
    if (rc2) return rc2;
  }
  // Solve on a renumbered copy (it needs the reverse positions, which give
  // every lower-direction copy its key id):
  //  - skewed degrees (the largest at least 32 times the mean): by
  //    decreasing degree, so the hubs' entries share lines;
  //  - scattered ids (under a tenth of the edges join ids at most 64 apart)
  //    on a graph of high diameter: by breadth-first clusters, so most
  //    edges join nearby ids.
  ctx->rl.ok = false;
  ctx->rl.ms = 0;
  ctx->rl.mode = 0;
  const int fl = ctx->cfg.flags;
  const bool skewed = g->nodes >= 1024 && maxdeg * (long long)g->nodes >= 32LL * g->edges;
  const bool scattered = g->nodes >= 65536 && near * 10 < (long long)g->edges;
  int order = 0, min_levels = 0;
  if (ctx->rev_ok && !(fl & EMST_F_NO_RELABEL)) {
    if (fl & EMST_F_RELABEL) order = kOrderDegree;
    else if (fl & EMST_F_CLUSTER) order = kOrderCluster;
    else if (skewed) order = kOrderDegree;
    else if (scattered) {
      order = kOrderCluster;
      min_levels = 16;
    }
  }
  if (order) {
    const int rc3 = build_relabel(ctx, g->nodes, g->edges, ctx->d_nindex, ctx->d_nlist, ctx->d_eweight, ctx->d_rev,
                                  ctx->stream, order, min_levels);
    if (rc3 == EMST_E_NOMEM) {
      // the copy is an optimisation: without room for it, solve in the
      // caller's numbering
      free_relabel(ctx->rl);
      ctx->err.clear();
      (void)hipGetLastError();  // the failed allocation's error is not the solve's
    } else if (rc3) {
      return rc3;
    }
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->n = g->nodes;
  ctx->m = g->edges;
  ctx->m_up = up;
  ctx->loaded = true;
  return ensure_buffers(ctx, g->nodes, g->edges, up);
}

int emst_solve(emst_ctx* ctx, int use_filter, int32_t threshold, emst_stats* stats)
{
  if (!ctx) return EMST_E_ARG;
  if (!ctx->loaded) return ctx_fail(ctx, EMST_E_STATE, "emst_solve before emst_load");
  HIPCHK(hipSetDevice(ctx->device));
  // the renumbered copy, unless the ladder (which keys by position) or
  // the caller's numbering is asked for
  const Relabel& rl = ctx->rl;
  const bool use_rl = rl.ok && !(ctx->cfg.flags & (kLadderFlags | EMST_F_NO_RELABEL));
  RevIndex rix;
  if (use_rl) {
    rix.rev = rl.kid;
    rix.mirrored = true;  // both copies carry the upper copy's weight
    rix.kid = true;
  } else if (ctx->rev_ok) {
    rix.rev = ctx->d_rev;
    rix.mirrored = ctx->rev_mirrored;
  }
  if (use_filter == EMST_FILTER_LEARNED && !ctx->learned_cached && !(ctx->cfg.flags & kLadderFlags)) {
    // The learned threshold is a property of the graph: decide it once per
    // loaded graph and model, before (outside) the timed region, as the
    // reference computes its threshold before calling gpuMST
    // (ECL_MST.cu:432-441).
    launch_learned_threshold(ctx->stream, ctx->n, ctx->m, ctx->d_nindex, ctx->d_eweight, ctx->d_mlp16,
                             ctx->have_model, ctx->d_thr, ctx->d_q, ctx->d_feat, ctx->d_samples, nullptr);
    HIPCHK(hipGetLastError());
    int thr[2];
    HIPCHK(hipMemcpyAsync(thr, ctx->d_thr, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->learned_filter = thr[1];
    ctx->learned_T = thr[0];
    ctx->learned_cached = true;
  }
  return run_solve(ctx, ctx->n, ctx->m, use_rl ? rl.nindex : ctx->d_nindex, use_rl ? rl.nlist : ctx->d_nlist,
                   use_rl ? rl.eweight : ctx->d_eweight, use_filter, threshold,
                   ctx->d_inMST, ctx->stream, ctx->m_up, nullptr, stats, nullptr, nullptr, &rix,
                   use_filter == EMST_FILTER_LEARNED && ctx->learned_cached);
}

int emst_fetch(emst_ctx* ctx, uint8_t* inMST)
{
  if (!ctx || (!inMST && ctx->m > 0)) return EMST_E_ARG;
  if (!ctx->loaded) return ctx_fail(ctx, EMST_E_STATE, "emst_fetch before emst_load");
  HIPCHK(hipSetDevice(ctx->device));
  if (ctx->m) HIPCHK(hipMemcpyAsync(inMST, ctx->d_inMST, (size_t)ctx->m, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return EMST_OK;
}

int emst_fetch_parent(emst_ctx* ctx, int32_t* parent)
{
  if (!ctx || !parent) return EMST_E_ARG;
  if (!ctx->last_parent || ctx->last_nodes < 1) return ctx_fail(ctx, EMST_E_STATE, "no solve has run");
  HIPCHK(hipSetDevice(ctx->device));
  if (ctx->last_relabelled) return relabel_parent(ctx, ctx->last_parent, ctx->last_nodes, parent, ctx->stream);
  HIPCHK(hipMemcpyAsync(parent, ctx->last_parent, (size_t)ctx->last_nodes * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return EMST_OK;
}

int emst_gpu_mst(emst_ctx* ctx, const emst_graph* g, int use_filter, int32_t threshold, uint8_t* inMST,
                 emst_stats* stats)
{
  if (!ctx || !g || (!inMST && g->edges > 0)) return EMST_E_ARG;
  int rc = emst_load(ctx, g);
  if (rc) return rc;
  rc = emst_solve(ctx, use_filter, threshold, stats);
  if (rc) return rc;
  rc = emst_fetch(ctx, inMST);
  if (rc) return rc;
  if (stats) {
    stats->mst_edges = 0;
    for (int32_t j = 0; j < g->edges; j++) stats->mst_edges += inMST[j] != 0;
    stats->weight = emst_mst_weight(g, inMST);
  }
  return EMST_OK;
}

int emst_set_filter_model(emst_ctx* ctx, const float* params)
{
  if (!ctx) return EMST_E_ARG;
  ctx->learned_cached = false;
  if (!params) {
    ctx->have_model = 0;
    return EMST_OK;
  }
  std::vector<uint16_t> h(kParams);
  for (int i = 0; i < kParams; i++) {
    if (!std::isfinite(params[i])) return ctx_fail(ctx, EMST_E_ARG, "non-finite model parameter");
    h[i] = emst_bf16_bits(params[i]);
  }
  HIPCHK(hipSetDevice(ctx->device));
  // the solve kernels read the parameters on the context's non-blocking
  // stream, so the copy is ordered on that stream and h outlives it
  HIPCHK(hipMemcpyAsync(ctx->d_mlp16, h.data(), kParams * 2, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->have_model = 1;
  return EMST_OK;
}

int emst_learned_threshold(emst_ctx* ctx, int* use_filter, int32_t* threshold, float* q, float* features,
                           int32_t* samples)
{
  if (!ctx) return EMST_E_ARG;
  if (!ctx->loaded) return ctx_fail(ctx, EMST_E_STATE, "emst_learned_threshold before emst_load");
  HIPCHK(hipSetDevice(ctx->device));
  ctx->learned_cached = false;
  launch_learned_threshold(ctx->stream, ctx->n, ctx->m, ctx->d_nindex, ctx->d_eweight, ctx->d_mlp16, ctx->have_model,
                           ctx->d_thr, ctx->d_q, ctx->d_feat, ctx->d_samples, nullptr);
  HIPCHK(hipGetLastError());
  int thr[2];
  float qq;
  HIPCHK(hipMemcpyAsync(thr, ctx->d_thr, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(&qq, ctx->d_q, 4, hipMemcpyDeviceToHost, ctx->stream));
  if (features) HIPCHK(hipMemcpyAsync(features, ctx->d_feat, kIn * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (samples) HIPCHK(hipMemcpyAsync(samples, ctx->d_samples, kWeightSamples * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (use_filter) *use_filter = thr[1];
  if (threshold) *threshold = thr[0];
  if (q) *q = qq;
  return EMST_OK;
}

int emst_solve_device(emst_ctx* ctx, int32_t nodes, int32_t edges, const int32_t* d_nindex, const int32_t* d_nlist,
                      const int32_t* d_eweight, int use_filter, int32_t threshold, uint8_t* d_inMST, void* stream,
                      emst_stats* stats)
{
  if (!ctx || !d_nindex || (edges > 0 && (!d_nlist || !d_eweight || !d_inMST))) return EMST_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  return run_solve(ctx, nodes, edges, d_nindex, d_nlist, d_eweight, use_filter, threshold, d_inMST, s, edges, nullptr,
                   stats);
}

// Coordinate-descent autotuner over the runtime shape knobs, timing
// repeated solves of the loaded graph (median of `reps`).
int emst_autotune(emst_ctx* ctx, int use_filter, int32_t threshold, int32_t reps, emst_config* best, double* best_ms)
{
  if (!ctx || reps < 1) return EMST_E_ARG;
  if (!ctx->loaded) return ctx_fail(ctx, EMST_E_STATE, "emst_autotune before emst_load");
  const emst_config saved = ctx->cfg;
  auto measure = [&](const emst_config& c, double* ms) -> int {
    int rc = emst_set_config(ctx, &c);
    if (rc) return rc;
    std::vector<double> t;
    for (int i = 0; i < reps + 1; i++) {
      emst_stats st;
      rc = emst_solve(ctx, use_filter, threshold, &st);
      if (rc) return rc;
      if (i) t.push_back(st.ms);  // first run warms caches and buffers
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return EMST_OK;
  };
  emst_config cur = saved;
  cur.flags &= ~(EMST_F_HOST_SYNC | kLadderFlags);
  double cur_ms = 0;
  int rc = measure(cur, &cur_ms);
  if (rc) {
    ctx->cfg = saved;
    return rc;
  }
  struct Knob {
    int32_t emst_config::*field;
    std::vector<int32_t> values;
  };
  const std::vector<Knob> knobs = {
      {&emst_config::tile_edges, {256, 512, 1024}},   {&emst_config::stage_recs, {128, 256, 512, 1024}},
      {&emst_config::init_grid_per_cu, {3, 4, 6, 8}},    {&emst_config::block, {256, 512, 1024}},
      {&emst_config::grid_per_cu, {1, 2, 4, 8}},       {&emst_config::tail_max, {-1, 1024, 4096, 16384}}};
  for (int pass = 0; pass < 2; pass++) {
    for (const Knob& k : knobs) {
      for (int32_t v : k.values) {
        if (cur.*(k.field) == v) continue;
        emst_config c = cur;
        c.*(k.field) = v;
        double ms = 0;
        if (measure(c, &ms) != EMST_OK) continue;  // e.g. LDS limits: skip
        if (ms < cur_ms * 0.99) {
          cur = c;
          cur_ms = ms;
        }
      }
    }
  }
  ctx->cfg = cur;
  if (best) *best = cur;
  if (best_ms) *best_ms = cur_ms;
  return EMST_OK;
}

}  // extern "C"
