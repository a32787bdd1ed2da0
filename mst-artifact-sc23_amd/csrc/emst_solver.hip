// emst_solver.hip -- MI355X (gfx950) minimum spanning forest solver with the
// semantics of ECL-MST's gpuMST path (ECL-MST/ECL_MST.cu:127-362).
//
// Device flow of one solve (the reference's timed region, ECL_MST.cu:301-345):
//   k_setup          parent[v]=v, minv[v]=max, per-task start vertices
//                    (initPM, ECL_MST.cu:147-154), inMST cleared by k_zero
//   k_init_wl<light> records {v, u, w, j} of upper-direction edges with w<=T
//                    (initWL<true>, ECL_MST.cu:157-207)
//   rounds           k_select: drop records inside one component, write back
//                    representatives, guarded 64-bit atomicMin of the key on
//                    both representatives (kernel1, ECL_MST.cu:210-226)
//                    k_join: the record holding a representative's minimum
//                    key clears that minimum, hooks, and marks inMST[j]
//                    (kernel2 + kernel3, ECL_MST.cu:229-251, fused)
//   k_flatten        parent[v] = root(v) before the heavy pass
//   k_init_wl<heavy> heavy upper-direction edges whose roots differ
//                    (initWL<false>, ECL_MST.cu:157-207, launch :326)
//   rounds           as above
//
// The host does not read the worklist size back every round (the reference
// does, ECL_MST.cu:318).  Round kernels run as fixed grid-stride grids that
// read sizes from device memory; k_join posts the surviving count into
// host-mapped memory and the host stops enqueueing rounds when it reads 0,
// a couple of rounds behind the device.  When a round's survivors fit
// cfg.tail_max, workgroup 0 of k_join finishes the whole phase with
// workgroup barriers instead of further launches.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/emst.h"
#include "emst_kernels.hpp"
#include "learned.hpp"
#include "host_util.hpp"

using namespace emst;

namespace {

constexpr int kInitWaves = 4;        // waves per workgroup in k_init_wl
constexpr int kRoundWaves = 8;       // waves per workgroup in k_select/k_join
constexpr int kHflagSlots = 512;     // host-visible round results per solve

// ---------------------------------------------------------------------------
// k_setup: initPM (ECL_MST.cu:147-154) plus the start vertex of every
// kGroupEdges-edge task: the vertex v whose adjacency range holds the task's
// first edge writes it.  Zero-degree vertices own no edge and write nothing.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_setup(int nodes, const int* __restrict__ nindex,
                                               int* __restrict__ parent, u64* __restrict__ minv,
                                               int* __restrict__ gsrc, int* __restrict__ cnt,
                                               long long* __restrict__ stats)
{
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v == 0) {
    cnt[0] = 0;
    cnt[1] = 0;
    for (int i = 0; i < ST_COUNT; i++) stats[i] = 0;
  }
  if (v < nodes) {
    parent[v] = (int)v;
    minv[v] = kNoKey;
    const int beg = nindex[v], end = nindex[v + 1];
    if (end > beg) {
      const long long g0 = ((long long)beg + kGroupEdges - 1) / kGroupEdges;
      const long long g1 = ((long long)end + kGroupEdges - 1) / kGroupEdges;
      for (long long g = g0; g < g1; g++) gsrc[g] = (int)v;
    }
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
  for (long long i = t; i < n16; i += stride) q[i] = make_int4(0, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// k_init_wl: streams the CSR and emits worklist records.
//   HEAVY=false: {v, u, w, j} for u > v and w <= T      (ECL_MST.cu:170-176)
//   HEAVY=true : {root(v), root(u), w, j} for u > v with different roots
//                (ECL_MST.cu:166, 173-174); parent is flat here.  The
//                reference also tests w > T (ECL_MST.cu:197); after the
//                light phase every light edge lies inside one component, so
//                "roots differ" already implies w > T and the weight is
//                read only for crossing edges.  bigmap marks the members of
//                one large component (2 MB for 2^24 vertices, resident in
//                L2); an edge with both ends marked is skipped without
//                touching the root array.
// One wave task covers kGroupEdges consecutive CSR positions.  Lane i takes
// positions base+i, base+64+i, ... so every load is one coalesced 256-byte
// request, and all kSub loads of a task are issued before the first use.
// The source vertex of a position is found against a window of 64 vertex
// boundaries held one per lane: src(p) = max{v : nindex[v] <= p}.
// ---------------------------------------------------------------------------
template <bool HEAVY>
__global__ __launch_bounds__(kInitWaves * kWave) void k_init_wl(
    const int* __restrict__ nindex, const int* __restrict__ nlist, const int* __restrict__ eweight,
    const int* __restrict__ gsrc, int nodes, int edges, int T, const int* __restrict__ label,
    const unsigned* __restrict__ bigmap, int4* __restrict__ wl, int* __restrict__ cnt, int stage_cap,
    long long* __restrict__ stats, const int* __restrict__ dthr)
{
  extern __shared__ int4 lds_stage[];
  // dthr: threshold and filter flag decided on the device (learned mode)
  if (dthr) {
    if (HEAVY && dthr[1] == 0) return;
    T = dthr[0];
  }
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  Stage st{lds_stage + wid * stage_cap, stage_cap, 0};
  const long long ngroups = ((long long)edges + kGroupEdges - 1) / kGroupEdges;
  const long long nwaves = (long long)gridDim.x * kInitWaves;
  long long emitted = 0;
  for (long long g = (long long)blockIdx.x * kInitWaves + wid; g < ngroups; g += nwaves) {
    const long long gbeg = g * kGroupEdges;
    int nl[kSub], w[kSub];
    int base = gsrc[g];
#pragma unroll
    for (int u = 0; u < kSub; u++) {
      const long long p = gbeg + u * kWave + lane;
      nl[u] = p < edges ? nlist[p] : -1;
      if (!HEAVY) w[u] = p < edges ? eweight[p] : 0;
    }
    long long bi = (long long)base + 1 + lane;
    int bnd = bi <= nodes ? nindex[bi] : INT_MAX;
    int src[kSub];
#pragma unroll
    for (int u = 0; u < kSub; u++) {
      const long long pl = gbeg + u * kWave + lane;
      const int p = pl < edges ? (int)pl : INT_MAX - 1;
      bool done = pl >= edges;
      int s = 0;
      for (;;) {
        const int c = count_le(bnd, p);
        const int b63 = __shfl(bnd, kWave - 1);
        if (!done && p < b63) {
          s = base + c;
          done = true;
        }
        if (__all(done)) break;
        base += kWave;
        bi = (long long)base + 1 + lane;
        bnd = bi <= nodes ? nindex[bi] : INT_MAX;
      }
      src[u] = s;
    }
    if (!HEAVY) {
#pragma unroll
      for (int u = 0; u < kSub; u++) {
        const int j = (int)(gbeg + u * kWave + lane);
        const bool keep = nl[u] > src[u] && w[u] <= T;
        emitted += keep;
        stage_push(st, keep, make_int4(src[u], nl[u], w[u], j), wl, cnt);
      }
    } else {
      bool both[kSub];
#pragma unroll
      for (int u = 0; u < kSub; u++) {
        both[u] = true;
        if (nl[u] > src[u]) {
          const unsigned bu = bigmap[(unsigned)nl[u] >> 5] >> (nl[u] & 31);
          const unsigned bv = bigmap[(unsigned)src[u] >> 5] >> (src[u] & 31);
          both[u] = (bu & bv & 1u) != 0;
        }
      }
      int a[kSub], b[kSub];
#pragma unroll
      for (int u = 0; u < kSub; u++) {
        a[u] = 0;
        b[u] = 0;
        if (!both[u]) {
          a[u] = label[src[u]];
          b[u] = label[nl[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < kSub; u++) {
        const long long p = gbeg + u * kWave + lane;
        const bool keep = !both[u] && a[u] != b[u];
        const int wt = keep ? eweight[p] : 0;
        emitted += keep;
        stage_push(st, keep, make_int4(a[u], b[u], wt, (int)p), wl, cnt);
      }
    }
  }
  stage_flush(st, wl, cnt);
  if (emitted) atomicAdd((unsigned long long*)&stats[HEAVY ? ST_HEAVY : ST_LIGHT], (unsigned long long)emitted);
}

// Picks the component that the heavy pass' bitmap marks: the most frequent
// root among 64 evenly spaced vertices (one wave).  Any choice is correct;
// a frequent one saves the most root lookups.
__global__ __launch_bounds__(64) void k_pick_big(int nodes, const int* __restrict__ parent, int* __restrict__ out)
{
  const int lane = lane_id();
  const int v = (int)(((long long)lane * nodes) / kWave + ((long long)nodes / (2 * kWave)));
  const int r = find_root(v < nodes ? v : nodes - 1, parent);
  int c = 0;
  for (int k = 0; k < kWave; k++) c += __shfl(r, k) == r;
  int best = (c << 8) | (kWave - 1 - lane);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) best = max(best, __shfl_xor(best, off));
  const int winner = kWave - 1 - (best & 0xff);
  const int big = __shfl(r, winner);
  if (lane == 0) out[0] = big;
}

// ---------------------------------------------------------------------------
// k_select (kernel1, ECL_MST.cu:210-226): for each record find both roots;
// records inside one component are dropped, the rest are written back with
// their roots (implicit path compression) and offer their key to both roots.
//
// Two MI355X-specific changes, both invisible in the results:
//  * Survivors of a 2048-record tile are compacted with a workgroup scan and
//    reserved with one atomicAdd per tile (a per-record or per-wave counter
//    atomic serialises on one address).
//  * Keys are first min-combined per workgroup in an LDS table keyed by
//    root (4-probe open addressing, claimed slots are never released).  A hub
//    component that appears in most records is claimed early and then
//    costs one global atomicMin per workgroup instead of one per record;
//    device-scope atomics and coherent loads are resolved beyond the
//    per-XCD L2, so a hot address otherwise serialises the whole grid.
//    Roots that find no slot fall back to the guarded global atomic.
// ---------------------------------------------------------------------------
constexpr int kSelItems = 4;
constexpr int kSelTile = kRoundWaves * kWave * kSelItems;
constexpr int kTbl = 1024;

__device__ __forceinline__ void offer_key(int* tags, u64* keys, u64* minv, int rep, u64 key, bool guard,
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
        atomicMin(&keys[sl], key);
        return;
      }
    }
  }
  min_key(minv, rep, key, guard);
}

template <bool FIRST>
__global__ __launch_bounds__(kRoundWaves * kWave) void k_select(
    const int4* __restrict__ wl_in, const int* __restrict__ cnt_in, int4* __restrict__ wl_out,
    int* __restrict__ cnt_out, const int* __restrict__ parent, u64* __restrict__ minv, int guard, int combine,
    long long* __restrict__ stats, int rounds_slot)
{
  __shared__ int s_tags[kTbl];
  __shared__ u64 s_keys[kTbl];
  __shared__ int s_wtot[kRoundWaves];
  __shared__ int s_woff[kRoundWaves];
  __shared__ int s_base;
  const int n_in = *cnt_in;
  if (n_in == 0) return;
  const int tid = threadIdx.x;
  if (blockIdx.x == 0 && tid == 0) {
    stats[rounds_slot] += 1;
    stats[ST_IN_SUM] += n_in;
  }
  for (int i = tid; i < kTbl; i += blockDim.x) {
    s_tags[i] = -1;
    s_keys[i] = kNoKey;
  }
  __syncthreads();
  const int lane = lane_id();
  const int wid = tid / kWave;
  const long long ntiles = ((long long)n_in + kSelTile - 1) / kSelTile;
  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const long long base = t * kSelTile;
    int4 r[kSelItems];
#pragma unroll
    for (int k = 0; k < kSelItems; k++) {
      const long long i = base + k * (kRoundWaves * kWave) + tid;
      r[k] = i < n_in ? wl_in[i] : make_int4(0, 0, 0, 0);
    }
    bool keep[kSelItems];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < kSelItems; k++) {
      const long long i = base + k * (kRoundWaves * kWave) + tid;
      keep[k] = false;
      if (i < n_in) {
        // In a phase's first round every endpoint is a root and every
        // record crosses (initialisation guarantees both), so the finds
        // and the copy are skipped: the list is joined in place.
        const int a = FIRST ? r[k].x : find_root(r[k].x, parent);
        const int c = FIRST ? r[k].y : find_root(r[k].y, parent);
        if (a != c) {
          keep[k] = true;
          r[k].x = a;
          r[k].y = c;
          const u64 key = make_key(r[k].z, r[k].w);
          offer_key(s_tags, s_keys, minv, a, key, guard, combine);
          offer_key(s_tags, s_keys, minv, c, key, guard, combine);
        }
      }
      mine += keep[k];
    }
    if (FIRST) continue;
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
      for (int w = 0; w < kRoundWaves; w++) {
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
  }
  // publish the workgroup's combined minima (after every offer of every
  // wave: the first-round pass has no per-tile barrier)
  __syncthreads();
  for (int i = tid; i < kTbl; i += blockDim.x) {
    const int rep = s_tags[i];
    if (rep >= 0) min_key(minv, rep, s_keys[i], guard);
  }
}

// One workgroup finishes a phase: join the current survivors, then repeat
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
      const u64 key = make_key(r.z, r.w);
      const u64 kx = __hip_atomic_load(&minv[r.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const u64 ky = __hip_atomic_load(&minv[r.y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const bool wx = key == kx, wy = key == ky;
      if (wx) __hip_atomic_store(&minv[r.x], kNoKey, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (wy) __hip_atomic_store(&minv[r.y], kNoKey, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (wx || wy) {
        hook(r.x, r.y, parent);
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
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kRoundWaves * kWave) void k_join(
    int4* __restrict__ wl, int* __restrict__ cnt, int4* __restrict__ wl_other,
    int* __restrict__ cnt_other, int* __restrict__ parent, u64* __restrict__ minv,
    uint8_t* __restrict__ inMST, int tail_max, int* __restrict__ hflag, long long* __restrict__ stats,
    int rounds_slot)
{
  const int n = *cnt;
  if (n > 0 && n <= tail_max) {
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
    __hip_atomic_store(hflag, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int4 r = wl[i];
    const u64 key = make_key(r.z, r.w);
    const u64 kx = minv[r.x];
    const u64 ky = minv[r.y];
    const bool wx = key == kx, wy = key == ky;
    if (wx) minv[r.x] = kNoKey;
    if (wy) minv[r.y] = kNoKey;
    if (wx || wy) {
      hook(r.x, r.y, parent);
      inMST[r.w] = 1;
    }
  }
}

// parent[v] = root(v), and bit v of bigmap = (root(v) == *big).  Concurrent
// writers only shorten chains toward the same root, so readers stay correct.
__global__ __launch_bounds__(256) void k_flatten(int nodes, int* __restrict__ parent, int* __restrict__ cnt,
                                                 const int* __restrict__ big, unsigned* __restrict__ bigmap,
                                                 const int* __restrict__ dthr)
{
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v == 0) {
    cnt[0] = 0;
    cnt[1] = 0;
  }
  if (dthr && dthr[1] == 0) return;  // unfiltered: no heavy phase
  const int bigroot = *big;
  int r = -1;
  if (v < nodes) {
    r = find_root((int)v, parent);
    parent[v] = r;
  }
  const u64 m = __ballot(r == bigroot && r >= 0);
  const long long v0 = v - lane_id();
  const long long words = ((long long)nodes + 31) / 32;
  if (lane_id() == 0 && v0 < nodes) {
    bigmap[v0 >> 5] = (unsigned)m;
    if ((v0 >> 5) + 1 < words) bigmap[(v0 >> 5) + 1] = (unsigned)(m >> 32);
  }
}

// Host-sync ablation: a plain kernel2 without the fused tail.
__global__ __launch_bounds__(256) void k_join_plain(const int4* __restrict__ wl, int n, int* __restrict__ parent,
                                                    u64* __restrict__ minv, uint8_t* __restrict__ inMST)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int4 r = wl[i];
    const u64 key = make_key(r.z, r.w);
    const u64 kx = minv[r.x];
    const u64 ky = minv[r.y];
    const bool wx = key == kx, wy = key == ky;
    if (wx) minv[r.x] = kNoKey;
    if (wy) minv[r.y] = kNoKey;
    if (wx || wy) {
      hook(r.x, r.y, parent);
      inMST[r.w] = 1;
    }
  }
}

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================
struct emst_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 0;
  size_t lds_per_block = 0;
  emst_config cfg{};
  std::string err;
  // graph owned by the context (emst_load)
  int32_t n = 0, m = 0;
  long long m_up = -1;
  bool loaded = false;
  int* d_nindex = nullptr;
  int* d_nlist = nullptr;
  int* d_eweight = nullptr;
  uint8_t* d_inMST = nullptr;
  size_t cap_nindex = 0, cap_edges = 0;
  // work buffers, grown on demand and reused
  int* d_parent = nullptr;
  u64* d_minv = nullptr;
  unsigned* d_bigmap = nullptr;
  size_t cap_nodes = 0;
  int4* d_wl[2] = {nullptr, nullptr};
  size_t cap_wl = 0;
  int* d_gsrc = nullptr;
  size_t cap_gsrc = 0;
  int* d_cnt = nullptr;
  long long* d_stats = nullptr;
  int* h_flag = nullptr;
  int* d_flag = nullptr;
  hipEvent_t ev[8] = {};
  int32_t last_nodes = 0;
  // learned threshold
  unsigned short* d_mlp16 = nullptr;  // bf16 parameters
  int have_model = 0;
  int* d_thr = nullptr;               // [threshold, use_filter] decided on the device
  float* d_q = nullptr;
  float* d_feat = nullptr;
  int* d_samples = nullptr;
};

namespace {

int fail(emst_ctx* c, int code, const std::string& what)
{
  if (c) c->err = what;
  return code;
}

int hip_fail(emst_ctx* c, hipError_t e, const char* what)
{
  char buf[256];
  std::snprintf(buf, sizeof(buf), "HIP error %d in %s: %s", (int)e, what, hipGetErrorString(e));
  return fail(c, EMST_E_HIP, buf);
}

#define HIPCHK(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
  } while (0)

emst_config default_config()
{
  emst_config c{};
  c.block = kRoundWaves * kWave;
  c.grid_per_cu = 4;
  c.tile_edges = kGroupEdges;
  c.stage_recs = 512;
  c.tail_max = 4096;
  c.flags = 0;
  return c;
}

}  // namespace

// The work buffers are managed in one function so their capacities stay in step.
static int ensure_buffers(emst_ctx* ctx, int32_t nodes, int32_t edges, long long wl_cap)
{
  if ((size_t)nodes > ctx->cap_nodes || !ctx->d_parent || !ctx->d_minv) {
    if (ctx->d_parent) (void)hipFree(ctx->d_parent);
    if (ctx->d_minv) (void)hipFree(ctx->d_minv);
    if (ctx->d_bigmap) (void)hipFree(ctx->d_bigmap);
    ctx->d_parent = nullptr;
    ctx->d_minv = nullptr;
    ctx->d_bigmap = nullptr;
    ctx->cap_nodes = 0;
    if (hipMalloc((void**)&ctx->d_parent, (size_t)std::max(nodes, 1) * 4) != hipSuccess ||
        hipMalloc((void**)&ctx->d_minv, (size_t)std::max(nodes, 1) * 8) != hipSuccess ||
        hipMalloc((void**)&ctx->d_bigmap, ((size_t)std::max(nodes, 1) + 63) / 32 * 4) != hipSuccess)
      return fail(ctx, EMST_E_NOMEM, "hipMalloc of parent/minv failed");
    ctx->cap_nodes = (size_t)std::max(nodes, 1);
  }
  const size_t groups = ((size_t)edges + kGroupEdges - 1) / kGroupEdges;
  if (groups > ctx->cap_gsrc || !ctx->d_gsrc) {
    if (ctx->d_gsrc) (void)hipFree(ctx->d_gsrc);
    ctx->d_gsrc = nullptr;
    ctx->cap_gsrc = 0;
    if (hipMalloc((void**)&ctx->d_gsrc, std::max<size_t>(groups, 1) * 4) != hipSuccess)
      return fail(ctx, EMST_E_NOMEM, "hipMalloc of task start vertices failed");
    ctx->cap_gsrc = std::max<size_t>(groups, 1);
  }
  if ((size_t)wl_cap > ctx->cap_wl || !ctx->d_wl[0]) {
    for (int i = 0; i < 2; i++) {
      if (ctx->d_wl[i]) (void)hipFree(ctx->d_wl[i]);
      ctx->d_wl[i] = nullptr;
    }
    ctx->cap_wl = 0;
    const size_t want = std::max<size_t>((size_t)wl_cap, 1);
    for (int i = 0; i < 2; i++)
      if (hipMalloc((void**)&ctx->d_wl[i], want * sizeof(int4)) != hipSuccess) {
        char buf[128];
        std::snprintf(buf, sizeof(buf), "hipMalloc of a %zu-record worklist failed", want);
        return fail(ctx, EMST_E_NOMEM, buf);
      }
    ctx->cap_wl = want;
  }
  return EMST_OK;
}

static int run_solve(emst_ctx* ctx, int32_t nodes, int32_t edges, const int* d_nindex, const int* d_nlist,
                     const int* d_eweight, int use_filter, int32_t threshold, uint8_t* d_inMST,
                     hipStream_t s, long long wl_cap, emst_stats* stats)
{
  if (nodes < 1 || edges < 0) return fail(ctx, EMST_E_ARG, "node or edge count too low");
  int rc = ensure_buffers(ctx, nodes, edges, wl_cap);
  if (rc) return rc;
  const emst_config& cfg = ctx->cfg;
  const bool learned = use_filter == EMST_FILTER_LEARNED;
  if (use_filter < 0 || use_filter > EMST_FILTER_LEARNED) return fail(ctx, EMST_E_ARG, "bad filter mode");
  const int T = use_filter == EMST_FILTER_PLAIN ? threshold : INT_MAX;
  const int* dthr = learned ? ctx->d_thr : nullptr;
  const bool host_sync = (cfg.flags & EMST_F_HOST_SYNC) != 0;
  const int guard = (cfg.flags & EMST_F_NO_GUARD) ? 0 : 1;
  const int combine = (cfg.flags & EMST_F_NO_COMBINE) ? 0 : 1;
  const int tail_max = cfg.tail_max < 0 ? 0 : cfg.tail_max;
  const int round_grid = std::max(1, ctx->cus * std::max(1, cfg.grid_per_cu));
  const int init_grid = std::max(1, ctx->cus * 8);
  const int stage_cap = std::max(64, cfg.stage_recs);
  const size_t init_lds = (size_t)kInitWaves * stage_cap * sizeof(int4);
  if (init_lds > ctx->lds_per_block) return fail(ctx, EMST_E_ARG, "stage_recs exceeds the LDS of a workgroup");
  int* d_cnt = ctx->d_cnt;
  long long* d_stats = ctx->d_stats;
  int4** wl = ctx->d_wl;
  ctx->last_nodes = nodes;

  HIPCHK(hipEventRecord(ctx->ev[0], s));
  if (learned)
    launch_learned_threshold(s, nodes, edges, d_nindex, d_eweight, ctx->d_mlp16, ctx->have_model, ctx->d_thr,
                             ctx->d_q, ctx->d_feat, ctx->d_samples);
  {
    const int blocks = (int)(((long long)nodes + 255) / 256);
    hipLaunchKernelGGL(k_setup, dim3(blocks), dim3(256), 0, s, nodes, d_nindex, ctx->d_parent, ctx->d_minv,
                       ctx->d_gsrc, d_cnt, d_stats);
    const long long z16 = ((long long)edges + 15) / 16;
    const int zb = (int)std::min<long long>(std::max<long long>((z16 + 255) / 256, 1), (long long)ctx->cus * 16);
    hipLaunchKernelGGL(k_zero, dim3(zb), dim3(256), 0, s, d_inMST, (long long)edges);
  }
  double init_ms = 0;
  int init_launches = 0;
  int slot = 0;
  for (int phase = 0; phase < (use_filter ? 2 : 1); phase++) {
    if (host_sync && learned) return fail(ctx, EMST_E_ARG, "learned filter needs the device-driven loop");
    if (phase == 1) {
      const int blocks = (int)(((long long)nodes + 255) / 256);
      hipLaunchKernelGGL(k_pick_big, dim3(1), dim3(kWave), 0, s, nodes, ctx->d_parent, &d_cnt[2]);
      hipLaunchKernelGGL(k_flatten, dim3(blocks), dim3(256), 0, s, nodes, ctx->d_parent, d_cnt, &d_cnt[2],
                         ctx->d_bigmap, dthr);
    }
    if (edges > 0) {
      HIPCHK(hipEventRecord(ctx->ev[2 + 2 * phase], s));
      if (phase == 0)
        hipLaunchKernelGGL(k_init_wl<false>, dim3(init_grid), dim3(kInitWaves * kWave), init_lds, s, d_nindex, d_nlist,
                           d_eweight, ctx->d_gsrc, nodes, edges, T, ctx->d_parent, ctx->d_bigmap, wl[0], &d_cnt[0], stage_cap, d_stats,
                           dthr);
      else
        hipLaunchKernelGGL(k_init_wl<true>, dim3(init_grid), dim3(kInitWaves * kWave), init_lds, s, d_nindex, d_nlist,
                           d_eweight, ctx->d_gsrc, nodes, edges, T, ctx->d_parent, ctx->d_bigmap, wl[0], &d_cnt[0], stage_cap, d_stats,
                           dthr);
      HIPCHK(hipEventRecord(ctx->ev[3 + 2 * phase], s));
      HIPCHK(hipGetLastError());
      init_launches++;
    }
    const int rounds_slot = phase == 0 ? ST_ROUNDS0 : ST_ROUNDS1;
    if (host_sync) {
      // Reference schedule: read the size back after every kernel1.
      int cur = 0;
      int size = 0;
      HIPCHK(hipMemcpyAsync(&size, &d_cnt[cur], 4, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (size > 0) {
        const int blocks = (int)std::min<long long>(((long long)size + kSelTile - 1) / kSelTile, round_grid);
        hipLaunchKernelGGL(k_select<true>, dim3(blocks), dim3(kRoundWaves * kWave), 0, s, wl[cur], &d_cnt[cur],
                           wl[cur ^ 1], &d_cnt[cur ^ 1], ctx->d_parent, ctx->d_minv, guard, combine, d_stats,
                           rounds_slot);
      }
      while (size > 0) {
        hipLaunchKernelGGL(k_join_plain, dim3((size + 255) / 256), dim3(256), 0, s, wl[cur], size, ctx->d_parent,
                           ctx->d_minv, d_inMST);
        HIPCHK(hipMemsetAsync(&d_cnt[cur ^ 1], 0, 4, s));
        const int blocks = (int)std::min<long long>(((long long)size + kSelTile - 1) / kSelTile, round_grid);
        hipLaunchKernelGGL(k_select<false>, dim3(blocks), dim3(kRoundWaves * kWave), 0, s, wl[cur], &d_cnt[cur],
                           wl[cur ^ 1], &d_cnt[cur ^ 1], ctx->d_parent, ctx->d_minv, guard, combine, d_stats,
                           rounds_slot);
        cur ^= 1;
        HIPCHK(hipMemcpyAsync(&size, &d_cnt[cur], 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
      }
      HIPCHK(hipMemsetAsync(d_cnt, 0, 8, s));
      continue;
    }
    // Device-driven rounds with a bounded look-ahead.
    const int lag = 2;
    int r = 0;
    int first_slot = slot;
    for (;; r++) {
      if (slot >= kHflagSlots) return fail(ctx, EMST_E_LIMIT, "round limit exceeded");
      // list `cur` holds the survivors of the last select; the phase's
      // first select is the offer-only pass over the initial list
      const int cur = r & 1, nxt = cur ^ 1;
      ctx->h_flag[slot] = -1;
      if (r == 0)
        hipLaunchKernelGGL(k_select<true>, dim3(round_grid), dim3(kRoundWaves * kWave), 0, s, wl[0], &d_cnt[0], wl[1],
                           &d_cnt[1], ctx->d_parent, ctx->d_minv, guard, combine, d_stats, rounds_slot);
      hipLaunchKernelGGL(k_join, dim3(round_grid), dim3(kRoundWaves * kWave), 0, s, wl[cur], &d_cnt[cur], wl[nxt],
                         &d_cnt[nxt], ctx->d_parent, ctx->d_minv, d_inMST, tail_max, ctx->d_flag + slot, d_stats,
                         rounds_slot);
      hipLaunchKernelGGL(k_select<false>, dim3(round_grid), dim3(kRoundWaves * kWave), 0, s, wl[cur], &d_cnt[cur],
                         wl[nxt], &d_cnt[nxt], ctx->d_parent, ctx->d_minv, guard, combine, d_stats, rounds_slot);
      HIPCHK(hipGetLastError());
      slot++;
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
  HIPCHK(hipEventRecord(ctx->ev[1], s));
  HIPCHK(hipEventSynchronize(ctx->ev[1]));
  HIPCHK(hipGetLastError());
  if (stats) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]));
    float tph[2] = {0, 0};
    for (int ph = 0; ph < init_launches; ph++) {
      HIPCHK(hipEventElapsedTime(&tph[ph], ctx->ev[2 + 2 * ph], ctx->ev[3 + 2 * ph]));
      init_ms += tph[ph];
    }
    long long hs[ST_COUNT];
    HIPCHK(hipMemcpy(hs, d_stats, sizeof(hs), hipMemcpyDeviceToHost));
    std::memset(stats, 0, sizeof(*stats));
    if (learned) {
      int thr[2];
      float q;
      HIPCHK(hipMemcpy(thr, ctx->d_thr, 8, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(&q, ctx->d_q, 4, hipMemcpyDeviceToHost));
      stats->threshold_used = thr[0];
      stats->filter_used = thr[1];
      stats->learned_q = q;
    } else {
      stats->threshold_used = T;
      stats->filter_used = use_filter ? 1 : 0;
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
  }
  return EMST_OK;
}

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
  ok = ok && hipMalloc((void**)&ctx->d_cnt, 16) == hipSuccess;
  ok = ok && hipMalloc((void**)&ctx->d_stats, ST_COUNT * sizeof(long long)) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&ctx->h_flag, kHflagSlots * sizeof(int), hipHostMallocMapped) == hipSuccess;
  ok = ok && hipHostGetDevicePointer((void**)&ctx->d_flag, ctx->h_flag, 0) == hipSuccess;
  for (int i = 0; ok && i < 8; i++) ok = hipEventCreate(&ctx->ev[i]) == hipSuccess;
  ok = ok && hipMalloc((void**)&ctx->d_mlp16, kParams * 2) == hipSuccess &&
       hipMalloc((void**)&ctx->d_thr, 8) == hipSuccess && hipMalloc((void**)&ctx->d_q, 4) == hipSuccess &&
       hipMalloc((void**)&ctx->d_feat, kIn * 4) == hipSuccess &&
       hipMalloc((void**)&ctx->d_samples, kWeightSamples * 4) == hipSuccess;
  if (ok) {
    ok = hipFuncSetAttribute((const void*)k_init_wl<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)ctx->lds_per_block) == hipSuccess &&
         hipFuncSetAttribute((const void*)k_init_wl<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)ctx->lds_per_block) == hipSuccess;
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
  void* ptrs[] = {ctx->d_nindex, ctx->d_nlist, ctx->d_eweight, ctx->d_inMST, ctx->d_parent, ctx->d_minv,
                  ctx->d_wl[0],  ctx->d_wl[1], ctx->d_gsrc,    ctx->d_cnt,   ctx->d_stats, ctx->d_bigmap, ctx->d_mlp16, ctx->d_thr, ctx->d_q, ctx->d_feat,
                  ctx->d_samples};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (ctx->h_flag) (void)hipHostFree(ctx->h_flag);
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
  c.flags = cfg->flags;
  if (c.grid_per_cu < 1 || c.grid_per_cu > 64) return fail(ctx, EMST_E_ARG, "grid_per_cu out of range");
  if (c.stage_recs < 64 || (size_t)kInitWaves * c.stage_recs * sizeof(int4) > ctx->lds_per_block)
    return fail(ctx, EMST_E_ARG, "stage_recs out of range");
  if (c.tail_max > (1 << 20)) return fail(ctx, EMST_E_ARG, "tail_max out of range");
  if (c.tile_edges != kGroupEdges) return fail(ctx, EMST_E_ARG, "tile_edges must be 256 in this build");
  if (c.block != kRoundWaves * kWave) return fail(ctx, EMST_E_ARG, "block must be 512 in this build");
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
  if (g->edges > 0 && !g->eweight) return fail(ctx, EMST_E_NOWEIGHT, "graph has no weights");
  int rc = emst_validate(g);
  if (rc) return fail(ctx, rc, "CSR failed validation");
  HIPCHK(hipSetDevice(ctx->device));
  const size_t ni = (size_t)g->nodes + 1, ne = (size_t)g->edges;
  if (ni > ctx->cap_nindex || !ctx->d_nindex) {
    if (ctx->d_nindex) (void)hipFree(ctx->d_nindex);
    ctx->d_nindex = nullptr;
    if (hipMalloc((void**)&ctx->d_nindex, ni * 4) != hipSuccess) return fail(ctx, EMST_E_NOMEM, "hipMalloc nindex");
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
      return fail(ctx, EMST_E_NOMEM, "hipMalloc of the CSR failed");
    ctx->cap_edges = want;
  }
  HIPCHK(hipMemcpyAsync(ctx->d_nindex, g->nindex, ni * 4, hipMemcpyHostToDevice, ctx->stream));
  if (ne) {
    HIPCHK(hipMemcpyAsync(ctx->d_nlist, g->nlist, ne * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_eweight, g->eweight, ne * 4, hipMemcpyHostToDevice, ctx->stream));
  }
  // Worklist capacity: the number of upper-direction edges (the reference
  // assumes a symmetric input and allocates edges/2, ECL_MST.cu:279, 285).
  std::vector<long long> part((size_t)emst::host_threads(), 0);
  emst::parallel_for(g->nodes, [&](int64_t b, int64_t e, int t) {
    long long c = 0;
    for (int64_t v = b; v < e; v++)
      for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) c += g->nlist[j] > v;
    part[t] = c;
  }, (int)part.size());
  long long up = 0;
  for (long long c : part) up += c;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->n = g->nodes;
  ctx->m = g->edges;
  ctx->m_up = up;
  ctx->loaded = true;
  rc = ensure_buffers(ctx, g->nodes, g->edges, up);
  return rc;
}

int emst_solve(emst_ctx* ctx, int use_filter, int32_t threshold, emst_stats* stats)
{
  if (!ctx) return EMST_E_ARG;
  if (!ctx->loaded) return fail(ctx, EMST_E_STATE, "emst_solve before emst_load");
  HIPCHK(hipSetDevice(ctx->device));
  return run_solve(ctx, ctx->n, ctx->m, ctx->d_nindex, ctx->d_nlist, ctx->d_eweight, use_filter, threshold,
                   ctx->d_inMST, ctx->stream, ctx->m_up, stats);
}

int emst_fetch(emst_ctx* ctx, uint8_t* inMST)
{
  if (!ctx || (!inMST && ctx->m > 0)) return EMST_E_ARG;
  if (!ctx->loaded) return fail(ctx, EMST_E_STATE, "emst_fetch before emst_load");
  HIPCHK(hipSetDevice(ctx->device));
  if (ctx->m) HIPCHK(hipMemcpyAsync(inMST, ctx->d_inMST, (size_t)ctx->m, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return EMST_OK;
}

int emst_fetch_parent(emst_ctx* ctx, int32_t* parent)
{
  if (!ctx || !parent) return EMST_E_ARG;
  if (!ctx->d_parent || ctx->last_nodes < 1) return fail(ctx, EMST_E_STATE, "no solve has run");
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipMemcpyAsync(parent, ctx->d_parent, (size_t)ctx->last_nodes * 4, hipMemcpyDeviceToHost, ctx->stream));
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
  if (!params) {
    ctx->have_model = 0;
    return EMST_OK;
  }
  std::vector<uint16_t> h(kParams);
  for (int i = 0; i < kParams; i++) {
    if (!std::isfinite(params[i])) return fail(ctx, EMST_E_ARG, "non-finite model parameter");
    h[i] = emst_bf16_bits(params[i]);
  }
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipMemcpy(ctx->d_mlp16, h.data(), kParams * 2, hipMemcpyHostToDevice));
  ctx->have_model = 1;
  return EMST_OK;
}

int emst_learned_threshold(emst_ctx* ctx, int* use_filter, int32_t* threshold, float* q, float* features,
                           int32_t* samples)
{
  if (!ctx) return EMST_E_ARG;
  if (!ctx->loaded) return fail(ctx, EMST_E_STATE, "emst_learned_threshold before emst_load");
  HIPCHK(hipSetDevice(ctx->device));
  launch_learned_threshold(ctx->stream, ctx->n, ctx->m, ctx->d_nindex, ctx->d_eweight, ctx->d_mlp16, ctx->have_model,
                           ctx->d_thr, ctx->d_q, ctx->d_feat, ctx->d_samples);
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
  return run_solve(ctx, nodes, edges, d_nindex, d_nlist, d_eweight, use_filter, threshold, d_inMST, s, edges, stats);
}

}  // extern "C"
