// relabel.hip -- renumbered layouts of a loaded graph.
//
// The solver's vertex-indexed arrays (parent, minv, the class map, the
// label and member flags) are gathered at every record's endpoints in every
// round, so the vertex numbering decides how many lines those gathers touch.
// emst_load therefore rebuilds the CSR of some graphs in a new numbering,
// with every list sorted by the new ids, and solves on that copy:
//  - degree order, for skewed degrees (RMAT).  Most records touch a few
//    hundred thousand hubs, but the caller's numbering scatters them over
//    the whole id range, so each gather fetches a line that holds one hot
//    entry.  Renumbered by decreasing degree (ties keep the caller's order),
//    the hubs' entries share lines and stay in L2; their lists also come
//    first, so the largest light-phase component (the hubs) owns a
//    contiguous run of positions that the sparse heavy initialisation skips
//    whole.
//  - cluster order, for scattered ids on a graph of high diameter (a
//    geometric graph numbered at random).  A breadth-first search from every
//    4096th vertex at once grows clusters of nearby vertices; numbered by
//    (cluster, level), most edges join ids of one cluster.
//
// Results are unchanged.  The reference keys an edge by its weight and the
// position j of its upper-direction copy (ECL_MST.cu:173-176, 222-223) and
// marks inMST[j]; with distinct keys the forest of every phase is the
// unique minimum forest under that order, whatever the vertex numbering.
// The new layout keeps, per position, its edge's key id (j in the caller's
// CSR: the position itself for an upper-direction copy, the reverse
// position for a lower one) and that copy's weight, so records, offers and
// inMST use the caller's keys and indices.  The passes that make the first
// offers read the key ids with their task; the plain streaming passes emit
// positions, and the phase's first select replaces each record's position by
// its key id before any key is formed (a gather there overlaps the select's
// own gathers; in the streaming pass it stalled the stream).  Both copies of
// an edge carry its upper copy's weight, the one the reference reads.  Root
// ids are the new numbering's: emst_fetch_parent maps them back.
//
// Built on the device from the uploaded CSR and its reverse positions: the
// vertex order (a radix sort of degrees, or the cluster search and a radix
// sort of its keys), then one 64-bit radix sort of (new source, new
// neighbour) keys with the caller's positions as values.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>

#include "emst_internal.hpp"

namespace {

using emst::u64;

// Degrees and vertex ids, the input of the degree sort.
__global__ __launch_bounds__(256) void k_rl_deg(int n, const int* __restrict__ nindex, int* __restrict__ deg,
                                                int* __restrict__ ids)
{
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v < n) {
    deg[v] = nindex[v + 1] - nindex[v];
    ids[v] = v;
  }
}

// newid is the inverse of order; ndeg[n] = 0 closes the scan to the new
// offsets.
__global__ __launch_bounds__(256) void k_rl_newid(int n, const int* __restrict__ order, int* __restrict__ newid,
                                                  int* __restrict__ ndeg)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) newid[order[i]] = i;
  if (i == 0) ndeg[n] = 0;
}

constexpr int kKeyPer = 1024;  // positions per workgroup of k_rl_keys

// key[p] = newid[src(p)] << B | newid[nlist[p]], val[p] = p.  The workgroup
// finds the source range of its positions once; each position then
// searches only that range of offsets.
__global__ __launch_bounds__(256) void k_rl_keys(int n, int m, int B, const int* __restrict__ nindex,
                                                 const int* __restrict__ nlist, const int* __restrict__ newid,
                                                 u64* __restrict__ key, int* __restrict__ val)
{
  __shared__ int s_lo, s_hi;
  const long long p0 = (long long)blockIdx.x * kKeyPer;
  // source of position p: the largest v in [lo, hi] with nindex[v] <= p
  auto src_of = [&](long long p, int lo, int hi) {
    while (lo < hi) {
      const int mid = (int)(((long long)lo + hi + 1) >> 1);
      if (nindex[mid] <= p) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  };
  if (threadIdx.x == 0) s_lo = src_of(p0, 0, n - 1);
  if (threadIdx.x == 1) {
    const long long p1 = p0 + kKeyPer < m ? p0 + kKeyPer - 1 : (long long)m - 1;
    s_hi = src_of(p1, 0, n - 1);
  }
  __syncthreads();
  const int lo = s_lo, hi = s_hi;
  for (int k = threadIdx.x; k < kKeyPer; k += 256) {
    const long long p = p0 + k;
    if (p >= m) break;
    const int v = src_of(p, lo, hi);
    key[p] = ((u64)newid[v] << B) | (u64)(unsigned)newid[nlist[p]];
    val[p] = (int)p;
  }
}

// The new lists from the sorted keys: neighbour, key id and weight per new
// position.
__global__ __launch_bounds__(256) void k_rl_fill(int m, int B, const u64* __restrict__ key,
                                                 const int* __restrict__ val, const int* __restrict__ order,
                                                 const int* __restrict__ nlist, const int* __restrict__ eweight,
                                                 const int* __restrict__ rev, int* __restrict__ nl2,
                                                 int* __restrict__ w2, int* __restrict__ kid)
{
  const long long stride = (long long)gridDim.x * 256;
  const u64 mask = (1ull << B) - 1;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < m; q += stride) {
    const u64 k = key[q];
    const int po = val[q];
    const int vo = order[(int)(k >> B)];
    const int yo = nlist[po];
    const int j = yo < vo ? rev[po] : po;  // a self-loop keeps its own position (it is never a record)
    nl2[q] = (int)(k & mask);
    kid[q] = j;
    w2[q] = eweight[j];
  }
}

// A parent array in the new numbering, in the caller's: out[v] is the
// caller's id of parent[newid[v]].
__global__ __launch_bounds__(256) void k_rl_parent(int n, const int* __restrict__ parent,
                                                   const int* __restrict__ newid, const int* __restrict__ order,
                                                   int* __restrict__ out)
{
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v < n) out[v] = order[parent[newid[v]]];
}

// Cluster order: a breadth-first search from every kClusterStride-th vertex
// at once.  A vertex takes the smallest seed label among its neighbours
// reached one level earlier (push, then take: deterministic whatever the
// schedule); vertices are then ordered by (label, level, caller id).
constexpr int kClusterStride = 4096;
constexpr int kUnreached = INT_MAX;
constexpr int kMaxLevels = 1 << 16;

__global__ __launch_bounds__(256) void k_cl_init(int n, int* __restrict__ lab, int* __restrict__ lev,
                                                 int* __restrict__ cand)
{
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v < n) {
    const bool seed = v % kClusterStride == 0;
    lab[v] = seed ? v / kClusterStride : kUnreached;
    lev[v] = seed ? 0 : kUnreached;
    cand[v] = kUnreached;
  }
}

__global__ __launch_bounds__(256) void k_cl_push(int n, int t, const int* __restrict__ nindex,
                                                 const int* __restrict__ nlist, const int* __restrict__ lab,
                                                 const int* __restrict__ lev, int* __restrict__ cand)
{
  const int u = blockIdx.x * 256 + threadIdx.x;
  if (u >= n || lev[u] != t - 1) return;
  const int l = lab[u];
  for (int p = nindex[u]; p < nindex[u + 1]; p++) {
    const int v = nlist[p];
    if (lev[v] == kUnreached) atomicMin(&cand[v], l);
  }
}

__global__ __launch_bounds__(256) void k_cl_take(int n, int t, int* __restrict__ lab, int* __restrict__ lev,
                                                 const int* __restrict__ cand, int* __restrict__ count)
{
  const int v = blockIdx.x * 256 + threadIdx.x;
  bool took = false;
  if (v < n && lev[v] == kUnreached && cand[v] != kUnreached) {
    lab[v] = cand[v];
    lev[v] = t;
    took = true;
  }
  const u64 b = __ballot(took);
  if (emst::lane_id() == 0 && b) atomicAdd(count, __popcll(b));
}

// Sort keys of the cluster order: (label, level); unreached vertices last.
__global__ __launch_bounds__(256) void k_cl_keys(int n, int nseeds, const int* __restrict__ lab,
                                                 const int* __restrict__ lev, u64* __restrict__ key,
                                                 int* __restrict__ ids)
{
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v < n) {
    const bool r = lev[v] != kUnreached;
    key[v] = ((u64)(r ? lab[v] : nseeds) << 32) | (u64)(unsigned)(r ? lev[v] : 0);
    ids[v] = v;
  }
}

// Degrees in the new order.
__global__ __launch_bounds__(256) void k_rl_ndeg(int n, const int* __restrict__ order, const int* __restrict__ nindex,
                                                 int* __restrict__ ndeg)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) ndeg[i] = nindex[order[i] + 1] - nindex[order[i]];
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf()
  {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

namespace emst {

void free_relabel(Relabel& rl)
{
  for (int* q : {rl.nindex, rl.nlist, rl.eweight, rl.kid, rl.order, rl.newid})
    if (q) (void)hipFree(q);
  rl = Relabel{};
}

int relabel_parent(emst_ctx* ctx, const int* parent, int32_t nodes, int32_t* out, hipStream_t s)
{
  const Relabel& rl = ctx->rl;
  if (!rl.ok || (size_t)nodes > rl.cap_nodes) return ctx_fail(ctx, EMST_E_STATE, "no renumbered layout");
  DevBuf tmp;
  if (hipMalloc(&tmp.p, (size_t)std::max(nodes, 1) * 4)) return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc failed");
  hipLaunchKernelGGL(k_rl_parent, dim3((nodes + 255) / 256), dim3(256), 0, s, nodes, parent, rl.newid, rl.order,
                     (int*)tmp.p);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(out, tmp.p, (size_t)nodes * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return ctx_fail(ctx, EMST_E_HIP, "mapping the parent array back failed");
  return EMST_OK;
}

int build_relabel(emst_ctx* ctx, int32_t nodes, int32_t edges, const int* d_nindex, const int* d_nlist,
                  const int* d_eweight, const int* rev, hipStream_t s, int mode, int min_levels)
{
  Relabel& rl = ctx->rl;
  rl.ok = false;
  rl.mode = 0;
  rl.levels = 0;
  if (nodes < 2 || edges < 1) return EMST_OK;
  const auto t0 = std::chrono::steady_clock::now();
  auto fail = [&](hipError_t e, const char* what) {
    char b[256];
    std::snprintf(b, sizeof(b), "HIP error %d in %s: %s", (int)e, what, hipGetErrorString(e));
    return ctx_fail(ctx, EMST_E_HIP, b);
  };
#define RLCHK(call)                               \
  do {                                            \
    hipError_t e_ = (call);                       \
    if (e_ != hipSuccess) return fail(e_, #call); \
  } while (0)
  const size_t n = (size_t)nodes, m = (size_t)edges;
  if (n > rl.cap_nodes || m > rl.cap_edges || !rl.nindex) {
    free_relabel(rl);
    if (hipMalloc((void**)&rl.nindex, (n + 1) * 4) || hipMalloc((void**)&rl.order, n * 4) ||
        hipMalloc((void**)&rl.newid, n * 4) || hipMalloc((void**)&rl.nlist, m * 4) ||
        hipMalloc((void**)&rl.eweight, m * 4) || hipMalloc((void**)&rl.kid, m * 4)) {
      free_relabel(rl);
      return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of the renumbered layout failed");
    }
    rl.cap_nodes = n;
    rl.cap_edges = m;
  }
  // bits of a vertex id
  int B = 1;
  while (B < 31 && ((size_t)1 << B) < n) B++;
  const int nseeds = (int)((n + kClusterStride - 1) / kClusterStride);
  int SB = 1;  // bits of a cluster label (unreached vertices take nseeds)
  while (SB < 31 && ((size_t)1 << SB) <= (size_t)nseeds) SB++;
  DevBuf deg, ids, ndeg, k0, k1, v0, v1, tmp, lab, lev, cand, vk0, vk1, cnt;
  if (hipMalloc(&deg.p, n * 4) || hipMalloc(&ids.p, n * 4) || hipMalloc(&ndeg.p, (n + 1) * 4) ||
      hipMalloc(&k0.p, m * 8) || hipMalloc(&k1.p, m * 8) || hipMalloc(&v0.p, m * 4) || hipMalloc(&v1.p, m * 4))
    return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of the relabelling buffers failed");
  if (mode == kOrderCluster && (hipMalloc(&lab.p, n * 4) || hipMalloc(&lev.p, n * 4) || hipMalloc(&cand.p, n * 4) ||
                                hipMalloc(&vk0.p, n * 8) || hipMalloc(&vk1.p, n * 8) || hipMalloc(&cnt.p, 4)))
    return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of the cluster search buffers failed");
  const int nb = (int)((n + 255) / 256);
  hipLaunchKernelGGL(k_rl_deg, dim3(nb), dim3(256), 0, s, nodes, d_nindex, (int*)deg.p, (int*)ids.p);
  RLCHK(hipGetLastError());
  size_t b_deg = 0, b_scan = 0, b_key = 0, b_cl = 0;
  hipcub::DoubleBuffer<u64> dk((u64*)k0.p, (u64*)k1.p);
  hipcub::DoubleBuffer<int> dv((int*)v0.p, (int*)v1.p);
  RLCHK(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b_deg, (const int*)deg.p, (int*)ndeg.p,
                                                     (const int*)ids.p, rl.order, nodes, 0, 32, s));
  RLCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, b_scan, (const int*)ndeg.p, rl.nindex, nodes + 1, s));
  RLCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, b_key, dk, dv, edges, 0, 2 * B, s));
  if (mode == kOrderCluster)
    RLCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, b_cl, (const u64*)vk0.p, (u64*)vk1.p, (const int*)ids.p,
                                             rl.order, nodes, 0, 32 + SB, s));
  if (hipMalloc(&tmp.p, std::max({b_deg, b_scan, b_key, b_cl, (size_t)16})))
    return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of the sort storage failed");
  if (mode == kOrderCluster) {
    // breadth-first levels from all seeds at once; one host read per level
    hipLaunchKernelGGL(k_cl_init, dim3(nb), dim3(256), 0, s, nodes, (int*)lab.p, (int*)lev.p, (int*)cand.p);
    RLCHK(hipGetLastError());
    int t = 1;
    for (;; t++) {
      int got = 0;
      RLCHK(hipMemsetAsync(cnt.p, 0, 4, s));
      hipLaunchKernelGGL(k_cl_push, dim3(nb), dim3(256), 0, s, nodes, t, d_nindex, d_nlist, (const int*)lab.p,
                         (const int*)lev.p, (int*)cand.p);
      hipLaunchKernelGGL(k_cl_take, dim3(nb), dim3(256), 0, s, nodes, t, (int*)lab.p, (int*)lev.p,
                         (const int*)cand.p, (int*)cnt.p);
      RLCHK(hipGetLastError());
      RLCHK(hipMemcpyAsync(&got, cnt.p, 4, hipMemcpyDeviceToHost, s));
      RLCHK(hipStreamSynchronize(s));
      if (got == 0 || t == kMaxLevels) break;  // (vertices still unreached go last)
    }
    rl.levels = t - 1;
    // a low-diameter graph gains nothing from clusters: keep the caller's numbering
    if (rl.levels < min_levels) return EMST_OK;
    hipLaunchKernelGGL(k_cl_keys, dim3(nb), dim3(256), 0, s, nodes, nseeds, (const int*)lab.p, (const int*)lev.p,
                       (u64*)vk0.p, (int*)ids.p);
    RLCHK(hipGetLastError());
    // stable: within a level of a cluster, the caller's order
    RLCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, b_cl, (const u64*)vk0.p, (u64*)vk1.p, (const int*)ids.p,
                                             rl.order, nodes, 0, 32 + SB, s));
    hipLaunchKernelGGL(k_rl_ndeg, dim3(nb), dim3(256), 0, s, nodes, rl.order, d_nindex, (int*)ndeg.p);
    RLCHK(hipGetLastError());
  } else {
    // vertices by decreasing degree; the radix sort is stable, so equal
    // degrees keep the caller's order
    RLCHK(hipcub::DeviceRadixSort::SortPairsDescending(tmp.p, b_deg, (const int*)deg.p, (int*)ndeg.p,
                                                       (const int*)ids.p, rl.order, nodes, 0, 32, s));
  }
  hipLaunchKernelGGL(k_rl_newid, dim3(nb), dim3(256), 0, s, nodes, rl.order, rl.newid, (int*)ndeg.p);
  RLCHK(hipGetLastError());
  // offsets of the new lists (ndeg holds the degrees in the new order)
  RLCHK(hipcub::DeviceScan::ExclusiveSum(tmp.p, b_scan, (const int*)ndeg.p, rl.nindex, nodes + 1, s));
  const long long kb = ((long long)edges + kKeyPer - 1) / kKeyPer;
  hipLaunchKernelGGL(k_rl_keys, dim3((unsigned)kb), dim3(256), 0, s, nodes, edges, B, d_nindex, d_nlist, rl.newid,
                     (u64*)k0.p, (int*)v0.p);
  RLCHK(hipGetLastError());
  RLCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, b_key, dk, dv, edges, 0, 2 * B, s));
  const int fg = (int)std::min<long long>(((long long)edges + 255) / 256, (long long)ctx->cus * 32);
  hipLaunchKernelGGL(k_rl_fill, dim3(fg), dim3(256), 0, s, edges, B, (const u64*)dk.Current(),
                     (const int*)dv.Current(), rl.order, d_nlist, d_eweight, rev, rl.nlist, rl.eweight, rl.kid);
  RLCHK(hipGetLastError());
  RLCHK(hipStreamSynchronize(s));
#undef RLCHK
  rl.ok = true;
  rl.mode = mode;
  rl.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return EMST_OK;
}

}  // namespace emst
