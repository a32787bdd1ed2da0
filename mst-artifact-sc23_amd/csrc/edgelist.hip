// edgelist.hip -- a second related-work GPU baseline on MI355X: Boruvka over
// a flat edge list, with components found by label propagation, the algorithm
// of the reference's Jucele GPU code (Jucele_GPU/JuceleMST.cu: kernels
// 456-656, timed host loop 191-355).  Per iteration, over the whole edge list:
//   1. every vertex finds its lightest live edge (EncontraMenorAresta1/2,
//      456-530);
//   2. those edges join the forest (MarcarArestas_Strut, Calcula_num_zerodiff,
//      532-562);
//   3. every vertex starts with its own label; each picked edge lowers the
//      larger label of its two ends to the smaller, one pass at a time, until
//      a pass changes nothing (Inicializa_arestasE_C, AtualizaC_1, 564-602,
//      host loop 286-293);
//   4. the vertices that kept their own label become the next iteration's
//      vertices (DefineNovosVU, AtualizaC_3, 604-633);
//   5. every edge is renamed to the new vertices of its ends, and an edge
//      whose ends now share one is retired in place (MarcarArestas, 635-656).
// Edges are never compacted: every iteration reads the whole list.
//
// Kept from the reference: the edge list holds each undirected edge once,
// and ties go to the smaller list position.  The list holds the upper-
// direction CSR positions (nlist[j] > source) in increasing j, so the order
// (weight, list position) is gpuMST's (weight, j): the forest is the same set
// of positions, bit for bit, for any input that gpuMST accepts.
//
// Changed, each for a stated reason:
//  * weights compare as unsigned 32-bit integers packed with the position
//    into one 64-bit key, as in gpuMST; the reference's float costs merge
//    integer weights above 2^24 into ties;
//  * the loop ends when no live edge is left; the reference's test (exactly
//    one mutual pair, 195) never holds on a disconnected graph, where a vertex
//    without an edge also indexes edge -1 (532-541);
//  * retired edges carry -1 instead of the vertex count (648);
//  * the two ends of a mutual pick both mark it (idempotent byte stores)
//    instead of counting picks per edge;
//  * new vertex numbers come from an exclusive scan in vertex order, not
//    from an atomic counter, and vertices left without edges are dropped;
//  * the lightest edge is one guarded 64-bit atomicMin per end instead of a
//    compare-and-swap loop, and a wavefront first combines the offers aimed
//    at its first live lane's target: late iterations send most offers to
//    one surviving vertex, and same-address atomics serialise at about 12 ns
//    each here.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>

#include "emst_internal.hpp"
#include "emst_kernels.hpp"

using namespace emst;

namespace {

constexpr int kEBlock = 256;
constexpr int kMaxIterations = 64;
constexpr long long kMaxLabelPasses = 1 << 20;

inline unsigned eblocks(long long n) { return (unsigned)std::max<long long>(1, (n + kEBlock - 1) / kEBlock); }

struct Buf {
  void* p = nullptr;
  ~Buf()
  {
    if (p) (void)hipFree(p);
  }
};

// Largest v with nindex[v] <= j: the source vertex of CSR position j < m
// (empty adjacencies share their start with the next vertex).
__device__ __forceinline__ int source_of(const int* __restrict__ nindex, int n, int j)
{
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (int)(((long long)lo + hi + 1) >> 1);
    if (nindex[mid] <= j) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// 1 for the upper-direction positions, the ones the list keeps
__global__ __launch_bounds__(kEBlock) void k_eflag(int n, int m, const int* __restrict__ nindex,
                                                   const int* __restrict__ nlist, uint8_t* __restrict__ flag)
{
  const long long j = (long long)blockIdx.x * kEBlock + threadIdx.x;
  if (j >= m) return;
  flag[j] = nlist[j] > source_of(nindex, n, (int)j) ? 1 : 0;
}

// list entry e: its two ends (source, target) and its weight
__global__ __launch_bounds__(kEBlock) void k_ebuild(int n, long long mu, const int* __restrict__ nindex,
                                                    const int* __restrict__ nlist, const int* __restrict__ eweight,
                                                    const int* __restrict__ pos, int2* __restrict__ ends,
                                                    int* __restrict__ wt)
{
  const long long e = (long long)blockIdx.x * kEBlock + threadIdx.x;
  if (e >= mu) return;
  const int j = pos[e];
  ends[e] = make_int2(source_of(nindex, n, j), nlist[j]);
  wt[e] = eweight[j];
}

__device__ __forceinline__ u64 wave_min_u64(u64 k)
{
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const u64 o = __shfl_xor(k, off);
    k = o < k ? o : k;
  }
  return k;
}

// Offers key to minkey[t]: lanes aimed at the first live lane's target
// combine in registers and send one atomic; the rest send their own.
__device__ __forceinline__ void offer(u64* __restrict__ minkey, int t, u64 key, bool live, bool* any)
{
  const u64 act = __ballot(live);
  if (!act) return;
  const int lead = __ffsll((long long)act) - 1;
  const int t0 = __shfl(t, lead);
  const bool same = live && t == t0;
  const u64 k0 = wave_min_u64(same ? key : ~0ull);
  if (lane_id() == lead) {
    if (minkey[t0] > k0) atomicMin(&minkey[t0], k0);
    *any = true;
  }
  if (live && !same && minkey[t] > key) atomicMin(&minkey[t], key);
}

// 1. lightest live edge of every vertex; sets *any when a live edge exists
__global__ __launch_bounds__(kEBlock) void k_emin(long long mu, const int2* __restrict__ ends,
                                                  const int* __restrict__ wt, u64* __restrict__ minkey,
                                                  int* __restrict__ any)
{
  const long long stride = (long long)gridDim.x * kEBlock;
  bool seen = false;
  // whole wavefronts iterate together, so the ballots see every lane
  const long long base = (long long)blockIdx.x * kEBlock + threadIdx.x;
  for (long long r = base - lane_id(); r < mu; r += stride) {
    const long long e = r + lane_id();
    int2 uv = make_int2(-1, -1);
    int w = 0;
    if (e < mu) {
      uv = ends[e];
      if (uv.x >= 0) w = wt[e];
    }
    const bool live = uv.x >= 0;
    const u64 key = make_key(w, (int)e);
    offer(minkey, uv.x, key, live, &seen);
    offer(minkey, uv.y, key, live, &seen);
  }
  if (seen) *any = 1;
}

// 2. + 3. (start): the picked edges join the forest, and every vertex takes
// its own label
__global__ __launch_bounds__(kEBlock) void k_epick(int n, const u64* __restrict__ minkey,
                                                   const int* __restrict__ pos, uint8_t* __restrict__ inMST,
                                                   int* __restrict__ label)
{
  const long long v = (long long)blockIdx.x * kEBlock + threadIdx.x;
  if (v >= n) return;
  const u64 k = minkey[v];
  if (k != ~0ull) inMST[pos[(unsigned)k]] = 1;
  label[v] = (int)v;
}

// 3. one propagation pass: each picked edge lowers its ends' larger label
__global__ __launch_bounds__(kEBlock) void k_eprop(int n, const u64* __restrict__ minkey,
                                                   const int2* __restrict__ ends, int* __restrict__ label,
                                                   int* __restrict__ changed)
{
  const long long v = (long long)blockIdx.x * kEBlock + threadIdx.x;
  if (v >= n) return;
  const u64 k = minkey[v];
  if (k == ~0ull) return;
  const int2 uv = ends[(unsigned)k];
  const int la = __hip_atomic_load(&label[uv.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int lb = __hip_atomic_load(&label[uv.y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (la < lb) {
    atomicMin(&label[uv.y], la);
    *changed = 1;
  } else if (lb < la) {
    atomicMin(&label[uv.x], lb);
    *changed = 1;
  }
}

// 4. vertices that kept their own label and have an edge (n + 1 flags, the
// last 0, for the exclusive scan's total)
__global__ __launch_bounds__(kEBlock) void k_eroots(int n, const u64* __restrict__ minkey,
                                                    const int* __restrict__ label, int* __restrict__ flag)
{
  const long long v = (long long)blockIdx.x * kEBlock + threadIdx.x;
  if (v > n) return;
  flag[v] = v < n && label[v] == (int)v && minkey[v] != ~0ull;
}

// 5. rename every live edge to its ends' new vertices; retire internal ones
__global__ __launch_bounds__(kEBlock) void k_erename(long long mu, const int* __restrict__ label,
                                                     const int* __restrict__ newid, int2* __restrict__ ends)
{
  const long long stride = (long long)gridDim.x * kEBlock;
  for (long long e = (long long)blockIdx.x * kEBlock + threadIdx.x; e < mu; e += stride) {
    const int2 uv = ends[e];
    if (uv.x < 0) continue;
    const int a = newid[label[uv.x]], b = newid[label[uv.y]];
    ends[e] = a == b ? make_int2(-1, -1) : make_int2(a, b);
  }
}

}  // namespace

extern "C" int emst_edgelist_mst(emst_ctx* ctx, const emst_graph* g, uint8_t* inMST, emst_edgelist_stats* stats)
{
  if (!ctx || !g || !inMST) return EMST_E_ARG;
  int rc = emst_load(ctx, g);
  if (rc) return rc;
  const int n0 = g->nodes, m0 = g->edges;
  hipStream_t s = ctx->stream;
  auto fail = [&](hipError_t e, const char* what) {
    char b[256];
    std::snprintf(b, sizeof(b), "HIP error %d in %s: %s", (int)e, what, hipGetErrorString(e));
    return ctx_fail(ctx, EMST_E_HIP, b);
  };
#define ECHK(call)                                \
  do {                                            \
    hipError_t e_ = (call);                       \
    if (e_ != hipSuccess) return fail(e_, #call); \
  } while (0)
  const size_t nv = (size_t)n0 + 1, ne = (size_t)std::max(m0, 1);
  const long long mu = ctx->m_up;
  const size_t nl = (size_t)std::max<long long>(mu, 1);
  Buf b_flag8, b_pos, b_ends, b_wt, b_cnt, b_minkey, b_label, b_vflag, b_newid, b_sel_tmp, b_scan_tmp;
  if (hipMalloc(&b_flag8.p, ne) || hipMalloc(&b_pos.p, nl * 4) || hipMalloc(&b_ends.p, nl * sizeof(int2)) ||
      hipMalloc(&b_wt.p, nl * 4) || hipMalloc(&b_cnt.p, 16) || hipMalloc(&b_minkey.p, nv * 8) ||
      hipMalloc(&b_label.p, nv * 4) || hipMalloc(&b_vflag.p, nv * 4) || hipMalloc(&b_newid.p, nv * 4))
    return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of the edge list failed");
  int* cnt = (int*)b_cnt.p;  // [0] selected count, [1] any live edge, [2] labels changed
  int* pos = (int*)b_pos.p;
  int2* ends = (int2*)b_ends.p;
  int* wt = (int*)b_wt.p;
  u64* minkey = (u64*)b_minkey.p;
  int* label = (int*)b_label.p;
  int* vflag = (int*)b_vflag.p;
  int* newid = (int*)b_newid.p;

  // the edge list, in increasing CSR position (outside the timed loop, as
  // the reference reads its list from a file before its timer starts, 134-191)
  const hipcub::CountingInputIterator<int> counting(0);
  size_t sel_bytes = 0, scan_bytes = 0;
  ECHK(hipcub::DeviceSelect::Flagged(nullptr, sel_bytes, counting, (const uint8_t*)b_flag8.p, pos, cnt,
                                     std::max(m0, 1), s));
  ECHK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, vflag, newid, (int)nv, s));
  if (hipMalloc(&b_sel_tmp.p, std::max<size_t>(sel_bytes, 16)) || hipMalloc(&b_scan_tmp.p, std::max<size_t>(scan_bytes, 16)))
    return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of scan storage failed");
  if (m0 > 0) {
    ECHK(hipMemsetAsync(ctx->d_inMST, 0, (size_t)m0, s));
    hipLaunchKernelGGL(k_eflag, dim3(eblocks(m0)), dim3(kEBlock), 0, s, n0, m0, (const int*)ctx->d_nindex,
                       (const int*)ctx->d_nlist, (uint8_t*)b_flag8.p);
    ECHK(hipcub::DeviceSelect::Flagged(b_sel_tmp.p, sel_bytes, counting, (const uint8_t*)b_flag8.p, pos, cnt, m0, s));
    int got = 0;
    ECHK(hipMemcpyAsync(&got, cnt, 4, hipMemcpyDeviceToHost, s));
    ECHK(hipStreamSynchronize(s));
    if (got != mu) return ctx_fail(ctx, EMST_E_HIP, "edge list size differs from the upper-direction count");
    if (mu > 0)
      hipLaunchKernelGGL(k_ebuild, dim3(eblocks(mu)), dim3(kEBlock), 0, s, n0, mu, (const int*)ctx->d_nindex,
                         (const int*)ctx->d_nlist, (const int*)ctx->d_eweight, (const int*)pos, ends, wt);
    ECHK(hipGetLastError());
  }
  ECHK(hipStreamSynchronize(s));

  const int grid = (int)std::min<long long>(eblocks(mu), (long long)ctx->cus * 16);
  ECHK(hipEventRecord(ctx->ev[0], s));
  int n = n0, it = 0;
  long long passes = 0;
  while (mu > 0 && n > 1) {
    if (++it > kMaxIterations) return ctx_fail(ctx, EMST_E_LIMIT, "edge-list Boruvka did not converge");
    ECHK(hipMemsetAsync(minkey, 0xff, (size_t)n * 8, s));
    ECHK(hipMemsetAsync(cnt + 1, 0, 4, s));
    hipLaunchKernelGGL(k_emin, dim3(grid), dim3(kEBlock), 0, s, mu, (const int2*)ends, (const int*)wt, minkey, cnt + 1);
    int any = 0;
    ECHK(hipMemcpyAsync(&any, cnt + 1, 4, hipMemcpyDeviceToHost, s));
    ECHK(hipStreamSynchronize(s));
    if (!any) {  // nothing left to join: this pass does not count
      it--;
      break;
    }
    hipLaunchKernelGGL(k_epick, dim3(eblocks(n)), dim3(kEBlock), 0, s, n, (const u64*)minkey, (const int*)pos,
                       ctx->d_inMST, label);
    for (;;) {
      if (++passes > kMaxLabelPasses) return ctx_fail(ctx, EMST_E_LIMIT, "labels did not converge");
      ECHK(hipMemsetAsync(cnt + 2, 0, 4, s));
      hipLaunchKernelGGL(k_eprop, dim3(eblocks(n)), dim3(kEBlock), 0, s, n, (const u64*)minkey, (const int2*)ends,
                         label, cnt + 2);
      int h = 0;
      ECHK(hipMemcpyAsync(&h, cnt + 2, 4, hipMemcpyDeviceToHost, s));
      ECHK(hipStreamSynchronize(s));
      if (!h) break;
    }
    hipLaunchKernelGGL(k_eroots, dim3(eblocks((long long)n + 1)), dim3(kEBlock), 0, s, n, (const u64*)minkey,
                       (const int*)label, vflag);
    ECHK(hipcub::DeviceScan::ExclusiveSum(b_scan_tmp.p, scan_bytes, vflag, newid, n + 1, s));
    int nn = 0;
    ECHK(hipMemcpyAsync(&nn, newid + n, 4, hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(k_erename, dim3(grid), dim3(kEBlock), 0, s, mu, (const int*)label, (const int*)newid, ends);
    ECHK(hipGetLastError());
    ECHK(hipStreamSynchronize(s));
    n = nn;
  }
  ECHK(hipEventRecord(ctx->ev[1], s));
  ECHK(hipEventSynchronize(ctx->ev[1]));
  float ms = 0;
  ECHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  if (m0 > 0) ECHK(hipMemcpy(inMST, ctx->d_inMST, (size_t)m0, hipMemcpyDeviceToHost));
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->ms = ms;
    stats->iterations = it;
    stats->label_passes = passes;
    for (int32_t j = 0; j < m0; j++)
      if (inMST[j]) {
        stats->mst_edges++;
        stats->weight += (uint64_t)(uint32_t)g->eweight[j];
      }
  }
#undef ECHK
  return EMST_OK;
}
