// cpu_mst.cpp -- host minimum spanning forest and result comparison used by
// the ecl_mst CLI, replacing cpuMST (ECL-MST/ECL_MST.cu:82-124) and verify
// (ECL_MST.cu:365-386).
//
// Kruskal over upper-direction edges in (signed weight, j) order: the
// reference sorts std::tuple<weight, j, from, to>, and j is unique, so the
// tuple order is exactly (weight, j).  Keys are built in parallel and
// sorted as 64-bit integers; union-find uses path halving and union by
// smaller id.

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/emst.h"
#include "host_util.hpp"

namespace {

int32_t uf_find(std::vector<int32_t>& p, int32_t v)
{
  while (p[v] != v) {
    p[v] = p[p[v]];
    v = p[v];
  }
  return v;
}

// Parallel sort: sort contiguous slices, then merge pairwise.
void psort(std::vector<uint64_t>& a)
{
  const int W = std::max(1, std::min(emst::host_threads(), 64));
  const int64_t n = (int64_t)a.size();
  if (n < (1 << 16) || W == 1) {
    std::sort(a.begin(), a.end());
    return;
  }
  std::vector<int64_t> cut(W + 1);
  for (int t = 0; t <= W; t++) cut[t] = n * t / W;
  emst::parallel_for(W, [&](int64_t b, int64_t e, int) {
    for (int64_t t = b; t < e; t++) std::sort(a.begin() + cut[t], a.begin() + cut[t + 1]);
  }, W);
  std::vector<uint64_t> tmp(a.size());
  for (int width = 1; width < W; width *= 2) {
    std::vector<int> starts;
    for (int t = 0; t < W; t += 2 * width) starts.push_back(t);
    emst::parallel_for((int64_t)starts.size(), [&](int64_t b, int64_t e, int) {
      for (int64_t k = b; k < e; k++) {
        const int t = starts[k];
        const int mid = std::min(t + width, W), hi = std::min(t + 2 * width, W);
        std::merge(a.begin() + cut[t], a.begin() + cut[mid], a.begin() + cut[mid], a.begin() + cut[hi],
                   tmp.begin() + cut[t]);
      }
    }, (int)starts.size());
    a.swap(tmp);
  }
}

}  // namespace

extern "C" {

int emst_cpu_mst(const emst_graph* g, uint8_t* inMST)
{
  if (!g || !g->nindex || g->nodes < 1 || g->edges < 0 || (g->edges > 0 && (!inMST || !g->nlist))) return EMST_E_ARG;
  if (g->edges > 0 && !g->eweight) return EMST_E_NOWEIGHT;
  const int32_t n = g->nodes;
  std::memset(inMST, 0, (size_t)g->edges);
  const int W = emst::host_threads();
  std::vector<int64_t> cnt((size_t)W + 1, 0);
  emst::parallel_for(n, [&](int64_t b, int64_t e, int t) {
    int64_t c = 0;
    for (int64_t v = b; v < e; v++)
      for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) c += g->nlist[j] > v;
    cnt[t + 1] = c;
  }, W);
  for (int t = 0; t < W; t++) cnt[t + 1] += cnt[t];
  std::vector<uint64_t> keys((size_t)cnt[W]);
  emst::parallel_for(n, [&](int64_t b, int64_t e, int t) {
    int64_t k = cnt[t];
    for (int64_t v = b; v < e; v++)
      for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++)
        if (g->nlist[j] > v) keys[k++] = ((uint64_t)((uint32_t)g->eweight[j] ^ 0x80000000u) << 32) | (uint32_t)j;
  }, W);
  psort(keys);
  std::vector<int32_t> src((size_t)std::max(g->edges, 1));
  emst::parallel_for(n, [&](int64_t b, int64_t e, int) {
    for (int64_t v = b; v < e; v++)
      for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) src[j] = (int32_t)v;
  });
  std::vector<int32_t> parent((size_t)n);
  for (int32_t v = 0; v < n; v++) parent[v] = v;
  int64_t remaining = (int64_t)n - 1;
  for (size_t k = 0; k < keys.size() && remaining > 0; k++) {
    const int32_t j = (int32_t)(uint32_t)keys[k];
    const int32_t a = uf_find(parent, src[j]), b = uf_find(parent, g->nlist[j]);
    if (a != b) {
      inMST[j] = 1;
      if (a < b) parent[b] = a; else parent[a] = b;
      remaining--;
    }
  }
  return EMST_OK;
}

int emst_verify(const emst_graph* g, const uint8_t* cpu, const uint8_t* gpu, uint64_t* cpu_weight,
                uint64_t* gpu_weight)
{
  if (!g || (g->edges > 0 && (!cpu || !gpu || !g->eweight))) return EMST_E_ARG;
  uint64_t cw = 0, gw = 0;
  int64_t only_cpu = 0, only_gpu = 0;
  for (int32_t j = 0; j < g->edges; j++) {
    const bool c = cpu[j] != 0, d = gpu[j] != 0;
    only_cpu += c && !d;
    only_gpu += d && !c;
    if (d) gw += (uint64_t)(int64_t)g->eweight[j];
    if (c) cw += (uint64_t)(int64_t)g->eweight[j];
  }
  if (cpu_weight) *cpu_weight = cw;
  if (gpu_weight) *gpu_weight = gw;
  return (cw != gw || only_cpu || only_gpu) ? 1 : 0;
}

}  // extern "C"
