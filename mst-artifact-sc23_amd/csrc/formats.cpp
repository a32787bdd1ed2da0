// formats.cpp -- writers of the graph formats the reference's external
// baselines read, from an ECLgraph CSR.  The reference converts with one
// small program per baseline (ecl2mtx_genweights_cugraph.cpp,
// ecl2pbbs_genweights.cpp, ecl2lsg_genweights.cpp,
// ecl2UMinho_genweights.cpp); these writers produce the same file layouts,
// so that a baseline built elsewhere reads the same graph the solver solves.
// Weights must be present (emst_synth_weights adds the reference's
// synthetic ones to a weightless input first).
#include <cinttypes>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/emst.h"

namespace {

bool valid(const emst_graph* g)
{
  return g && g->nodes >= 1 && g->edges >= 0 && g->nindex && (g->edges == 0 || (g->nlist && g->eweight));
}

// Buffered text output: one fwrite per megabyte instead of one fprintf per
// number, and a sticky error flag.
struct TextOut {
  FILE* f = nullptr;
  std::vector<char> buf;
  bool ok = true;
  explicit TextOut(FILE* file) : f(file) { buf.reserve(1 << 20); }
  void put(const char* s, size_t n)
  {
    buf.insert(buf.end(), s, s + n);
    if (buf.size() >= (1u << 20)) flush();
  }
  void printf_(const char* fmt, ...) __attribute__((format(printf, 2, 3)))
  {
    char tmp[128];
    va_list ap;
    va_start(ap, fmt);
    const int n = std::vsnprintf(tmp, sizeof(tmp), fmt, ap);
    va_end(ap);
    if (n < 0 || n >= (int)sizeof(tmp)) {
      ok = false;
      return;
    }
    put(tmp, (size_t)n);
  }
  void flush()
  {
    if (!buf.empty() && std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) ok = false;
    buf.clear();
  }
  int close()
  {
    flush();
    if (std::fclose(f) != 0) ok = false;
    return ok ? EMST_OK : EMST_E_WRITE;
  }
};

}  // namespace

extern "C" {

// MatrixMarket coordinate, real, symmetric: one entry per undirected edge,
// the lower-triangle copy (row > column), 1-based, weight with one decimal
// (the layout of ecl2mtx_genweights_cugraph.cpp).
int emst_write_mtx(const char* path, const emst_graph* g)
{
  if (!path || !valid(g)) return EMST_E_ARG;
  FILE* f = std::fopen(path, "wt");
  if (!f) return EMST_E_OPEN;
  TextOut o(f);
  o.printf_("%%%%MatrixMarket matrix coordinate real symmetric\n");
  o.printf_("%d %d %d\n", g->nodes, g->nodes, g->edges / 2);
  for (int32_t v = 0; v < g->nodes; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++)
      if (v > g->nlist[j]) o.printf_("%d %d %.1f\n", v + 1, g->nlist[j] + 1, (double)(float)g->eweight[j]);
  return o.close();
}

// PBBS WeightedEdgeArray: every directed CSR edge as "src dst weight" with
// the weight printed as a float (ecl2pbbs_genweights.cpp).
int emst_write_pbbs(const char* path, const emst_graph* g)
{
  if (!path || !valid(g)) return EMST_E_ARG;
  FILE* f = std::fopen(path, "wt");
  if (!f) return EMST_E_OPEN;
  TextOut o(f);
  o.printf_("WeightedEdgeArray\n");
  for (int32_t v = 0; v < g->nodes; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++)
      o.printf_("%d %d %f\n", v, g->nlist[j], (double)(float)g->eweight[j]);
  return o.close();
}

// Galois/Lonestar binary graph, version 1 (ecl2lsg_genweights.cpp): u64
// version = 1, u64 edge-data size = 4, u64 nodes, u64 edges, u64 end offset
// of every node's adjacency (nindex[1..nodes]), u32 neighbours, a u32 pad
// when the edge count is odd, then u32 weights.
int emst_write_galois(const char* path, const emst_graph* g)
{
  if (!path || !valid(g)) return EMST_E_ARG;
  FILE* f = std::fopen(path, "wb");
  if (!f) return EMST_E_OPEN;
  const uint64_t head[4] = {1, 4, (uint64_t)g->nodes, (uint64_t)g->edges};
  bool ok = std::fwrite(head, 8, 4, f) == 4;
  std::vector<uint64_t> ends((size_t)g->nodes);
  for (int32_t v = 0; v < g->nodes; v++) ends[(size_t)v] = (uint64_t)g->nindex[v + 1];
  ok = ok && std::fwrite(ends.data(), 8, ends.size(), f) == ends.size();
  const size_t ne = (size_t)g->edges;
  ok = ok && (ne == 0 || std::fwrite(g->nlist, 4, ne, f) == ne);
  if (ok && (ne & 1)) {
    const uint32_t pad = 0;
    ok = std::fwrite(&pad, 4, 1, f) == 1;
  }
  ok = ok && (ne == 0 || std::fwrite(g->eweight, 4, ne, f) == ne);
  if (std::fclose(f) != 0) ok = false;
  return ok ? EMST_OK : EMST_E_WRITE;
}

// BoruvkaUMinho text edge list (ecl2UMinho_genweights.cpp): "nodes edges",
// then every directed CSR edge as "src dst weight".
int emst_write_uminho(const char* path, const emst_graph* g)
{
  if (!path || !valid(g)) return EMST_E_ARG;
  FILE* f = std::fopen(path, "wt");
  if (!f) return EMST_E_OPEN;
  TextOut o(f);
  o.printf_("%d %d\n", g->nodes, g->edges);
  for (int32_t v = 0; v < g->nodes; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) o.printf_("%d %d %d\n", v, g->nlist[j], g->eweight[j]);
  return o.close();
}

// Jucele edge list (the layout JuceleMST.cu:410-440 reads): the vertex
// count and the edge count on two lines, then each undirected edge once, as
// "src dst weight", here the upper-direction CSR positions (dst > src) in
// order, the edges gpuMST and emst_edgelist_mst work on.
int emst_write_jucele(const char* path, const emst_graph* g)
{
  if (!path || !valid(g)) return EMST_E_ARG;
  int64_t up = 0;
  for (int32_t v = 0; v < g->nodes; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) up += g->nlist[j] > v;
  FILE* f = std::fopen(path, "wt");
  if (!f) return EMST_E_OPEN;
  TextOut o(f);
  o.printf_("%d\n%" PRId64 "\n", g->nodes, up);
  for (int32_t v = 0; v < g->nodes; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++)
      if (g->nlist[j] > v) o.printf_("%d %d %d\n", v, g->nlist[j], g->eweight[j]);
  return o.close();
}

}  // extern "C"
