// egr_io.cpp -- ECLgraph binary CSR reader/writer and the host-side pieces
// of the reference's main(): synthetic weights and the plain threshold.
//
// File layout (ECLgraph.h:29-64, little-endian int32): nodes, edges,
// nindex[nodes+1], nlist[edges], then an optional eweight[edges].  The
// reader keeps the reference's acceptance rules and reports each of its
// failure branches as a distinct error code instead of exiting.

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/emst.h"
#include "host_util.hpp"

namespace emst {
static std::atomic<int> g_threads{0};
int host_threads()
{
  const int t = g_threads.load();
  if (t > 0) return t;
  const unsigned hc = std::thread::hardware_concurrency();
  return hc ? (int)std::min(hc, 64u) : 1;
}
void set_host_threads(int t) { g_threads.store(t < 0 ? 0 : t); }
}  // namespace emst

extern "C" {

const char* emst_version(void) { return "emst-mi355x 1.0"; }

void emst_gen_set_threads(int32_t threads) { emst::set_host_threads(threads); }

int emst_read_egr(const char* path, emst_graph_buf* out)
{
  if (!path || !out) return EMST_E_ARG;
  std::memset(out, 0, sizeof(*out));
  FILE* f = std::fopen(path, "rb");                                   // ECLgraph.h:42
  if (!f) return EMST_E_OPEN;
  int32_t nodes = 0, edges = 0;
  if (std::fread(&nodes, 4, 1, f) != 1) { std::fclose(f); return EMST_E_NODES; }  // :43
  if (std::fread(&edges, 4, 1, f) != 1) { std::fclose(f); return EMST_E_EDGES; }  // :44
  if (nodes < 1 || edges < 0) { std::fclose(f); return EMST_E_COUNTS; }           // :45
  const size_t ni = (size_t)nodes + 1, ne = (size_t)edges;
  int32_t* nindex = (int32_t*)std::malloc(ni * 4);
  int32_t* nlist = (int32_t*)std::malloc((ne ? ne : 1) * 4);
  int32_t* ew = (int32_t*)std::malloc((ne ? ne : 1) * 4);
  if (!nindex || !nlist || !ew) {                                     // :50
    std::free(nindex); std::free(nlist); std::free(ew); std::fclose(f);
    return EMST_E_NOMEM;
  }
  int rc = 0;
  if (std::fread(nindex, 4, ni, f) != ni) rc = EMST_E_NINDEX;         // :52
  else if (std::fread(nlist, 4, ne, f) != ne) rc = EMST_E_NLIST;      // :53
  size_t cnt = 0;
  if (!rc) {
    cnt = std::fread(ew, 4, ne, f);                                   // :54
    if (cnt != 0 && cnt != ne) rc = EMST_E_WEIGHTS;                   // :59
  }
  std::fclose(f);
  if (rc) { std::free(nindex); std::free(nlist); std::free(ew); return rc; }
  if (cnt == 0) { std::free(ew); ew = nullptr; }                      // :55-57
  out->nodes = nodes;
  out->edges = edges;
  out->nindex = nindex;
  out->nlist = nlist;
  out->eweight = ew;
  return EMST_OK;
}

int emst_write_egr(const char* path, const emst_graph* g)
{
  if (!path || !g) return EMST_E_ARG;
  if (g->nodes < 1 || g->edges < 0) return EMST_E_COUNTS;            // ECLgraph.h:68
  if (!g->nindex || (g->edges > 0 && !g->nlist)) return EMST_E_ARG;
  FILE* f = std::fopen(path, "wb");                                   // :70
  if (!f) return EMST_E_OPEN;
  const size_t ni = (size_t)g->nodes + 1, ne = (size_t)g->edges;
  bool ok = std::fwrite(&g->nodes, 4, 1, f) == 1 && std::fwrite(&g->edges, 4, 1, f) == 1 &&
            std::fwrite(g->nindex, 4, ni, f) == ni && (ne == 0 || std::fwrite(g->nlist, 4, ne, f) == ne);
  if (ok && g->eweight && ne) ok = std::fwrite(g->eweight, 4, ne, f) == ne;  // :76-78
  if (std::fclose(f) != 0) ok = false;
  return ok ? EMST_OK : EMST_E_WRITE;
}

void emst_free_graph(emst_graph_buf* g)
{
  if (!g) return;
  std::free(g->nindex);
  std::free(g->nlist);
  std::free(g->eweight);
  g->nindex = g->nlist = g->eweight = nullptr;
}

int emst_synth_weights(emst_graph_buf* g)
{
  if (!g || !g->nindex || g->nodes < 1) return EMST_E_ARG;
  if (g->eweight) return EMST_OK;
  int32_t* w = (int32_t*)std::malloc((size_t)(g->edges ? g->edges : 1) * 4);
  if (!w) return EMST_E_NOMEM;
  const int32_t n = g->nodes;
  const int32_t* nindex = g->nindex;
  const int32_t* nlist = g->nlist;
  emst::parallel_for(n, [&](int64_t b, int64_t e, int) {
    for (int32_t i = (int32_t)b; i < (int32_t)e; i++)
      for (int32_t j = nindex[i]; j < nindex[i + 1]; j++) {
        // ECL_MST.cu:414-416, with the int multiply made explicitly wrapping.
        const int32_t prod = (int32_t)((uint32_t)i * (uint32_t)nlist[j]);
        int32_t v = 1 + prod % n;
        w[j] = v < 0 ? -v : v;
      }
  });
  g->eweight = w;
  return EMST_OK;
}

int emst_validate(const emst_graph* g)
{
  if (!g || !g->nindex) return EMST_E_ARG;
  if (g->nodes < 1 || g->edges < 0) return EMST_E_GRAPH;
  if (g->edges > 0 && (!g->nlist || !g->eweight)) return EMST_E_ARG;
  const int32_t n = g->nodes, m = g->edges;
  if (g->nindex[0] != 0 || g->nindex[n] != m) return EMST_E_GRAPH;
  std::atomic<int> bad{0};
  emst::parallel_for(n, [&](int64_t b, int64_t e, int) {
    for (int64_t v = b; v < e && !bad.load(std::memory_order_relaxed); v++) {
      const int32_t lo = g->nindex[v], hi = g->nindex[v + 1];
      if (hi < lo || lo < 0 || hi > m) { bad = 1; break; }
      for (int32_t j = lo; j < hi; j++)
        if ((uint32_t)g->nlist[j] >= (uint32_t)n) { bad = 1; break; }
    }
  });
  return bad.load() ? EMST_E_GRAPH : EMST_OK;
}

static inline uint32_t ecl_hash(uint32_t v)  // ECL_MST.cu:390-395
{
  v = ((v >> 16) ^ v) * 0x45d9f3bu;
  v = ((v >> 16) ^ v) * 0x45d9f3bu;
  return (v >> 16) ^ v;
}

int emst_plain_threshold(const emst_graph* g, int* use_filter, int32_t* threshold)
{
  if (!g || !use_filter || !threshold || g->nodes < 1 || g->edges < 0) return EMST_E_ARG;
  if (g->edges > 0 && !g->eweight) return EMST_E_NOWEIGHT;
  const int32_t avg_deg = g->edges / g->nodes;                       // ECL_MST.cu:408
  if (avg_deg >= 4) {                                                 // :432
    int32_t s[20];
    const int32_t ns = std::min(g->edges, 20);                        // :436
    for (int32_t i = 0; i < ns; i++) s[i] = g->eweight[ecl_hash((uint32_t)i) % (uint32_t)g->edges];
    std::sort(s, s + ns);                                             // :438
    // :440 -- the index expression is evaluated in double, left to right.
    const int32_t idx = std::min(19, (int32_t)(3.0 * g->nodes * ns / g->edges));
    *use_filter = 1;
    *threshold = s[idx];
  } else {
    *use_filter = 0;                                                  // :443
    *threshold = INT_MAX;
  }
  return EMST_OK;
}

uint64_t emst_mst_weight(const emst_graph* g, const uint8_t* inMST)
{
  if (!g || !inMST || !g->eweight) return 0;
  uint64_t s = 0;
  for (int32_t j = 0; j < g->edges; j++)
    if (inMST[j]) s += (uint64_t)(int64_t)g->eweight[j];             // ECL_MST.cu:377
  return s;
}

}  // extern "C"
