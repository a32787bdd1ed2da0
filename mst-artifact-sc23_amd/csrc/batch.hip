// batch.hip -- many independent graphs solved together (no counterpart in
// the reference, which runs one process per input, run_all_ECL-MST.sh:7).
//
// The graphs are packed into HBM arenas.  Each arena ("chunk") is the
// disjoint union of consecutive graphs with fewer than 2^31 directed edges
// in total: vertex ids and edge positions of graph g are shifted by the
// chunk's running vertex and edge offsets.  The minimum spanning forest of
// a disjoint union is the union of the forests, and inside one graph the
// key order (w, offset + j) equals the order (w, j) of a standalone solve,
// so every graph's inMST slice is bit-identical to solving it alone.
// Per-graph light-edge thresholds enter the union solve through a
// vertex -> graph map.  The launch sequence of a chunk is captured once in
// a hipGraph (rounds are a deterministic function of the input) and then
// replayed; a replay whose last round did not end the phase falls back to
// the dynamic round loop.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/emst.h"
#include "emst_internal.hpp"
#include "host_util.hpp"
#include "learned.hpp"

using namespace emst;

namespace {

struct Chunk {
  int32_t g0 = 0, g1 = 0;          // graphs [g0, g1)
  int32_t nodes = 0, edges = 0;    // union sizes
  long long m_up = 0;              // upper-direction edges
  int* d_nindex = nullptr;
  int* d_nlist = nullptr;
  int* d_eweight = nullptr;
  uint8_t* d_inMST = nullptr;
  int* d_vgid = nullptr;           // chunk-local graph index of each vertex
  int* d_gthr = nullptr;           // per-graph threshold (INT_MAX: unfiltered)
  int* d_gflag = nullptr;          // per-graph filter flag
  long long* d_voff = nullptr;     // per-graph vertex offsets (chunk-local), g1-g0+1 entries
  long long* d_eoff = nullptr;     // per-graph edge offsets
  int* d_rev = nullptr;            // reverse positions (sorted, symmetric lists), else null
  bool rev_mirrored = false;
  std::vector<long long> voff, eoff;
  // captured schedule
  hipGraphExec_t exec = nullptr;
  int cap_mode = -1;
  unsigned long long cap_gen = 0;
  emst_config cap_cfg{};
  FixedSchedule sched;
  bool any_filter = false;
};

}  // namespace

struct emst_batch {
  emst_ctx* ctx = nullptr;
  int32_t ngraphs = 0;
  std::vector<Chunk> chunks;
  std::vector<int32_t> graph_chunk;        // chunk of each graph
  std::vector<int32_t> plain_thr, plain_flag;  // plain rule per graph (host)
  std::vector<int32_t> last_thr, last_flag;
  std::vector<double> last_arena_ms;  // device time of each arena in the last solve
  long long total_nodes = 0, total_edges = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  unsigned short* d_mlp16 = nullptr;  // the batch's own network (bf16), if set
  int have_model = 0;
};

namespace {

void free_chunk(Chunk& c)
{
  for (void* p : {(void*)c.d_nindex, (void*)c.d_nlist, (void*)c.d_eweight, (void*)c.d_inMST, (void*)c.d_vgid,
                  (void*)c.d_gthr, (void*)c.d_gflag, (void*)c.d_voff, (void*)c.d_eoff, (void*)c.d_rev})
    if (p) (void)hipFree(p);
  if (c.exec) (void)hipGraphExecDestroy(c.exec);
  c = Chunk();
}

int plain_rule(const emst_graph& g, int32_t* T, int32_t* flag)
{
  int uf = 0;
  int32_t t = INT_MAX;
  const int rc = emst_plain_threshold(&g, &uf, &t);
  *T = uf ? t : INT_MAX;
  *flag = uf;
  return rc;
}

}  // namespace

extern "C" {

int emst_batch_create(emst_ctx* ctx, int32_t ngraphs, const emst_graph* graphs, int64_t max_arena_edges,
                      emst_batch** out)
{
  if (!ctx || !out || ngraphs < 1 || !graphs || max_arena_edges < 0) return EMST_E_ARG;
  const long long cap = (max_arena_edges == 0 || max_arena_edges >= INT32_MAX) ? (long long)INT32_MAX - 1
                                                                               : max_arena_edges;
  *out = nullptr;
  for (int32_t i = 0; i < ngraphs; i++) {
    const emst_graph& g = graphs[i];
    if (g.edges > 0 && !g.eweight) return ctx_fail(ctx, EMST_E_NOWEIGHT, "a batch graph has no weights");
    const int rc = emst_validate(&g);
    if (rc) return ctx_fail(ctx, rc, "a batch graph failed CSR validation");
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return EMST_E_HIP;
  emst_batch* b = new emst_batch();
  b->ctx = ctx;
  b->ngraphs = ngraphs;
  b->graph_chunk.resize(ngraphs);
  b->plain_thr.resize(ngraphs);
  b->plain_flag.resize(ngraphs);
  b->last_thr.assign(ngraphs, INT_MAX);
  b->last_flag.assign(ngraphs, 0);
  // chunking: consecutive graphs while the union stays below 2^31 edges/vertices
  int32_t g = 0;
  while (g < ngraphs) {
    Chunk c;
    c.g0 = g;
    long long n = 0, m = 0;
    while (g < ngraphs && n + graphs[g].nodes < INT32_MAX && (g == c.g0 || m + graphs[g].edges <= cap)) {
      n += graphs[g].nodes;
      m += graphs[g].edges;
      g++;
    }
    if (g == c.g0) {  // one graph alone too large cannot happen (int32 sizes)
      delete b;
      return EMST_E_LIMIT;
    }
    c.g1 = g;
    c.nodes = (int32_t)n;
    c.edges = (int32_t)m;
    b->chunks.push_back(c);
  }
  int rc = EMST_OK;
  for (size_t ci = 0; ci < b->chunks.size() && !rc; ci++) {
    Chunk& c = b->chunks[ci];
    const int32_t G = c.g1 - c.g0;
    c.voff.resize(G + 1);
    c.eoff.resize(G + 1);
    c.voff[0] = c.eoff[0] = 0;
    for (int32_t k = 0; k < G; k++) {
      c.voff[k + 1] = c.voff[k] + graphs[c.g0 + k].nodes;
      c.eoff[k + 1] = c.eoff[k] + graphs[c.g0 + k].edges;
      b->graph_chunk[c.g0 + k] = (int32_t)ci;
      b->total_nodes += graphs[c.g0 + k].nodes;
      b->total_edges += graphs[c.g0 + k].edges;
    }
    std::vector<int32_t> nindex((size_t)c.nodes + 1), nlist((size_t)std::max(c.edges, 1)),
        ew((size_t)std::max(c.edges, 1)), vgid((size_t)c.nodes);
    std::vector<long long> ups(G, 0), lows(G, 0);
    emst::parallel_for(G, [&](int64_t lo, int64_t hi, int) {
      for (int64_t k = lo; k < hi; k++) {
        const emst_graph& gg = graphs[c.g0 + k];
        const long long vo = c.voff[k], eo = c.eoff[k];
        long long up = 0, low = 0;
        for (int32_t v = 0; v < gg.nodes; v++) {
          nindex[vo + v] = (int32_t)(eo + gg.nindex[v]);
          vgid[vo + v] = (int32_t)k;
          for (int32_t j = gg.nindex[v]; j < gg.nindex[v + 1]; j++) {
            up += gg.nlist[j] > v;
            low += gg.nlist[j] < v;
          }
        }
        lows[k] = low;
        for (int32_t j = 0; j < gg.edges; j++) {
          nlist[eo + j] = (int32_t)(vo + gg.nlist[j]);
          ew[eo + j] = gg.eweight[j];
        }
        ups[k] = up;
        plain_rule(gg, &b->plain_thr[c.g0 + k], &b->plain_flag[c.g0 + k]);
      }
    });
    nindex[c.nodes] = c.edges;
    long long m_low = 0;
    for (long long u : ups) c.m_up += u;
    for (long long l : lows) m_low += l;
    const size_t ne = (size_t)std::max(c.edges, 1);
    bool ok = hipMalloc((void**)&c.d_nindex, ((size_t)c.nodes + 1) * 4) == hipSuccess &&
              hipMalloc((void**)&c.d_nlist, ne * 4) == hipSuccess && hipMalloc((void**)&c.d_eweight, ne * 4) == hipSuccess &&
              hipMalloc((void**)&c.d_inMST, ne + 16) == hipSuccess &&
              hipMalloc((void**)&c.d_vgid, (size_t)c.nodes * 4) == hipSuccess &&
              hipMalloc((void**)&c.d_gthr, (size_t)G * 4) == hipSuccess &&
              hipMalloc((void**)&c.d_gflag, (size_t)G * 4) == hipSuccess &&
              hipMalloc((void**)&c.d_voff, (size_t)(G + 1) * 8) == hipSuccess &&
              hipMalloc((void**)&c.d_eoff, (size_t)(G + 1) * 8) == hipSuccess;
    if (!ok) {
      rc = ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of a batch arena failed");
      break;
    }
    // the arena's kernels run on the context's non-blocking stream: copy on
    // that stream and wait before the host staging vectors go out of scope
    hipStream_t cs = ctx->stream;
    ok = hipMemcpyAsync(c.d_nindex, nindex.data(), ((size_t)c.nodes + 1) * 4, hipMemcpyHostToDevice, cs) == hipSuccess &&
         hipMemcpyAsync(c.d_nlist, nlist.data(), (size_t)c.edges * 4, hipMemcpyHostToDevice, cs) == hipSuccess &&
         hipMemcpyAsync(c.d_eweight, ew.data(), (size_t)c.edges * 4, hipMemcpyHostToDevice, cs) == hipSuccess &&
         hipMemcpyAsync(c.d_vgid, vgid.data(), (size_t)c.nodes * 4, hipMemcpyHostToDevice, cs) == hipSuccess &&
         hipMemcpyAsync(c.d_voff, c.voff.data(), (size_t)(G + 1) * 8, hipMemcpyHostToDevice, cs) == hipSuccess &&
         hipMemcpyAsync(c.d_eoff, c.eoff.data(), (size_t)(G + 1) * 8, hipMemcpyHostToDevice, cs) == hipSuccess &&
         hipStreamSynchronize(cs) == hipSuccess;
    if (!ok) {
      rc = ctx_fail(ctx, EMST_E_HIP, "batch arena upload failed");
      break;
    }
    // reverse positions (sorted, symmetric lists with one weight per edge):
    // the light initialisation then also makes the first offers
    if (c.edges > 0 && c.m_up == m_low) {
      if (hipMalloc((void**)&c.d_rev, (size_t)c.edges * 4) != hipSuccess) {
        rc = ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of an arena's reverse positions failed");
        break;
      }
      bool rok = false, mir = false;
      rc = build_rev(ctx, c.nodes, c.edges, c.d_nindex, c.d_nlist, c.d_eweight, c.d_rev, cs, &rok, &mir);
      if (rc) break;
      if (!rok || !mir) {
        (void)hipFree(c.d_rev);
        c.d_rev = nullptr;
      }
      c.rev_mirrored = mir;
    }
  }
  if (!rc) {
    // size the shared work buffers for the largest arena up front, so that a
    // later arena never reallocates buffers an earlier capture points at
    int32_t mn = 1, me = 0;
    long long mu = 1;
    for (const Chunk& c : b->chunks) {
      mn = std::max(mn, c.nodes);
      me = std::max(me, c.edges);
      mu = std::max(mu, c.m_up);
    }
    rc = ensure_buffers(ctx, mn, me, mu);
  }
  if (!rc && (hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess))
    rc = ctx_fail(ctx, EMST_E_HIP, "hipEventCreate failed");
  if (rc) {
    emst_batch_destroy(b);
    return rc;
  }
  *out = b;
  return EMST_OK;
}

void emst_batch_destroy(emst_batch* b)
{
  if (!b) return;
  if (b->ctx) (void)hipSetDevice(b->ctx->device);
  for (Chunk& c : b->chunks) free_chunk(c);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->d_mlp16) (void)hipFree(b->d_mlp16);
  delete b;
}

int emst_batch_set_filter_model(emst_batch* b, const float* params)
{
  if (!b) return EMST_E_ARG;
  emst_ctx* ctx = b->ctx;
  if (!params) {
    b->have_model = 0;
    return EMST_OK;
  }
  std::vector<uint16_t> h(kParams);
  for (int i = 0; i < kParams; i++) {
    if (!std::isfinite(params[i])) return ctx_fail(ctx, EMST_E_ARG, "non-finite model parameter");
    h[i] = emst_bf16_bits(params[i]);
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx_fail(ctx, EMST_E_HIP, "hipSetDevice failed");
  if (!b->d_mlp16 && hipMalloc((void**)&b->d_mlp16, kParams * 2) != hipSuccess)
    return ctx_fail(ctx, EMST_E_NOMEM, "hipMalloc of the batch network failed");
  if (hipMemcpyAsync(b->d_mlp16, h.data(), kParams * 2, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return ctx_fail(ctx, EMST_E_HIP, "upload of the batch network failed");
  b->have_model = 1;
  return EMST_OK;
}

int emst_batch_info(const emst_batch* b, int32_t* chunks, int64_t* total_nodes, int64_t* total_edges)
{
  if (!b) return EMST_E_ARG;
  if (chunks) *chunks = (int32_t)b->chunks.size();
  if (total_nodes) *total_nodes = b->total_nodes;
  if (total_edges) *total_edges = b->total_edges;
  return EMST_OK;
}

int emst_batch_solve(emst_batch* b, int use_filter, const int32_t* thresholds, int use_graph, emst_batch_stats* st)
{
  if (!b || use_filter < EMST_FILTER_OFF || use_filter > EMST_FILTER_LEARNED) return EMST_E_ARG;
  emst_ctx* ctx = b->ctx;
  if (hipSetDevice(ctx->device) != hipSuccess) return EMST_E_HIP;
  if (ctx->cfg.flags & EMST_F_HOST_SYNC) return ctx_fail(ctx, EMST_E_ARG, "batches use the device-driven loop");
  hipStream_t s = ctx->stream;
  if (st) std::memset(st, 0, sizeof(*st));
  double total_ms = 0;
  b->last_arena_ms.assign(b->chunks.size(), 0.0);
  for (size_t ci = 0; ci < b->chunks.size(); ci++) {
    Chunk& c = b->chunks[ci];
    const int32_t G = c.g1 - c.g0;
    // per-graph thresholds
    std::vector<int32_t> thr(G), flag(G);
    if (use_filter == EMST_FILTER_LEARNED) {
      if (b->have_model)
        launch_learned_batch(s, G, c.d_nindex, c.d_eweight, c.d_voff, c.d_eoff, b->d_mlp16, 1, c.d_gthr, c.d_gflag);
      else
        launch_learned_batch(s, G, c.d_nindex, c.d_eweight, c.d_voff, c.d_eoff, ctx->d_mlp16, ctx->have_model,
                             c.d_gthr, c.d_gflag);
      if (hipMemcpyAsync(thr.data(), c.d_gthr, (size_t)G * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipMemcpyAsync(flag.data(), c.d_gflag, (size_t)G * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        return ctx_fail(ctx, EMST_E_HIP, "learned batch thresholds failed");
    } else {
      for (int32_t k = 0; k < G; k++) {
        if (use_filter == EMST_FILTER_OFF) {
          thr[k] = INT_MAX;
          flag[k] = 0;
        } else if (thresholds) {
          thr[k] = thresholds[c.g0 + k];
          flag[k] = thr[k] != INT_MAX;
        } else {
          thr[k] = b->plain_thr[c.g0 + k];
          flag[k] = b->plain_flag[c.g0 + k];
        }
      }
      if (hipMemcpyAsync(c.d_gthr, thr.data(), (size_t)G * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        return ctx_fail(ctx, EMST_E_HIP, "batch threshold upload failed");
    }
    bool any = false;
    for (int32_t k = 0; k < G; k++) {
      any = any || flag[k];
      b->last_thr[c.g0 + k] = thr[k];
      b->last_flag[c.g0 + k] = flag[k];
    }
    VertexThresholds vt;
    vt.vgid = c.d_vgid;
    vt.gthr = c.d_gthr;
    vt.any_filter = any;
    emst_stats cs{};
    // a captured schedule is valid for the same arena, mode and config
    const int mode_key = use_filter * 2 + (any ? 1 : 0);
    if (c.exec && (c.cap_gen != ctx->buf_gen || std::memcmp(&c.cap_cfg, &ctx->cfg, sizeof(emst_config)) != 0)) {
      (void)hipGraphExecDestroy(c.exec);  // buffers or shape changed since the capture
      c.exec = nullptr;
    }
    if (use_graph && c.exec && c.cap_mode == mode_key) {
      for (int p = 0; p < 2; p++)
        if (c.sched.flag_slot[p] >= 0) ctx->h_flag[c.sched.flag_slot[p]] = -1;
      if (hipEventRecord(b->e0, s) != hipSuccess || hipGraphLaunch(c.exec, s) != hipSuccess ||
          hipEventRecord(b->e1, s) != hipSuccess || hipEventSynchronize(b->e1) != hipSuccess)
        return ctx_fail(ctx, EMST_E_HIP, "hipGraph replay failed");
      bool done = true;
      for (int p = 0; p < (any ? 2 : 1); p++)
        if (c.sched.flag_slot[p] < 0 || ctx->h_flag[c.sched.flag_slot[p]] != 0) done = false;
      float ms = 0;
      (void)hipEventElapsedTime(&ms, b->e0, b->e1);
      if (done) {
        long long hs[ST_COUNT];
        if (hipMemcpy(hs, ctx->d_stats, sizeof(hs), hipMemcpyDeviceToHost) != hipSuccess)
          return ctx_fail(ctx, EMST_E_HIP, "stats read failed");
        cs.ms = ms;
        cs.light = hs[ST_LIGHT];
        cs.heavy = hs[ST_HEAVY];
        cs.light_rounds = (int32_t)hs[ST_ROUNDS0];
        cs.heavy_rounds = (int32_t)hs[ST_ROUNDS1];
        cs.tail_rounds = (int32_t)hs[ST_TAIL];
        total_ms += ms;
        b->last_arena_ms[ci] = ms;
        if (st) {
          st->light += cs.light;
          st->heavy += cs.heavy;
          st->light_rounds = std::max(st->light_rounds, cs.light_rounds);
          st->heavy_rounds = std::max(st->heavy_rounds, cs.heavy_rounds);
          st->replays++;
        }
        continue;
      }
      // stale capture: fall through to a dynamic solve and recapture
      (void)hipGraphExecDestroy(c.exec);
      c.exec = nullptr;
    }
    RevIndex rix;
    if (c.d_rev) {
      rix.rev = c.d_rev;
      rix.mirrored = c.rev_mirrored;
    }
    int rc = run_solve(ctx, c.nodes, c.edges, c.d_nindex, c.d_nlist, c.d_eweight,
                       any ? EMST_FILTER_PLAIN : EMST_FILTER_OFF, INT_MAX, c.d_inMST, s, c.m_up, &vt, &cs, nullptr,
                       nullptr, &rix);
    if (rc) return rc;
    total_ms += cs.ms;
    b->last_arena_ms[ci] = cs.ms;
    if (st) {
      st->light += cs.light;
      st->heavy += cs.heavy;
      st->light_rounds = std::max(st->light_rounds, cs.light_rounds);
      st->heavy_rounds = std::max(st->heavy_rounds, cs.heavy_rounds);
    }
    if (use_graph) {
      // capture the same schedule for later replays
      FixedSchedule fs;
      fs.rounds[0] = std::max(1, cs.light_rounds + 1);
      fs.rounds[1] = std::max(1, cs.heavy_rounds + 1);
      hipGraph_t graph = nullptr;
      if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return ctx_fail(ctx, EMST_E_HIP, "hipStreamBeginCapture failed");
      rc = run_solve(ctx, c.nodes, c.edges, c.d_nindex, c.d_nlist, c.d_eweight, any ? EMST_FILTER_PLAIN : EMST_FILTER_OFF,
                     INT_MAX, c.d_inMST, s, c.m_up, &vt, nullptr, &fs, nullptr, &rix);
      const hipError_t ec = hipStreamEndCapture(s, &graph);
      if (rc || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc ? rc : ctx_fail(ctx, EMST_E_HIP, "hipStreamEndCapture failed");
      }
      hipGraphExec_t exec = nullptr;
      const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) return ctx_fail(ctx, EMST_E_HIP, "hipGraphInstantiate failed");
      if (c.exec) (void)hipGraphExecDestroy(c.exec);  // a capture of another mode
      c.exec = exec;
      c.cap_mode = mode_key;
      c.cap_gen = ctx->buf_gen;
      c.cap_cfg = ctx->cfg;
      c.sched = fs;
    }
  }
  if (st) {
    st->ms = total_ms;
    st->chunks = (int32_t)b->chunks.size();
    st->graphs = b->ngraphs;
    st->nodes = b->total_nodes;
    st->edges = b->total_edges;
    int32_t nf = 0;
    for (int32_t f : b->last_flag) nf += f;
    st->filtered_graphs = nf;
  }
  return EMST_OK;
}

int emst_batch_fetch(emst_batch* b, int32_t g, uint8_t* inMST)
{
  if (!b || g < 0 || g >= b->ngraphs) return EMST_E_ARG;
  const Chunk& c = b->chunks[b->graph_chunk[g]];
  const int32_t k = g - c.g0;
  const size_t m = (size_t)(c.eoff[k + 1] - c.eoff[k]);
  if (m == 0) return EMST_OK;
  if (!inMST) return EMST_E_ARG;
  if (hipSetDevice(b->ctx->device) != hipSuccess ||
      hipMemcpy(inMST, c.d_inMST + c.eoff[k], m, hipMemcpyDeviceToHost) != hipSuccess)
    return ctx_fail(b->ctx, EMST_E_HIP, "batch fetch failed");
  return EMST_OK;
}

int emst_batch_thresholds(const emst_batch* b, int32_t* thresholds, int32_t* use_filter)
{
  if (!b) return EMST_E_ARG;
  for (int32_t g = 0; g < b->ngraphs; g++) {
    if (thresholds) thresholds[g] = b->last_thr[g];
    if (use_filter) use_filter[g] = b->last_flag[g];
  }
  return EMST_OK;
}

int emst_batch_arena_ms(const emst_batch* b, double* ms, int32_t* first_graph, int32_t cap, int32_t* arenas)
{
  if (!b || (cap > 0 && (!ms || !first_graph))) return EMST_E_ARG;
  const int32_t n = (int32_t)b->chunks.size();
  if (arenas) *arenas = n;
  for (int32_t i = 0; i < n && i < cap; i++) {
    ms[i] = i < (int32_t)b->last_arena_ms.size() ? b->last_arena_ms[i] : 0.0;
    first_graph[i] = b->chunks[i].g0;
  }
  return EMST_OK;
}

}  // extern "C"
