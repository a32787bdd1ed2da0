// emst_internal.hpp -- context layout and the solve entry shared by the
// single-graph C-ABI (emst_solver.hip) and the batched arena (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/emst.h"
#include "emst_kernels.hpp"

struct Relabel {
  bool ok = false;
  int* nindex = nullptr;
  int* nlist = nullptr;
  int* eweight = nullptr;
  int* kid = nullptr;
  int* order = nullptr;
  int* newid = nullptr;
  size_t cap_nodes = 0, cap_edges = 0;
  double ms = 0;   // build time at the last load
  int mode = 0;    // the order in use: emst::kOrderDegree or emst::kOrderCluster
  int levels = 0;  // breadth-first levels of the cluster search (cluster order)
};

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
  emst::u64* d_minv = nullptr;
  unsigned* d_vclass = nullptr;
  size_t cap_nodes = 0;
  int4* d_wl[2] = {nullptr, nullptr};
  size_t cap_wl[2] = {0, 0};
  emst::u64* d_cand = nullptr;        // first-round candidate bitmasks, joined-record masks
  size_t cap_cand = 0;
  int* d_gsrc = nullptr;
  size_t cap_gsrc = 0;
  int* d_segcnt = nullptr;            // records per initialisation segment
  size_t cap_seg = 0;
  unsigned char* d_skip = nullptr;    // per-task skip flags of the sparse heavy initialisation
  size_t cap_skip = 0;
  unsigned char* d_hasmem = nullptr;  // per-vertex flag: a root with other members (sparse heavy initialisation)
  size_t cap_hasmem = 0;
  // reverse positions of the loaded graph (emst_load), when its lists are
  // sorted and symmetric
  int* d_rev = nullptr;
  size_t cap_rev = 0;
  bool rev_ok = false;
  bool rev_mirrored = false;
  // renumbered layout of the loaded graph (emst_load: degree order for
  // skewed degrees, cluster order for scattered ids on a deep graph): lists
  // sorted by the new ids, each position's key id (its edge's
  // upper-direction position in the caller's CSR) and that copy's weight;
  // order maps new ids to the caller's, newid the reverse
  Relabel rl;
  int* d_cnt = nullptr;
  long long* d_stats = nullptr;
  int* h_flag = nullptr;
  int* d_flag = nullptr;
  long long* h_report = nullptr;      // host-mapped solve report (k_report)
  long long* d_report = nullptr;
  int* h_decide = nullptr;            // host-mapped learned filter decision
  int* d_decide = nullptr;
  // de-optimization ladder (ladder.hip): worklists, per-edge sources, counters
  void* d_lad[2] = {nullptr, nullptr};
  size_t cap_lad[2] = {0, 0};
  int* d_slist = nullptr;
  size_t cap_slist = 0;
  int* d_lad_cnt = nullptr;
  size_t cap_lad_cnt = 0;
  hipEvent_t ev[8] = {};
  int32_t last_nodes = 0;
  const int* last_parent = nullptr;   // parent array of the last solve (context or scratch)
  bool last_relabelled = false;       // ... in the renumbered layout's numbering
  // learned threshold
  unsigned short* d_mlp16 = nullptr;  // bf16 parameters
  int have_model = 0;
  int* d_thr = nullptr;               // [threshold, use_filter] decided on the device
  float* d_q = nullptr;
  float* d_feat = nullptr;
  int* d_samples = nullptr;
  // the learned threshold of the loaded graph, decided once outside the
  // timed region (as the reference decides its threshold before gpuMST,
  // ECL_MST.cu:432-441); d_thr and d_q hold it while learned_cached is set
  bool learned_cached = false;
  int learned_filter = 0;
  int learned_T = INT_MAX;  // the cached decision's threshold (a solve then runs as a plain one at it)
  // incremented whenever a work buffer is (re)allocated: captured launch
  // schedules hold raw buffer addresses and are valid for one generation
  unsigned long long buf_gen = 0;
};

namespace emst {

// Device counter words of one solve (WorkBufs::cnt).
constexpr int kCntWords = 8;

// Device work buffers of one solve: the context's own (grown on demand) or
// carved from the library's static scratch array.
struct WorkBufs {
  int* parent = nullptr;        // nodes
  u64* minv = nullptr;          // nodes
  unsigned* vclass = nullptr;   // two bits per vertex
  int* gsrc = nullptr;          // start vertex of every initialisation task
  int4* wl[2] = {nullptr, nullptr};
  u64* cand = nullptr;          // two bitmasks of cand_words words
  int* cnt = nullptr;           // kCntWords: [list 0 size, list 1 size, marked component, sparse, 1, -, -, -]
  long long* stats = nullptr;   // ST_COUNT counters
  int* segcnt = nullptr;        // records per initialisation segment
  unsigned char* skip = nullptr;  // per-task skip flags (sparse heavy initialisation)
  unsigned char* hasmem = nullptr;  // nodes: root-with-members flags (sparse heavy initialisation)
};

// Reverse positions of a graph whose lists are sorted and symmetric: rev[p]
// is the position of the upper-direction copy of the edge at lower-direction
// position p.  They enable the sparse heavy initialisation.
struct RevIndex {
  const int* rev = nullptr;
  bool mirrored = false;  // every edge's two copies carry the same weight
  // rev holds the key id of every position (a relabelled layout): the
  // fused passes key records and offers by rev[p] in both directions; the
  // streaming passes write positions, which the phase's first select
  // replaces by their key ids
  bool kid = false;
};

// Element counts of the work buffers for a solve of the given size under a
// configuration.  The initialisation kernels run init_grid workgroups, each
// appending to its own segment of seg_cap records at the start of list 0.
struct WorkSizes {
  long long nodes = 0, vclass_words = 0, gsrc = 0, wl[2] = {0, 0}, cand_words = 0, segs = 0;
  int init_grid = 0;
  long long seg_cap = 0;
};

WorkSizes work_sizes(int cus, const emst_config& cfg, int32_t nodes, int32_t edges, long long wl_cap);

// Per-vertex thresholds for a union of graphs solved together: the
// threshold of vertex v's graph is gthr[vgid[v]].  Both NULL for one graph.
struct VertexThresholds {
  const int* vgid = nullptr;
  const int* gthr = nullptr;
  bool any_filter = true;  // run the heavy phase
};

// Fixed launch schedule for stream capture: exactly rounds[p] join/select
// iterations per phase, no host polling, no events or synchronisation.
// Round counts are a deterministic function of the input, so the counts of
// an earlier solve of the same input replay it exactly; the caller checks
// the posted flags (flag_slot[p] = slot of the phase's last join) after a
// replay.
struct FixedSchedule {
  int rounds[2] = {0, 0};
  int flag_slot[2] = {-1, -1};
};

int ctx_fail(emst_ctx* c, int code, const std::string& what);

// Grows the context's work buffers to hold a solve of this size under the
// context's configuration (worklists of at least wl_cap records).
int ensure_buffers(emst_ctx* ctx, int32_t nodes, int32_t edges, long long wl_cap);

// One solve.  With wb == nullptr the context's buffers are grown and used;
// otherwise wb must hold work_sizes(...) of this solve and nothing is
// allocated.
int run_solve(emst_ctx* ctx, int32_t nodes, int32_t edges, const int* d_nindex, const int* d_nlist,
              const int* d_eweight, int use_filter, int32_t threshold, uint8_t* d_inMST, hipStream_t s,
              long long wl_cap, const VertexThresholds* vt, emst_stats* stats, FixedSchedule* fixed = nullptr,
              const WorkBufs* wb = nullptr, const RevIndex* rix = nullptr, bool learned_cached = false);

// Builds the reverse positions of a device-resident graph into rev (edges
// ints) on stream s; *ok is set when the lists are sorted and every
// lower-direction copy has a partner (the caller also checks that the
// graph has as many lower- as upper-direction positions), *mirrored when
// every pair carries one weight.
int build_rev(emst_ctx* ctx, int32_t nodes, int32_t edges, const int* d_nindex, const int* d_nlist,
              const int* d_eweight, int* rev, hipStream_t s, bool* ok, bool* mirrored);

// Builds the context's renumbered layout (ctx->rl) of the loaded graph from
// its device copy and reverse positions (relabel.hip); sets rl.ok.
// kOrderDegree: by decreasing degree.  kOrderCluster: by breadth-first
// clusters grown from every 4096th vertex, kept only when the search takes
// at least min_levels levels (a graph of high diameter).
constexpr int kOrderDegree = 1, kOrderCluster = 2;
int build_relabel(emst_ctx* ctx, int32_t nodes, int32_t edges, const int* d_nindex, const int* d_nlist,
                  const int* d_eweight, const int* rev, hipStream_t s, int mode, int min_levels);
void free_relabel(Relabel& rl);
// Copies a parent array of the renumbered layout to host memory in the
// caller's numbering.
int relabel_parent(emst_ctx* ctx, const int* parent, int32_t nodes, int32_t* out, hipStream_t s);

// The reference's kernel sequence with switchable optimizations (the
// EMST_F_LADDER and EMST_F_NO_* bits of the context's flags): one graph, a
// given threshold or none.
int run_ladder(emst_ctx* ctx, int32_t nodes, int32_t edges, const int* d_nindex, const int* d_nlist,
               const int* d_eweight, int use_filter, int32_t threshold, uint8_t* d_inMST, hipStream_t s,
               long long wl_cap, emst_stats* stats);

// Takes the device address of this device's static scratch array for the
// allocation-free solve (noalloc.hip); called once per device, with the
// device current, when a context is created.
int noalloc_setup(int device);

// Flags that select the ladder schedule.
constexpr int kLadderFlags = EMST_F_LADDER | EMST_F_NO_WARP | EMST_F_NO_IMPLICIT_PC | EMST_F_NO_ONE_DIRECTION |
                             EMST_F_NO_TUPLES | EMST_F_NO_DATA_DRIVEN | EMST_F_NO_EDGE_BASED;
constexpr int kAllFlags = EMST_F_HOST_SYNC | EMST_F_NO_GUARD | EMST_F_NO_COMBINE | EMST_F_NO_PAIR_DEDUP |
                          EMST_F_NO_SPARSE_HEAVY | EMST_F_NO_LIGHT_OFFERS | EMST_F_NO_RELABEL | EMST_F_RELABEL |
                          EMST_F_CLUSTER |
                          kLadderFlags;

}  // namespace emst
