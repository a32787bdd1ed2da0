// emst_internal.hpp -- context layout and the solve entry shared by the
// single-graph C-ABI (emst_solver.hip) and the batched arena (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/emst.h"
#include "emst_kernels.hpp"

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
  emst::u64* d_cand = nullptr;        // first-round candidate bitmasks
  size_t cap_wl = 0;
  int* d_gsrc = nullptr;
  size_t cap_gsrc = 0;
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
  // learned threshold
  unsigned short* d_mlp16 = nullptr;  // bf16 parameters
  int have_model = 0;
  int* d_thr = nullptr;               // [threshold, use_filter] decided on the device
  float* d_q = nullptr;
  float* d_feat = nullptr;
  int* d_samples = nullptr;
  // incremented whenever a work buffer is (re)allocated: captured launch
  // schedules hold raw buffer addresses and are valid for one generation
  unsigned long long buf_gen = 0;
};

namespace emst {

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

// Grows the context's work buffers to hold a solve of this size.
int ensure_buffers(emst_ctx* ctx, int32_t nodes, int32_t edges, long long wl_cap, int task_edges);

int run_solve(emst_ctx* ctx, int32_t nodes, int32_t edges, const int* d_nindex, const int* d_nlist,
              const int* d_eweight, int use_filter, int32_t threshold, uint8_t* d_inMST, hipStream_t s,
              long long wl_cap, const VertexThresholds* vt, emst_stats* stats, FixedSchedule* fixed = nullptr);

// The reference's kernel sequence with switchable optimizations (the
// EMST_F_LADDER and EMST_F_NO_* bits of the context's flags): one graph, a
// given threshold or none.
int run_ladder(emst_ctx* ctx, int32_t nodes, int32_t edges, const int* d_nindex, const int* d_nlist,
               const int* d_eweight, int use_filter, int32_t threshold, uint8_t* d_inMST, hipStream_t s,
               long long wl_cap, emst_stats* stats);

// Flags that select the ladder schedule.
constexpr int kLadderFlags = EMST_F_LADDER | EMST_F_NO_WARP | EMST_F_NO_IMPLICIT_PC | EMST_F_NO_ONE_DIRECTION |
                             EMST_F_NO_TUPLES | EMST_F_NO_DATA_DRIVEN | EMST_F_NO_EDGE_BASED;
constexpr int kAllFlags = EMST_F_HOST_SYNC | EMST_F_NO_GUARD | EMST_F_NO_COMBINE | EMST_F_NO_PAIR_DEDUP | kLadderFlags;

}  // namespace emst
