// noalloc.hip -- the allocation-free solve: emst_solve_device's work
// buffers carved from one static device array per GPU instead of from
// hipMalloc, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  It has no
// counterpart in the reference, whose gpuMST allocates and frees every
// buffer per call (ECL-MST/ECL_MST.cu:268-297, 349-357).
//
// The array lives in the code object; each GPU that loads the library holds
// its own copy until the process ends.  Host code never passes the array's
// name to a kernel or a copy: the name is only a host-side handle.  The
// device address comes from hipGetSymbolAddress, taken once per device at
// context creation with that device current, and is used everywhere the
// host needs the array's memory.  Every call on one GPU shares that GPU's
// array and leaves its contents behind; a solve clears what must start at
// zero itself (counters, parents, keys), and calls that overlap on one GPU
// race.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/emst.h"
#include "emst_internal.hpp"

static __device__ __align__(256) unsigned char scratch[EMST_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices] = {};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

namespace emst {

// Set-up: the device address of this device's scratch array (the device
// must be current).  Idempotent per device.
int noalloc_setup(int device)
{
  if (device < 0 || device >= kMaxDevices) return EMST_E_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch_ptr[device]) return EMST_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return EMST_E_HIP;
  g_scratch_ptr[device] = p;
  return EMST_OK;
}

}  // namespace emst

extern "C" int emst_solve_device_noalloc(emst_ctx* ctx, int32_t nodes, int32_t edges, const int32_t* d_nindex,
                                         const int32_t* d_nlist, const int32_t* d_eweight, int use_filter,
                                         int32_t threshold, uint8_t* d_inMST, void* stream, emst_stats* stats)
{
  using namespace emst;
  if (!ctx || !d_nindex || (edges > 0 && (!d_nlist || !d_eweight || !d_inMST))) return EMST_E_ARG;
  if (nodes < 1 || edges < 0) return ctx_fail(ctx, EMST_E_ARG, "node or edge count too low");
  if (ctx->cfg.flags & kLadderFlags)
    return ctx_fail(ctx, EMST_E_ARG, "the ablation ladder allocates its own lists; not available without allocation");
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
  }
  if (!base) return ctx_fail(ctx, EMST_E_STATE, "scratch address was not set up for this device");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx_fail(ctx, EMST_E_HIP, "hipSetDevice failed");
  // The worklists hold up to one record per directed edge: the upper
  // direction of a device-resident graph is not counted on the host.
  const WorkSizes z = work_sizes(ctx->cus, ctx->cfg, nodes, edges, edges);
  const size_t sz[] = {(size_t)z.nodes * 4,         (size_t)z.nodes * 8,          (size_t)z.vclass_words * 4,
                       (size_t)z.gsrc * 4,          (size_t)z.wl[0] * 16,         (size_t)z.wl[1] * 16,
                       (size_t)z.cand_words * 2 * 8, (size_t)kCntWords * 4,               (size_t)ST_COUNT * 8,
                       (size_t)z.segs * 4};
  size_t off[sizeof(sz) / sizeof(sz[0])];
  size_t total = 0;
  for (size_t i = 0; i < sizeof(sz) / sizeof(sz[0]); i++) {
    off[i] = total;
    total += align256(sz[i]);
  }
  if (total > (size_t)EMST_SCRATCH_BYTES) {
    char buf[192];
    std::snprintf(buf, sizeof(buf),
                  "graph too large for the allocation-free solve: %d nodes and %d edges need %zu bytes of scratch, "
                  "the static array holds %zu",
                  nodes, edges, total, (size_t)EMST_SCRATCH_BYTES);
    return ctx_fail(ctx, EMST_E_LIMIT, buf);
  }
  unsigned char* b = static_cast<unsigned char*>(base);
  WorkBufs wb;
  wb.parent = reinterpret_cast<int*>(b + off[0]);
  wb.minv = reinterpret_cast<u64*>(b + off[1]);
  wb.vclass = reinterpret_cast<unsigned*>(b + off[2]);
  wb.gsrc = reinterpret_cast<int*>(b + off[3]);
  wb.wl[0] = reinterpret_cast<int4*>(b + off[4]);
  wb.wl[1] = reinterpret_cast<int4*>(b + off[5]);
  wb.cand = reinterpret_cast<u64*>(b + off[6]);
  wb.cnt = reinterpret_cast<int*>(b + off[7]);
  wb.stats = reinterpret_cast<long long*>(b + off[8]);
  wb.segcnt = reinterpret_cast<int*>(b + off[9]);
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  return run_solve(ctx, nodes, edges, d_nindex, d_nlist, d_eweight, use_filter, threshold, d_inMST, s, edges,
                   nullptr, stats, nullptr, &wb);
}
