/*
 * emst.h -- C-ABI of the MI355X ECL-MST solver (libemst.so).
 *
 * The reference (alexfallin/MST-Artifact-SC23, ECL-MST/ECL_MST.cu) has no
 * FFI: its boundary is the in-file function
 *     template<bool filter> static bool* gpuMST(const ECLgraph& g, int threshold)
 * (ECL_MST.cu:265-362) called from main (ECL_MST.cu:432-444) after the host
 * picks the light-edge threshold (ECL_MST.cu:432-440).  The entry points
 * below replace that function and its helpers one for one; each cites the
 * reference code it replaces.  Plain pointers and sizes only.
 *
 * Conventions
 *  - Every entry point returns EMST_OK (0) or a negative EMST_E_* code and
 *    never exits the process (the reference calls exit(-1) in CheckCuda,
 *    ECL_MST.cu:254-262, and in readECLgraph, ECLgraph.h:42-59).
 *  - inMST is one byte per directed CSR edge, 0 or 1, byte-compatible with
 *    the reference's bool[edges] (ECL_MST.cu:269-270).  Only the
 *    upper-direction position j (nlist[j] > source) of an MSF edge is 1.
 *  - A context is bound to one device and one HIP stream and is not
 *    reentrant; use one context per host thread.
 */
#ifndef EMST_H
#define EMST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMST_OK 0
#define EMST_E_ARG (-1)      /* NULL pointer, bad size, bad option            */
#define EMST_E_GRAPH (-2)    /* CSR fails validation (see emst_validate)      */
#define EMST_E_HIP (-3)      /* a HIP runtime call failed                      */
#define EMST_E_NOMEM (-4)    /* device or host allocation failed               */
#define EMST_E_LIMIT (-5)    /* size beyond a documented limit                 */
#define EMST_E_STATE (-6)    /* call order (e.g. emst_solve before emst_load)  */
#define EMST_E_NOWEIGHT (-7) /* graph has no weights and none were synthesized */
/* .egr reader errors, one per failure branch of ECLgraph.h:42-59 */
#define EMST_E_OPEN (-10)    /* ECLgraph.h:42 could not open file              */
#define EMST_E_NODES (-11)   /* ECLgraph.h:43 failed to read nodes             */
#define EMST_E_EDGES (-12)   /* ECLgraph.h:44 failed to read edges             */
#define EMST_E_COUNTS (-13)  /* ECLgraph.h:45 node or edge count too low       */
#define EMST_E_NINDEX (-14)  /* ECLgraph.h:52 failed to read neighbor index    */
#define EMST_E_NLIST (-15)   /* ECLgraph.h:53 failed to read neighbor list     */
#define EMST_E_WEIGHTS (-16) /* ECLgraph.h:59 failed to read edge weights      */
#define EMST_E_WRITE (-17)   /* ECLgraph.h:70-78 write failures                */

/* Read-only view of a graph in the ECLgraph 32-bit CSR layout
 * (ECLgraph.h:29-35): nindex[nodes+1], nlist[edges], eweight[edges]. */
typedef struct {
  int32_t nodes;
  int32_t edges;
  const int32_t* nindex;
  const int32_t* nlist;
  const int32_t* eweight;
} emst_graph;

/* Owning graph returned by the reader and the generators. */
typedef struct {
  int32_t nodes;
  int32_t edges;
  int32_t* nindex;
  int32_t* nlist;
  int32_t* eweight; /* NULL when the file had no weight section */
} emst_graph_buf;

/* Per-solve statistics.  Counts are exact; times come from HIP events on
 * the context's stream. */
typedef struct {
  double ms;                 /* device time, first kernel to last (ECL_MST.cu:301-345) */
  double init_ms;            /* time inside the worklist-initialisation kernels        */
  double init_light_ms;      /* of which the light (first) initialisation launch       */
  double init_heavy_ms;      /* of which the heavy (second) initialisation launch      */
  int32_t init_launches;     /* number of worklist-initialisation launches timed      */
  int32_t light_rounds;      /* kernel1-equivalent launches, light phase               */
  int32_t heavy_rounds;      /* same, heavy phase (0 without the filter)               */
  int32_t tail_rounds;       /* rounds finished inside the one-workgroup tail kernel   */
  int64_t light;             /* L: records emitted by the light initialisation         */
  int64_t heavy;             /* H: heavy records whose components differ               */
  int64_t mst_edges;         /* edges marked in inMST                                  */
  uint64_t weight;           /* 64-bit wrapping sum of marked weights (ECL_MST.cu:377) */
  int64_t round_in_sum;      /* sum over rounds of worklist sizes entering a round     */
  int64_t round_out_sum;     /* sum over rounds of surviving records                   */
  int32_t threshold_used;    /* light-edge threshold the solve used (INT32_MAX = none)  */
  int32_t filter_used;       /* 1 when the two-phase light/heavy path ran              */
  float learned_q;           /* quantile predicted by the network, -1 if not used      */
  int32_t sparse_heavy;      /* 1 when the sparse heavy initialisation ran             */
  int32_t light_offers;      /* 1 when the light initialisation made the first offers  */
  int32_t relabelled;        /* renumbered copy the solve ran on: 0 none, 1 by degree,   */
                             /* 2 by breadth-first clusters                              */
} emst_stats;

/* Tunables.  Zero in a field selects the built-in default. */
typedef struct {
  int32_t block;            /* threads per workgroup of the round kernels: 256, 512 or 1024 */
  int32_t grid_per_cu;      /* workgroups per CU of the grid-stride round kernels            */
  int32_t tile_edges;       /* CSR edges per wave task of the initialisation kernels:        */
                            /* 256, 512 or 1024 (the edge-centric replacement of the         */
                            /* reference's thread/warp degree cutoff, ECL_MST.cu:167, 185)   */
  int32_t stage_recs;       /* LDS compaction staging records per wave before a flush        */
  int32_t tail_max;         /* worklist size at or below which one workgroup finishes the    */
                            /* phase (-1 disables the tail)                                  */
  int32_t flags;            /* EMST_F_* bits                                                 */
  int32_t init_grid_per_cu; /* workgroups per CU of the initialisation kernels               */
} emst_config;

#define EMST_F_HOST_SYNC 1  /* read the worklist size back every round, as ECL_MST.cu:318 */
#define EMST_F_NO_GUARD 2   /* unconditional atomicMin (ablation of ECL_MST.cu:222-223) */
#define EMST_F_NO_COMBINE 4 /* no per-workgroup LDS min-combining of keys (ablation)     */
/* Keep parallel records.  By default, after a phase's first round, each
 * k_select tile keeps only the lightest of the records that join the same
 * two components (the others close a cycle with it and cannot be in the
 * forest).  Results and round counts are unchanged; with this bit the
 * worklist sizes (round_in_sum, round_out_sum) are the reference schedule's
 * (ECL_MST.cu:210-226 keeps every crossing record). */
#define EMST_F_NO_PAIR_DEDUP 1024
/* Always use the full heavy initialisation.  By default, when the loaded
 * graph's lists are sorted and symmetric (checked by emst_load, which then
 * builds the position of every edge's upper-direction copy), the heavy
 * initialisation examines only the adjacency of vertices outside the
 * largest light-phase component and skips the rest of the CSR.  Results,
 * records and round counts are the same either way (ablation). */
#define EMST_F_NO_SPARSE_HEAVY 2048
/* Make the light phase's first offers in a separate select.  By default,
 * when the loaded graph's lists are sorted and symmetric with the same
 * weight on both copies of every edge, the light initialisation also makes
 * the first round's offers (every vertex is then a component of its own and
 * sees all its edges in its own list), and that round's select is skipped.
 * Results, records and round counts are the same either way (ablation). */
#define EMST_F_NO_LIGHT_OFFERS 4096
/* Keep the caller's vertex numbering.  By default emst_load renumbers a
 * graph with sorted, symmetric lists and solves on a copy of its CSR in the
 * new numbering (rebuilt on the device, once per load):
 *  - skewed degrees (largest degree at least 32 times the mean): by
 *    decreasing degree, so the hubs' per-vertex entries share cache lines;
 *  - scattered ids (under a tenth of the edges join ids at most 64 apart,
 *    at least 65536 vertices): by clusters grown breadth-first from every
 *    4096th vertex, ordered by (cluster, level), when the search takes at
 *    least 16 levels (a graph of high diameter, such as a geometric one).
 * Keys, records and inMST keep the caller's edge positions, so results and
 * round counts are the same either way; only the root ids that
 * emst_fetch_parent reports differ.  Read at emst_load (ablation). */
#define EMST_F_NO_RELABEL 8192
/* Renumber by decreasing degree at emst_load whenever the lists are sorted
 * and symmetric, whatever the degrees (testing). */
#define EMST_F_RELABEL 16384
/* Renumber in cluster order at emst_load whenever the lists are sorted and
 * symmetric, whatever the ids and the diameter (testing). */
#define EMST_F_CLUSTER 32768

/* The de-optimization ladder (reference: Deoptimizations/ECL_MST_No*.cu,
 * removing one optimization per file).  EMST_F_LADDER runs the reference's
 * kernel sequence instead of the fused schedule: initWL, then kernel1,
 * kernel2 and kernel3 per round with the worklist size read back by the host
 * after every kernel1 (ECL_MST.cu:304-340).  Each EMST_F_NO_* bit below
 * removes one optimization from that sequence and implies EMST_F_LADDER;
 * EMST_F_NO_GUARD and the filter mode apply to both schedules.  Any
 * combination gives the same inMST bytes as the default schedule (with
 * EMST_F_NO_ONE_DIRECTION, for a symmetric CSR, as the reference's inputs
 * are).  Filter
 * mode EMST_FILTER_LEARNED and the batched entry points take none of these
 * bits. */
#define EMST_F_LADDER 8              /* reference kernel sequence (ECL_MST.cu:304-340)             */
#define EMST_F_NO_WARP 16            /* one thread walks every adjacency in initWL; no wave        */
                                     /* cooperation at degree >= 4 (ECL_MST.cu:167, 180-205)       */
#define EMST_F_NO_IMPLICIT_PC 32     /* kernel1 keeps the original endpoints; kernel2 and kernel3  */
                                     /* find roots again with path splitting (ECL_MST.cu:218-219)  */
#define EMST_F_NO_ONE_DIRECTION 64   /* both directions of every edge enter the worklist           */
                                     /* (ECL_MST.cu:170, 195)                                      */
#define EMST_F_NO_TUPLES 128         /* worklist of edge ids plus a source array instead of        */
                                     /* {x, y, weight, j} records (ECL_MST.cu:176, 201)            */
#define EMST_F_NO_DATA_DRIVEN 256    /* every round revisits the phase's whole edge set; a flag    */
                                     /* ends the loop (ECL_MST.cu:313-323)                         */
#define EMST_F_NO_EDGE_BASED 512     /* every round loops over each vertex's adjacency; no         */
                                     /* worklist at all (ECL_MST.cu:210-251)                       */

/* Filter modes for the use_filter argument of the solve entry points. */
#define EMST_FILTER_OFF 0     /* gpuMST<false>(g, INT_MAX)                                  */
#define EMST_FILTER_PLAIN 1   /* gpuMST<true>(g, threshold), threshold from the caller       */
#define EMST_FILTER_LEARNED 2 /* threshold predicted on the device by the loaded network;    */
                              /* without a network, the plain rule of ECL_MST.cu:432-443     */

typedef struct emst_ctx emst_ctx;

/* Context on HIP device `device` with its own stream.  (ECL_MST.cu:421-427
 * hard-codes device 0 and the default stream.) */
int emst_create(int device, emst_ctx** out);
void emst_destroy(emst_ctx* ctx);
/* Text of the last error on this context; never NULL. */
const char* emst_last_error(const emst_ctx* ctx);
int emst_set_config(emst_ctx* ctx, const emst_config* cfg);
int emst_get_config(const emst_ctx* ctx, emst_config* cfg);

/* Checks nindex[0]==0, nindex non-decreasing, nindex[nodes]==edges,
 * 0<=nlist[j]<nodes.  The reference trusts its input; this library does
 * not, because a bad CSR would make the kernels read out of bounds. */
int emst_validate(const emst_graph* g);

/* Replaces the threshold selection at ECL_MST.cu:408, 432-443 (hash at
 * 389-395): *use_filter = (edges/nodes >= 4); *threshold is the
 * min(19,(int)(3.0*n*s/m))-th smallest of s=min(m,20) hashed weight
 * samples, or INT32_MAX without the filter. */
int emst_plain_threshold(const emst_graph* g, int* use_filter, int32_t* threshold);

/* Replaces gpuMST<filter>(g, threshold) (ECL_MST.cu:265-362) as a whole:
 * copies the CSR to the device, solves, copies inMST[g->edges] back.
 * use_filter is an EMST_FILTER_* mode; EMST_FILTER_OFF runs
 * gpuMST<false>(g, INT_MAX) regardless of threshold. */
int emst_gpu_mst(emst_ctx* ctx, const emst_graph* g, int use_filter, int32_t threshold,
                 uint8_t* inMST, emst_stats* stats);

/* The same call split at the reference's timed region (ECL_MST.cu:301-345):
 * emst_load does the allocations and host-to-device copies of lines
 * 268-297 (validating the CSR first), emst_solve is the device-resident
 * timed region, emst_fetch is the device-to-host copy of line 347.  A
 * loaded graph stays resident across any number of solves.  With
 * EMST_FILTER_LEARNED, emst_solve decides the threshold of a loaded graph
 * once (the network's inference on the device), before its timed region,
 * and reuses it until the next emst_load or emst_set_filter_model, as the
 * reference decides its threshold before calling gpuMST (ECL_MST.cu:432-441);
 * emst_solve_device decides it inside every solve.  For sorted, symmetric
 * lists emst_load also builds on the device the reverse positions (the
 * sparse heavy pass and the fused light pass use them) and, by the rules of
 * EMST_F_NO_RELABEL, a renumbered copy of the CSR that the solves then run
 * on; results are in the caller's layout either way. */
int emst_load(emst_ctx* ctx, const emst_graph* g);
int emst_solve(emst_ctx* ctx, int use_filter, int32_t threshold, emst_stats* stats);
int emst_fetch(emst_ctx* ctx, uint8_t* inMST);

/* Project autotuner (no counterpart in the reference, whose shape is fixed
 * at compile time, ECL_MST.cu:47-49): coordinate descent over block,
 * tile_edges, stage_recs, init_grid_per_cu, grid_per_cu and tail_max,
 * timing `reps` solves of the loaded graph per candidate (median).  The
 * best configuration is installed on the context and returned. */
int emst_autotune(emst_ctx* ctx, int use_filter, int32_t threshold, int32_t reps, emst_config* best,
                  double* best_ms);

/* Device-pointer form of emst_solve for callers that own device memory
 * (e.g. tensors).  All pointers are device pointers on the context's
 * device; d_inMST receives edges bytes.  The CSR is not validated here
 * (validate on the host first).  stream may be NULL for the context's
 * own stream. */
int emst_solve_device(emst_ctx* ctx, int32_t nodes, int32_t edges, const int32_t* d_nindex,
                      const int32_t* d_nlist, const int32_t* d_eweight, int use_filter,
                      int32_t threshold, uint8_t* d_inMST, void* stream, emst_stats* stats);

/* Allocation-free form of emst_solve_device.  No counterpart in the
 * reference (its gpuMST allocates every buffer per call, ECL_MST.cu:268-297):
 * it is for callers that cannot allocate device memory -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments and results as emst_solve_device, but every work buffer
 * (parents, minimum keys, worklists, bitmasks, counters) is carved from one
 * static device array of EMST_SCRATCH_BYTES per GPU, held until the process
 * ends, so the call makes no hipMalloc, hipFree or other allocating call.
 * A solve needs about 12.3 bytes per vertex plus 32.3 bytes per directed
 * edge with the default configuration (a 512x512 grid: 37 MB); a larger
 * graph is refused with EMST_E_LIMIT, whose message gives the bytes needed.
 * Every call on one GPU shares that GPU's array and leaves its contents
 * behind (each solve clears what must start at zero): calls that overlap on
 * one GPU, from any context or stream, race, and the caller must serialise
 * them.  The ablation ladder flags are refused (EMST_E_ARG): the ladder
 * allocates its own lists. */
#define EMST_SCRATCH_BYTES ((size_t)64 << 20)
int emst_solve_device_noalloc(emst_ctx* ctx, int32_t nodes, int32_t edges, const int32_t* d_nindex,
                              const int32_t* d_nlist, const int32_t* d_eweight, int use_filter,
                              int32_t threshold, uint8_t* d_inMST, void* stream, emst_stats* stats);

/* Parent array of the last solve (ints, nodes entries), copied to host
 * memory in the caller's numbering: after a solve every vertex's root is
 * the smallest vertex id of its tree in the numbering the solve ran in
 * (join() at ECL_MST.cu:137-144 hooks the larger root under the smaller).
 * For a graph solved on its degree-ordered copy (stats.relabelled) that is
 * the tree's member of largest degree (ties: the smallest caller id). */
int emst_fetch_parent(emst_ctx* ctx, int32_t* parent);

/* ---- batched solving (no counterpart in the reference) --------------------
 * Many independent graphs packed into HBM arenas and solved together: each
 * arena is the disjoint union of consecutive graphs with fewer than 2^31
 * edges, solved with per-graph thresholds; every graph's result is
 * bit-identical to solving it alone.  With use_graph != 0 the launch
 * sequence of each arena is captured once in a hipGraph and replayed. */
typedef struct emst_batch emst_batch;
typedef struct {
  double ms;                 /* device time of all arenas                       */
  int32_t chunks;            /* arenas (union solves)                           */
  int32_t graphs;
  int64_t nodes, edges;      /* totals over the batch                           */
  int64_t light, heavy;      /* records over all arenas                         */
  int32_t light_rounds;      /* largest per-arena round counts                  */
  int32_t heavy_rounds;
  int32_t replays;           /* arenas served by a hipGraph replay              */
  int32_t filtered_graphs;   /* graphs that used the light/heavy split          */
} emst_batch_stats;
/* Validates and uploads the graphs (host arrays stay owned by the caller).
 * max_arena_edges caps the edges of one arena (0: 2^31-2, the int32
 * limit); a single graph above the cap gets an arena of its own. */
int emst_batch_create(emst_ctx* ctx, int32_t ngraphs, const emst_graph* graphs, int64_t max_arena_edges,
                      emst_batch** out);
void emst_batch_destroy(emst_batch* b);
int emst_batch_info(const emst_batch* b, int32_t* chunks, int64_t* total_nodes, int64_t* total_edges);
/* use_filter: EMST_FILTER_OFF, EMST_FILTER_PLAIN (thresholds[g] per graph,
 * INT32_MAX for unfiltered; NULL applies the plain rule of
 * ECL_MST.cu:432-443 per graph) or EMST_FILTER_LEARNED (per-graph network
 * inference on the device). */
int emst_batch_solve(emst_batch* b, int use_filter, const int32_t* thresholds, int use_graph, emst_batch_stats* st);
/* inMST of graph g (its own edge count of bytes). */
int emst_batch_fetch(emst_batch* b, int32_t g, uint8_t* inMST);
/* Per-graph thresholds and filter flags of the last solve. */
int emst_batch_thresholds(const emst_batch* b, int32_t* thresholds, int32_t* use_filter);
/* Device time of each arena in the last solve (arenas run one after another
 * on the context's stream, and every graph of an arena completes with it),
 * and the index of each arena's first graph (arena i holds graphs
 * first_graph[i] .. first_graph[i+1]-1): up to cap entries each, the arena
 * count in *arenas.  No counterpart in the reference, which solves one graph
 * per run. */
int emst_batch_arena_ms(const emst_batch* b, double* ms, int32_t* first_graph, int32_t cap, int32_t* arenas);
/* Network used by EMST_FILTER_LEARNED for this batch (EMST_MLP_PARAMS fp32
 * values, as emst_set_filter_model); NULL falls back to the context's.  The
 * rounds of an arena are shared by all its graphs, so a graph's cost there
 * is its bytes, not its round count: a network trained for that cost
 * (filter_mlp_union.f32) belongs here. */
int emst_batch_set_filter_model(emst_batch* b, const float* params);

/* ---- learned light-edge threshold (no counterpart in the reference) -----
 * A 32-64-64-1 network predicts log2 of the light-edge quantile from 32
 * features of the resident graph; the threshold is that quantile of 1024
 * hashed weight samples.  Any threshold gives the same exact MSF; the
 * prediction only changes the work.  Parameters are EMST_MLP_PARAMS fp32
 * values (W1[32][64], b1[64], W2[64][64], b2[64], W3[64], b3), used as
 * bf16 with fp32 accumulation. */
#define EMST_MLP_PARAMS 6337
#define EMST_MLP_FEATURES 32
#define EMST_WEIGHT_SAMPLES 1024
int emst_set_filter_model(emst_ctx* ctx, const float* params); /* NULL removes the model */
/* Runs feature extraction and inference on the loaded graph.  Any output
 * pointer may be NULL; features has EMST_MLP_FEATURES floats, samples
 * EMST_WEIGHT_SAMPLES sorted weights. */
int emst_learned_threshold(emst_ctx* ctx, int* use_filter, int32_t* threshold, float* q, float* features,
                           int32_t* samples);
/* Training kernels (device `device`): forward of a batch, gradient of the
 * mean squared error of a batch, and full-batch Adam training for `epochs`
 * steps from params_inout (fp32 master weights, bf16 working copy). */
int emst_mlp_forward_batch(int device, int32_t batch, const float* params, const float* X, float* Y);
int emst_mlp_grad_batch(int device, int32_t batch, const float* params, const float* X, const float* T,
                        float* grad, float* loss);
int emst_mlp_train(int device, int32_t batch, const float* X, const float* T, int32_t epochs, float lr,
                   uint64_t seed, float* params_inout, float* final_loss);
/* bf16 bits (round to nearest even) of an fp32 value. */
uint16_t emst_bf16_bits(float x);

/* ---- ECLgraph file format: replaces ECLgraph.h:37-90 ------------------- */
int emst_read_egr(const char* path, emst_graph_buf* out);          /* ECLgraph.h:37-64 */
int emst_write_egr(const char* path, const emst_graph* g);         /* ECLgraph.h:66-80 */
void emst_free_graph(emst_graph_buf* g);                           /* ECLgraph.h:82-90 */
/* Formats of the reference's external baselines, written from a weighted
 * CSR (the layouts of its converters, <baseline>/inputs/ecl2..._genweights.cpp):
 *   emst_write_mtx     MatrixMarket real symmetric, lower triangle, 1-based
 *                      (cuGraph_GPU/inputs/ecl2mtx_genweights_cugraph.cpp)
 *   emst_write_pbbs    PBBS WeightedEdgeArray text (PBBS_CPU/inputs/ecl2pbbs_genweights.cpp)
 *   emst_write_galois  Galois/Lonestar binary, version 1 (Lonestar_CPU/inputs/ecl2lsg_genweights.cpp)
 *   emst_write_uminho  BoruvkaUMinho edge list (UMinho_BOTH/inputs/ecl2UMinho_genweights.cpp)
 *   emst_write_jucele  Jucele edge list, each undirected edge once (the layout
 *                      Jucele_GPU/JuceleMST.cu:410-440 reads; its converter is
 *                      not in the reference)
 * EMST_E_ARG without weights, EMST_E_OPEN / EMST_E_WRITE on file errors. */
int emst_write_mtx(const char* path, const emst_graph* g);
int emst_write_pbbs(const char* path, const emst_graph* g);
int emst_write_galois(const char* path, const emst_graph* g);
int emst_write_uminho(const char* path, const emst_graph* g);
int emst_write_jucele(const char* path, const emst_graph* g);
/* Weights for weightless inputs, ECL_MST.cu:410-419: eweight[j] =
 * |1 + ((i*k) % nodes)| with a wrapping 32-bit multiply.  Allocates
 * g->eweight (freed by emst_free_graph). */
int emst_synth_weights(emst_graph_buf* g);
/* Wrapping 64-bit sum of eweight[j] over inMST[j] != 0 (ECL_MST.cu:377). */
uint64_t emst_mst_weight(const emst_graph* g, const uint8_t* inMST);

/* Host minimum spanning forest, replacing cpuMST (ECL_MST.cu:82-124):
 * Kruskal over upper-direction edges in (signed weight, j) order. */
int emst_cpu_mst(const emst_graph* g, uint8_t* inMST);
/* Replaces verify (ECL_MST.cu:365-386): returns 0 when the edge sets and
 * the 64-bit weight sums agree, 1 when they differ, <0 on bad arguments. */
int emst_verify(const emst_graph* g, const uint8_t* cpu, const uint8_t* gpu, uint64_t* cpu_weight,
                uint64_t* gpu_weight);

/* ---- Related-work GPU baseline ------------------------------------------
 * Vertex-based Boruvka with graph contraction, the algorithm of the
 * reference's UMinho GPU code (UMinho_BOTH/src/BoruvkaUMinho_GPU.cu:10-455),
 * run on the context's device for comparison with gpuMST.  Loads g (as
 * emst_load), times the contraction loop with device events, and marks a
 * minimum spanning forest in inMST[edges]: its weight is the MSF weight, but
 * under equal weights its edges can differ from gpuMST's, and either
 * direction's position of an edge may be marked.  Weights compare as
 * unsigned 32-bit values.  Needs a symmetric CSR, as the reference's inputs
 * are (every edge (v, u, w) has a reverse (u, v, w), with the same
 * multiplicity): one-way edges can close a cycle of picks, so any other input
 * is refused with EMST_E_GRAPH before the device is used. */
typedef struct {
  double ms;              /* device time of the contraction loop            */
  int32_t iterations;     /* contraction iterations                         */
  int32_t color_passes;   /* pointer-jumping passes, all iterations         */
  int64_t mst_edges;      /* marked forest edges                            */
  uint64_t weight;        /* their total weight, as unsigned 32-bit values  */
} emst_contract_stats;
int emst_contract_mst(emst_ctx* ctx, const emst_graph* g, uint8_t* inMST, emst_contract_stats* stats);

/* Edge-list Boruvka with components found by label propagation, the
 * algorithm of the reference's Jucele GPU code (Jucele_GPU/JuceleMST.cu:
 * 191-355 timed host loop, 456-656 kernels), run on the context's device for
 * comparison with gpuMST.  Loads g (as emst_load), builds the list of
 * upper-direction positions (nlist[j] > source) outside the timed loop, times
 * the loop with device events and marks the forest in inMST[edges].  Ties go
 * to the smaller position and weights compare as unsigned 32-bit values, as
 * in gpuMST, so inMST is byte-identical to emst_gpu_mst's for the same graph.
 * Any CSR that emst_validate accepts is allowed. */
typedef struct {
  double ms;              /* device time of the Boruvka loop                */
  int64_t iterations;     /* iterations that joined edges                   */
  int64_t label_passes;   /* label-propagation passes, all iterations       */
  int64_t mst_edges;      /* marked forest edges                            */
  uint64_t weight;        /* their total weight, as unsigned 32-bit values  */
} emst_edgelist_stats;
int emst_edgelist_mst(emst_ctx* ctx, const emst_graph* g, uint8_t* inMST, emst_edgelist_stats* stats);

/* Library version string. */
const char* emst_version(void);

#ifdef __cplusplus
}
#endif
#endif /* EMST_H */
