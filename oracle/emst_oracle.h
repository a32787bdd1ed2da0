/*
 * emst_oracle.h -- CPU restatement of ECL-MST's gpuMST path.
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; the product never links it.
 */
#ifndef EMST_ORACLE_H
#define EMST_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_E_OPEN = -1, ORC_E_NODES = -2, ORC_E_EDGES = -3, ORC_E_COUNTS = -4,
  ORC_E_ALLOC = -5, ORC_E_NINDEX = -6, ORC_E_NLIST = -7, ORC_E_WEIGHTS = -8
};

#define ORC_MAX_ROUNDS 256

typedef struct {
  int32_t nodes, edges;
  const int32_t* nindex;
  const int32_t* nlist;
  const int32_t* eweight;
} orc_graph;

typedef struct {
  int32_t nodes, edges;
  int32_t* nindex;
  int32_t* nlist;
  int32_t* eweight;   /* NULL when the file holds no weights */
} orc_graph_buf;

/* Deterministic quantities of the GPU round schedule. */
typedef struct {
  int64_t light;            /* L: records written by the light pass            */
  int64_t heavy_all;        /* H_all: upper-direction edges with w > T         */
  int64_t heavy;            /* H: heavy records whose reps differ              */
  int light_rounds, heavy_rounds;
  int64_t light_w[ORC_MAX_ROUNDS], light_s[ORC_MAX_ROUNDS];   /* kernel1 in/out sizes */
  int64_t heavy_w[ORC_MAX_ROUNDS], heavy_s[ORC_MAX_ROUNDS];
  int64_t components;       /* c: trees in the forest (isolated vertices count) */
  int64_t mst_edges;        /* n - c                                            */
  uint64_t weight;          /* 64-bit wrapping sum of marked weights            */
} orc_sched;

int orc_read_egr(const char* path, orc_graph_buf* out);
void orc_free_graph(orc_graph_buf* g);
void orc_synth_weights(int32_t nodes, const int32_t* nindex, const int32_t* nlist, int32_t* out);
uint32_t orc_hash(uint32_t v);
void orc_plain_threshold(int32_t nodes, int32_t edges, const int32_t* eweight,
                         int* use_filter, int32_t* threshold);
int orc_kruskal(const orc_graph* g, uint8_t* inMST);
int orc_ecl_schedule(const orc_graph* g, int use_filter, int32_t threshold,
                     uint8_t* inMST, orc_sched* st);
/* The same schedule with position j keyed by kid[j] and inMST indexed by
 * those ids (a renumbered copy of a graph keyed by its original positions). */
int orc_ecl_schedule_keyed(const orc_graph* g, int use_filter, int32_t threshold, const int32_t* kid,
                           uint8_t* inMST, orc_sched* st);
int orc_verify(const orc_graph* g, const uint8_t* cpu, const uint8_t* gpu,
               uint64_t* cpu_weight, uint64_t* gpu_weight, int64_t* only_cpu, int64_t* only_gpu);

int orc_forest_check(const orc_graph* g, const uint8_t* inMST, int64_t* marked, int64_t* components,
                     int* is_forest);

/* Parallel CPU baselines (cpu_baselines.c); threads <= 0 keeps the OpenMP default. */
int orc_boruvka_omp(const orc_graph* g, uint8_t* inMST, int threads);
int orc_filter_kruskal_omp(const orc_graph* g, uint8_t* inMST, int threads);

#ifdef __cplusplus
}
#endif
#endif
