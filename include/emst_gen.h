/*
 * emst_gen.h -- seeded synthetic graph generators (part of libemst.so).
 *
 * No counterpart in the reference, which downloads its 17 inputs
 * (Inputs/download_inputs.sh:4-20).  These produce graphs in the same
 * ECLgraph CSR layout so that the solver's workloads can be built offline:
 * symmetric, no self-loops, no duplicate edges, each adjacency list sorted
 * ascending, and each undirected edge's weight drawn once from a
 * counter-based hash of (seed, min(u,v), max(u,v)) and mirrored to both
 * directions.  Output is a pure function of the arguments (any thread
 * count gives identical bytes).
 */
#ifndef EMST_GEN_H
#define EMST_GEN_H

#include <stdint.h>
#include "emst.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rows x cols 4-neighbour grid, vertex id r*cols+c; weights uniform in
 * [wlo, whi). */
int emst_gen_grid(int32_t rows, int32_t cols, uint64_t seed, int32_t wlo, int32_t whi,
                  emst_graph_buf* out);

/* RMAT: edge_factor * 2^scale samples, each descending `scale` levels
 * with quadrant probabilities (a, b, c, 1-a-b-c); symmetrized,
 * self-loops and duplicates removed; weights uniform in [wlo, whi). */
int emst_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                  uint64_t seed, int32_t wlo, int32_t whi, emst_graph_buf* out);

/* Random geometric graph: n points uniform in the unit square, radius
 * r = sqrt(mean_degree / (pi * n)), edge iff distance < r; weight
 * 1 + floor((d / r) * (wmax - 1)) in [1, wmax). */
int emst_gen_rgg(int32_t n, double mean_degree, uint64_t seed, int32_t wmax, emst_graph_buf* out);

/* Erdos-Renyi G(n, M) with M = round(n * mean_degree / 2) uniform vertex
 * pairs, symmetrized, self-loops and duplicates removed; weights uniform
 * in [wlo, whi). */
int emst_gen_er(int32_t n, double mean_degree, uint64_t seed, int32_t wlo, int32_t whi,
                emst_graph_buf* out);

/* Worker threads used by the generators (0 = hardware concurrency). */
void emst_gen_set_threads(int32_t threads);

#ifdef __cplusplus
}
#endif
#endif /* EMST_GEN_H */
