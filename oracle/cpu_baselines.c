/*
 * cpu_baselines.c -- parallel CPU MST codes used as timed baselines and as
 * independent cross-checks of the oracle's Kruskal.  TEST INFRASTRUCTURE /
 * BASELINE ONLY: the product never links this file.
 *
 * orc_boruvka_omp: contraction-based Boruvka in the structure of the
 *   reference's OpenMP code (UMinho_BOTH/src/BoruvkaUMinho_OMP.cpp:151-449):
 *   per super-vertex minimum edge (find_min_per_vertex, :41-61), breaking of
 *   mirrored pairs, colour propagation by pointer jumping, marking of MST
 *   edges, new ids by prefix sum, counting and inserting the surviving
 *   edges into a contracted graph.  Two deliberate changes: ties break on
 *   (weight, upper-direction CSR position j) instead of (weight, dst) at
 *   :51, so the forest equals ECL-MST's; and the loop ends when no edge
 *   crosses super-vertices (the original `next_nnodes == 1` exit at :315
 *   never fires on a forest).
 * orc_filter_kruskal_omp: the Filter-Kruskal method (Osipov, Sanders,
 *   Singler, "The Filter-Kruskal Minimum Spanning Tree Algorithm",
 *   ALENEX 2009), the method of PBBS's parallelFilterKruskal, restated:
 *   partition by a sampled pivot key, solve the light part, filter the
 *   heavy part against the current components in parallel, recurse.
 *   It is a restatement, not PBBS code.
 * Both order edges by the key ((w ^ 0x80000000) << 32) | j, i.e. signed
 * weight then j, the order of the oracle's Kruskal (ECL_MST.cu:93-102).
 */
#include <limits.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "emst_oracle.h"

static inline uint64_t okey(int32_t w, int32_t j)
{
  return ((uint64_t)((uint32_t)w ^ 0x80000000u) << 32) | (uint32_t)j;
}

/* ------------------------------------------------------------------------ */
/* Contraction Boruvka                                                      */
/* ------------------------------------------------------------------------ */

/* In-place exclusive prefix sum of a[0..n) in parallel (per-thread block
 * sums, then their offsets); returns the total. */
static int64_t par_exscan(int64_t* a, int64_t n)
{
  int64_t total = 0;
  const int nt = omp_get_max_threads();
  int64_t* part = (int64_t*)calloc((size_t)nt + 1, sizeof(int64_t));
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num(), T = omp_get_num_threads();
    const int64_t b = n * t / T, e = n * (t + 1) / T;
    int64_t acc = 0;
    for (int64_t i = b; i < e; i++) acc += a[i];
    part[t + 1] = acc;
#pragma omp barrier
#pragma omp single
    for (int i = 0; i < T; i++) part[i + 1] += part[i];
    acc = part[t];
    for (int64_t i = b; i < e; i++) {
      const int64_t x = a[i];
      a[i] = acc;
      acc += x;
    }
  }
  total = part[nt];
  free(part);
  return total;
}

int orc_boruvka_omp(const orc_graph* g, uint8_t* inMST, int threads)
{
  if (threads > 0) omp_set_num_threads(threads);
  const int32_t n0 = g->nodes;
  memset(inMST, 0, (size_t)g->edges);
  /* The current graph is a CSR over super-vertices, each entry (dst, key).
   * The reference's MST only considers upper-direction entries (n > v); a
   * lower entry v->u (u < v) stands for the same undirected edge only on a
   * symmetric input, so here each upper entry j contributes (v->u, j) and
   * (u->v, j).  Entries are placed with atomic cursors: the order inside a
   * list is not fixed, and need not be, since every key is unique. */
  int64_t* idx = (int64_t*)calloc((size_t)n0 + 1, sizeof(int64_t));
  if (!idx) return ORC_E_ALLOC;
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t v = 0; v < n0; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) {
      const int32_t u = g->nlist[j];
      if (u > v) {
        __atomic_fetch_add(&idx[v], 1, __ATOMIC_RELAXED);
        __atomic_fetch_add(&idx[u], 1, __ATOMIC_RELAXED);
      }
    }
  const int64_t m = par_exscan(idx, (int64_t)n0 + 1);
  int32_t* dst = (int32_t*)malloc((size_t)(m ? m : 1) * sizeof(int32_t));
  uint64_t* key = (uint64_t*)malloc((size_t)(m ? m : 1) * sizeof(uint64_t));
  int64_t* cur = (int64_t*)malloc(((size_t)n0 + 1) * sizeof(int64_t));
  if (!dst || !key || !cur) { free(idx); free(dst); free(key); free(cur); return ORC_E_ALLOC; }
  memcpy(cur, idx, ((size_t)n0 + 1) * sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t v = 0; v < n0; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) {
      const int32_t u = g->nlist[j];
      if (u > v) {
        const uint64_t k = okey(g->eweight[j], j);
        const int64_t a = __atomic_fetch_add(&cur[v], 1, __ATOMIC_RELAXED);
        const int64_t b = __atomic_fetch_add(&cur[u], 1, __ATOMIC_RELAXED);
        dst[a] = u; key[a] = k;
        dst[b] = v; key[b] = k;
      }
    }
  int32_t n = n0;
  int32_t* succ = (int32_t*)malloc((size_t)(n0 ? n0 : 1) * sizeof(int32_t));
  int32_t* color = (int32_t*)malloc((size_t)(n0 ? n0 : 1) * sizeof(int32_t));
  int64_t* newid = (int64_t*)malloc(((size_t)n0 + 1) * sizeof(int64_t));
  uint64_t* minkey = (uint64_t*)malloc((size_t)(n0 ? n0 : 1) * sizeof(uint64_t));
  int64_t* cnt = (int64_t*)malloc(((size_t)n0 + 1) * sizeof(int64_t));
  while (n > 1 && idx[n] > 0) {
    /* find_min_per_vertex */
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t u = 0; u < n; u++) {
      uint64_t best = UINT64_MAX;
      int32_t bd = u;
      for (int64_t e = idx[u]; e < idx[u + 1]; e++)
        if (key[e] < best) { best = key[e]; bd = dst[e]; }
      minkey[u] = best;
      succ[u] = bd;
    }
    /* break mirrored pairs: the smaller id of a mutual pair becomes root */
#pragma omp parallel for
    for (int32_t u = 0; u < n; u++) {
      const int32_t s = succ[u];
      if (s != u && succ[s] == u && u < s) succ[u] = u;
    }
    /* mark MST edges (each non-root's chosen edge) */
#pragma omp parallel for
    for (int32_t u = 0; u < n; u++)
      if (succ[u] != u) inMST[(uint32_t)minkey[u]] = 1;
    /* colour propagation by pointer jumping */
#pragma omp parallel for
    for (int32_t u = 0; u < n; u++) color[u] = succ[u];
    int changed = 1;
    while (changed) {
      changed = 0;
#pragma omp parallel for reduction(|: changed)
      for (int32_t u = 0; u < n; u++) {
        const int32_t c = color[color[u]];
        if (c != color[u]) { color[u] = c; changed = 1; }
      }
    }
    /* new ids for the roots: exclusive scan of the root flags */
#pragma omp parallel for
    for (int32_t u = 0; u < n; u++) newid[u] = color[u] == u;
    const int32_t nn = (int32_t)par_exscan(newid, n);
    /* count the surviving edges of every new super-vertex */
#pragma omp parallel for
    for (int32_t v = 0; v <= nn; v++) cnt[v] = 0;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t u = 0; u < n; u++) {
      const int32_t cu = color[u];
      int64_t c = 0;
      for (int64_t e = idx[u]; e < idx[u + 1]; e++)
        if (color[dst[e]] != cu) c++;
      if (c) __atomic_fetch_add(&cnt[newid[cu]], c, __ATOMIC_RELAXED);
    }
    const int64_t mm = par_exscan(cnt, (int64_t)nn + 1);
    int32_t* ndst = (int32_t*)malloc((size_t)(mm ? mm : 1) * sizeof(int32_t));
    uint64_t* nkey = (uint64_t*)malloc((size_t)(mm ? mm : 1) * sizeof(uint64_t));
    int64_t* pos = (int64_t*)malloc(((size_t)nn + 1) * sizeof(int64_t));
    memcpy(pos, cnt, ((size_t)nn + 1) * sizeof(int64_t));
    /* copy: each vertex reserves a block of its super-vertex's list */
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t u = 0; u < n; u++) {
      const int32_t cu = color[u];
      int64_t c = 0;
      for (int64_t e = idx[u]; e < idx[u + 1]; e++) c += color[dst[e]] != cu;
      if (!c) continue;
      int64_t at = __atomic_fetch_add(&pos[newid[cu]], c, __ATOMIC_RELAXED);
      for (int64_t e = idx[u]; e < idx[u + 1]; e++) {
        const int32_t cv = color[dst[e]];
        if (cv != cu) { ndst[at] = (int32_t)newid[cv]; nkey[at++] = key[e]; }
      }
    }
    free(pos);
    free(dst); free(key);
    dst = ndst; key = nkey;
    memcpy(idx, cnt, ((size_t)nn + 1) * sizeof(int64_t));
    n = nn;
  }
  free(idx); free(dst); free(key); free(succ); free(color); free(newid); free(minkey); free(cnt); free(cur);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Filter-Kruskal                                                           */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t k; int32_t a, b; } fk_edge;

static int32_t fk_find(int32_t v, int32_t* p)
{
  while (p[v] != v) { p[v] = p[p[v]]; v = p[v]; }
  return v;
}

static int32_t fk_find_ro(int32_t v, const int32_t* p)
{
  while (p[v] != v) v = p[v];
  return v;
}

static int fk_cmp(const void* x, const void* y)
{
  const uint64_t a = ((const fk_edge*)x)->k, b = ((const fk_edge*)y)->k;
  return (a > b) - (a < b);
}

static void fk_kruskal(fk_edge* e, int64_t n, int32_t* p, uint8_t* inMST)
{
  qsort(e, (size_t)n, sizeof(fk_edge), fk_cmp);
  for (int64_t i = 0; i < n; i++) {
    const int32_t a = fk_find(e[i].a, p), b = fk_find(e[i].b, p);
    if (a != b) {
      inMST[(uint32_t)e[i].k] = 1;
      if (a < b) p[b] = a; else p[a] = b;
    }
  }
}

static int64_t fk_partition(fk_edge* e, int64_t n, uint64_t pivot)
{
  int64_t i = 0, j = n - 1;
  while (i <= j) {
    while (i <= j && e[i].k <= pivot) i++;
    while (i <= j && e[j].k > pivot) j--;
    if (i < j) { fk_edge t = e[i]; e[i] = e[j]; e[j] = t; i++; j--; }
  }
  return i;
}

/* Removes edges whose ends already share a component (parallel, read-only
 * finds), compacting in place; returns the new length. */
static int64_t fk_filter(fk_edge* e, int64_t n, const int32_t* p, uint8_t* keep)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) keep[i] = fk_find_ro(e[i].a, p) != fk_find_ro(e[i].b, p);
  int64_t w = 0;
  for (int64_t i = 0; i < n; i++)
    if (keep[i]) e[w++] = e[i];
  return w;
}

static void fk_rec(fk_edge* e, int64_t n, int32_t* p, uint8_t* inMST, uint8_t* keep, uint64_t seed)
{
  if (n <= (1 << 16)) {
    fk_kruskal(e, n, p, inMST);
    return;
  }
  /* pivot: median of a 63-element sample */
  uint64_t s[63];
  for (int i = 0; i < 63; i++) {
    seed = seed * 6364136223846793005ull + 1442695040888963407ull;
    s[i] = e[(seed >> 17) % (uint64_t)n].k;
  }
  for (int i = 1; i < 63; i++) {
    uint64_t v = s[i];
    int k = i;
    while (k > 0 && s[k - 1] > v) { s[k] = s[k - 1]; k--; }
    s[k] = v;
  }
  const uint64_t pivot = s[31];
  const int64_t nl = fk_partition(e, n, pivot);
  if (nl == 0 || nl == n) {  /* degenerate split */
    fk_kruskal(e, n, p, inMST);
    return;
  }
  fk_rec(e, nl, p, inMST, keep, seed ^ 0x9e3779b97f4a7c15ull);
  const int64_t nh = fk_filter(e + nl, n - nl, p, keep);
  fk_rec(e + nl, nh, p, inMST, keep, seed + 1);
}

int orc_filter_kruskal_omp(const orc_graph* g, uint8_t* inMST, int threads)
{
  if (threads > 0) omp_set_num_threads(threads);
  memset(inMST, 0, (size_t)g->edges);
  int64_t m = 0;
  for (int32_t v = 0; v < g->nodes; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) m += g->nlist[j] > v;
  fk_edge* e = (fk_edge*)malloc((size_t)(m ? m : 1) * sizeof(fk_edge));
  int32_t* p = (int32_t*)malloc((size_t)g->nodes * sizeof(int32_t));
  uint8_t* keep = (uint8_t*)malloc((size_t)(m ? m : 1));
  if (!e || !p || !keep) { free(e); free(p); free(keep); return ORC_E_ALLOC; }
  int64_t k = 0;
  for (int32_t v = 0; v < g->nodes; v++) {
    p[v] = v;
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++)
      if (g->nlist[j] > v) { e[k].k = okey(g->eweight[j], j); e[k].a = v; e[k].b = g->nlist[j]; k++; }
  }
  fk_rec(e, m, p, inMST, keep, 12345);
  free(e); free(p); free(keep);
  return 0;
}
