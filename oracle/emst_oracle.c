/*
 * emst_oracle.c -- CPU restatement of ECL-MST's gpuMST path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker (or as the timed CPU
 * baseline), never as the thing measured or shipped.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference checkout (alexfallin/MST-Artifact-SC23).
 *
 * Pinning (what checks each function, besides retyping it):
 *   orc_read_egr       : the reference's own reader (ECL-MST/ECLgraph.h),
 *                        compiled from its source into oracle/_ref, on byte
 *                        fixtures covering every error path; plus the format
 *                        arithmetic (file size = 8 + 4(n+1) + 8m).
 *   orc_synth_weights  : hand-computed golden vectors incl. int32 wraparound
 *                        and negative-remainder cases (tests/golden).
 *   orc_hash           : known answers computed independently with big
 *                        integers reduced mod 2^32 (tests/golden/golden.json).
 *   orc_plain_threshold: closed-form sample indices worked by hand for
 *                        (n, m) = (10, 40), (100, 6000), (1, 4) and m/n < 4.
 *   orc_kruskal        : brute-force MSF enumeration for n <= 7, closed forms
 *                        (trees, all-equal weights, isolated vertices), the
 *                        independent Boruvka/filter-Kruskal in cpu_baselines.c.
 *   orc_ecl_schedule   : equals orc_kruskal whenever weights are
 *                        non-negative (the MSF under the (w, j) order is
 *                        unique); per-round sizes checked against a hand
 *                        trace of a path graph and against invariants.
 *   orc_ecl_schedule_keyed: with kid[j] = j it is orc_ecl_schedule; on a
 *                        renumbered copy keyed by the original positions it
 *                        reproduces the original bytes and round sizes (the
 *                        keys are distinct, so each phase's forest and round
 *                        structure are fixed), while keying the copy by its
 *                        own positions does not.
 *   orc_verify         : direct cases (equal, one flipped bit, weight sum).
 *   orc_forest_check   : Kruskal's forest passes; dropping an edge breaks
 *                        spanning, adding a non-tree edge makes a cycle.
 */
#include <limits.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "emst_oracle.h"

/* ------------------------------------------------------------------------ */
/* ECLgraph binary format reader.  Restates ECL-MST/ECLgraph.h:37-64.        */
/* Layout: int32 nodes, int32 edges, int32 nindex[nodes+1],                  */
/* int32 nlist[edges], optional int32 eweight[edges].                        */
/* ------------------------------------------------------------------------ */
int orc_read_egr(const char* path, orc_graph_buf* out)
{
  memset(out, 0, sizeof(*out));
  FILE* f = fopen(path, "rb");                                   /* :42 */
  if (!f) return ORC_E_OPEN;
  int32_t nodes, edges;
  if (fread(&nodes, 4, 1, f) != 1) { fclose(f); return ORC_E_NODES; } /* :43 */
  if (fread(&edges, 4, 1, f) != 1) { fclose(f); return ORC_E_EDGES; } /* :44 */
  if (nodes < 1 || edges < 0) { fclose(f); return ORC_E_COUNTS; }     /* :45 */
  int32_t* nindex = (int32_t*)malloc(((size_t)nodes + 1) * 4);
  int32_t* nlist = (int32_t*)malloc((size_t)(edges ? edges : 1) * 4);
  int32_t* ew = (int32_t*)malloc((size_t)(edges ? edges : 1) * 4);
  if (!nindex || !nlist || !ew) {                                     /* :50 */
    free(nindex); free(nlist); free(ew); fclose(f); return ORC_E_ALLOC;
  }
  if (fread(nindex, 4, (size_t)nodes + 1, f) != (size_t)nodes + 1) {  /* :52 */
    free(nindex); free(nlist); free(ew); fclose(f); return ORC_E_NINDEX;
  }
  if (fread(nlist, 4, (size_t)edges, f) != (size_t)edges) {           /* :53 */
    free(nindex); free(nlist); free(ew); fclose(f); return ORC_E_NLIST;
  }
  size_t cnt = fread(ew, 4, (size_t)edges, f);                        /* :54 */
  fclose(f);
  if (cnt == 0) {                                                     /* :55-57 */
    free(ew);
    ew = NULL;
  } else if (cnt != (size_t)edges) {                                  /* :59 */
    free(nindex); free(nlist); free(ew); return ORC_E_WEIGHTS;
  }
  out->nodes = nodes;
  out->edges = edges;
  out->nindex = nindex;
  out->nlist = nlist;
  out->eweight = ew;
  return 0;
}

void orc_free_graph(orc_graph_buf* g)
{
  free(g->nindex); free(g->nlist); free(g->eweight);
  memset(g, 0, sizeof(*g));
}

/* Synthetic weights for weightless inputs.  Restates ECL_MST.cu:410-419:
 * w = 1 + ((i * k) % nodes) with a wrapping 32-bit multiply and C's
 * truncating remainder, then the absolute value. */
void orc_synth_weights(int32_t nodes, const int32_t* nindex, const int32_t* nlist, int32_t* out)
{
  for (int32_t i = 0; i < nodes; i++) {
    for (int32_t j = nindex[i]; j < nindex[i + 1]; j++) {
      const int32_t k = nlist[j];
      const int32_t prod = (int32_t)((uint32_t)i * (uint32_t)k);
      int32_t w = 1 + (prod % nodes);
      if (w < 0) w = -w;
      out[j] = w;
    }
  }
}

/* Integer hash used to pick threshold samples.  Restates ECL_MST.cu:390-395. */
uint32_t orc_hash(uint32_t v)
{
  v = ((v >> 16) ^ v) * 0x45d9f3bu;
  v = ((v >> 16) ^ v) * 0x45d9f3bu;
  return (v >> 16) ^ v;
}

static int cmp_i32(const void* a, const void* b)
{
  const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* Plain light-edge threshold.  Restates ECL_MST.cu:408, 432-443: filter only
 * when the integer average degree is at least 4; the threshold is the
 * min(19, (int)(3.0*n*s/m))-th smallest of s = min(m, 20) hashed samples. */
void orc_plain_threshold(int32_t nodes, int32_t edges, const int32_t* eweight,
                         int* use_filter, int32_t* threshold)
{
  const int32_t avg_deg = edges / nodes;
  if (avg_deg >= 4) {
    int32_t sw[20];
    const int32_t s = edges < 20 ? edges : 20;
    for (int32_t i = 0; i < s; i++) sw[i] = eweight[orc_hash((uint32_t)i) % (uint32_t)edges];
    qsort(sw, (size_t)s, 4, cmp_i32);
    int32_t idx = (int32_t)(3.0 * nodes * s / edges);
    if (idx > 19) idx = 19;
    *use_filter = 1;
    *threshold = sw[idx];
  } else {
    *use_filter = 0;
    *threshold = INT_MAX;
  }
}

/* ------------------------------------------------------------------------ */
/* Serial Kruskal.  Restates ECL_MST.cu:55-124.                              */
/* ------------------------------------------------------------------------ */
static int32_t serial_find(int32_t idx, int32_t* parent)         /* :55-67 */
{
  int32_t curr = parent[idx];
  if (curr != idx) {
    int32_t next, prev = idx;
    while (curr != (next = parent[curr])) {
      parent[prev] = next;
      prev = curr;
      curr = next;
    }
  }
  return curr;
}

static void serial_join(int32_t a, int32_t b, int32_t* parent)   /* :70-79 */
{
  const int32_t arep = serial_find(a, parent);
  const int32_t brep = serial_find(b, parent);
  if (arep > brep) parent[brep] = arep; else parent[arep] = brep;
}

/* LSD radix sort of 64-bit keys, 16-bit digits. */
static void radix_sort_u64(uint64_t* a, uint64_t* tmp, size_t n)
{
  size_t* cnt = (size_t*)malloc(65536 * sizeof(size_t));
  for (int pass = 0; pass < 4; pass++) {
    const int sh = pass * 16;
    memset(cnt, 0, 65536 * sizeof(size_t));
    for (size_t i = 0; i < n; i++) cnt[(a[i] >> sh) & 0xffff]++;
    size_t s = 0;
    for (int d = 0; d < 65536; d++) { size_t c = cnt[d]; cnt[d] = s; s += c; }
    for (size_t i = 0; i < n; i++) tmp[cnt[(a[i] >> sh) & 0xffff]++] = a[i];
    uint64_t* t = a; a = tmp; tmp = t;
  }
  free(cnt);  /* 4 passes: result is back in the caller's array */
}

/* The reference sorts std::tuple<weight, j, from, to> lexicographically
 * (:93-102).  j is unique per upper-direction edge, so the order is
 * (signed weight, j); the key below encodes exactly that order. */
int orc_kruskal(const orc_graph* g, uint8_t* inMST)
{
  const int32_t n = g->nodes;
  memset(inMST, 0, (size_t)g->edges);
  int32_t* parent = (int32_t*)malloc((size_t)n * 4);
  int32_t* src = (int32_t*)malloc((size_t)(g->edges ? g->edges : 1) * 4);
  if (!parent || !src) { free(parent); free(src); return ORC_E_ALLOC; }
  size_t cnt = 0;
  for (int32_t i = 0; i < n; i++) {
    parent[i] = i;
    for (int32_t j = g->nindex[i]; j < g->nindex[i + 1]; j++) {
      src[j] = i;
      if (g->nlist[j] > i) cnt++;                               /* :97 */
    }
  }
  uint64_t* keys = (uint64_t*)malloc((cnt ? cnt : 1) * 8);
  uint64_t* tmp = (uint64_t*)malloc((cnt ? cnt : 1) * 8);
  if (!keys || !tmp) { free(parent); free(src); free(keys); free(tmp); return ORC_E_ALLOC; }
  size_t k = 0;
  for (int32_t i = 0; i < n; i++)
    for (int32_t j = g->nindex[i]; j < g->nindex[i + 1]; j++)
      if (g->nlist[j] > i)
        keys[k++] = ((uint64_t)((uint32_t)g->eweight[j] ^ 0x80000000u) << 32) | (uint32_t)j;
  radix_sort_u64(keys, tmp, cnt);                                 /* :102 */
  int32_t count = n - 1;                                          /* :104 */
  for (size_t p = 0; p < cnt && count > 0; p++) {
    const int32_t j = (int32_t)(uint32_t)keys[p];
    const int32_t a = src[j], b = g->nlist[j];
    const int32_t arep = serial_find(a, parent);
    const int32_t brep = serial_find(b, parent);
    if (arep != brep) {                                           /* :110-115 */
      inMST[j] = 1;
      serial_join(arep, brep, parent);
      count--;
    }
  }
  free(parent); free(src); free(keys); free(tmp);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Emulator of the GPU round schedule.  Restates the semantics of            */
/* ECL_MST.cu:127-251 (find, join, initWL, kernel1/2/3) and the host loop    */
/* 304-340.  The GPU orders candidate edges by the unsigned 64-bit key       */
/* ((uint64)w << 32) | j built at :221 and :234 (an int weight converted to  */
/* unsigned long long, so negative weights order above all non-negative      */
/* ones), splits light/heavy by the signed test w <= T at :172/:197, and     */
/* marks only the upper-direction position j (:170, :195, :237).             */
/* ------------------------------------------------------------------------ */
static int32_t uf_find(int32_t v, int32_t* p)
{
  while (p[v] != v) { p[v] = p[p[v]]; v = p[v]; }
  return v;
}

static void uf_union_min(int32_t a, int32_t b, int32_t* p)
{
  a = uf_find(a, p); b = uf_find(b, p);
  if (a == b) return;
  if (a < b) p[b] = a; else p[a] = b;   /* the root is the smallest id, as join() at :137-144 */
}

typedef struct { int32_t x, y; uint64_t key; } orc_rec;

static uint64_t ecl_key(int32_t w, int32_t j)
{
  return ((uint64_t)(int64_t)w << 32) | (uint64_t)(uint32_t)j;   /* :221 */
}

/* Runs kernel1/kernel2 rounds until the worklist is empty (:313-323). */
static int run_rounds(orc_rec* wl, size_t len, int32_t* parent, uint64_t* minv,
                      uint8_t* inMST, int32_t n, int64_t* wr, int64_t* sr,
                      int* nrounds, int maxr)
{
  (void)n;
  int r = 0;
  while (len > 0) {
    /* kernel1 (:210-226): drop intra-component records, compress reps. */
    size_t s = 0;
    for (size_t i = 0; i < len; i++) {
      const int32_t a = uf_find(wl[i].x, parent), b = uf_find(wl[i].y, parent);
      if (a != b) {
        wl[s].x = a; wl[s].y = b; wl[s].key = wl[i].key; s++;
        if (minv[a] > wl[i].key) minv[a] = wl[i].key;
        if (minv[b] > wl[i].key) minv[b] = wl[i].key;
      }
    }
    if (r < maxr) { wr[r] = (int64_t)len; sr[r] = (int64_t)s; }
    r++;
    len = s;
    if (len == 0) break;
    /* kernel2 (:229-240): join on min edges; kernel3 (:243-251): reset. */
    for (size_t i = 0; i < len; i++) {
      const uint64_t key = wl[i].key;
      if (key == minv[wl[i].x] || key == minv[wl[i].y]) {
        uf_union_min(wl[i].x, wl[i].y, parent);
        inMST[(uint32_t)key] = 1;
      }
    }
    for (size_t i = 0; i < len; i++) { minv[wl[i].x] = UINT64_MAX; minv[wl[i].y] = UINT64_MAX; }
  }
  *nrounds = r;
  return 0;
}

/* The schedule with every position keyed by an id of the caller's choosing:
 * kid == NULL keys position j by j (the reference, :173-176); otherwise by
 * kid[j], the key id a renumbered copy of a graph carries (the position of
 * the edge's upper-direction copy in the original CSR), and inMST is indexed
 * by those ids. */
static int schedule_impl(const orc_graph* g, int use_filter, int32_t threshold, const int32_t* kid,
                         uint8_t* inMST, orc_sched* st)
{
  const int32_t n = g->nodes, m = g->edges;
  memset(st, 0, sizeof(*st));
  memset(inMST, 0, (size_t)m);
  int32_t* parent = (int32_t*)malloc((size_t)n * 4);
  uint64_t* minv = (uint64_t*)malloc((size_t)n * 8);
  size_t ups = 0;
  for (int32_t v = 0; v < n; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++)
      if (g->nlist[j] > v) ups++;
  orc_rec* wl = (orc_rec*)malloc((ups ? ups : 1) * sizeof(orc_rec));
  if (!parent || !minv || !wl) { free(parent); free(minv); free(wl); return ORC_E_ALLOC; }
  for (int32_t v = 0; v < n; v++) { parent[v] = v; minv[v] = UINT64_MAX; }   /* initPM :147-154 */
  const int32_t T = use_filter ? threshold : INT_MAX;

  /* initWL<true> (:157-207): every upper-direction edge with w <= T. */
  size_t len = 0;
  for (int32_t v = 0; v < n; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) {
      const int32_t u = g->nlist[j];
      if (u > v && g->eweight[j] <= T) {
        wl[len].x = v; wl[len].y = u; wl[len].key = ecl_key(g->eweight[j], kid ? kid[j] : j); len++;
      }
    }
  st->light = (int64_t)len;
  run_rounds(wl, len, parent, minv, inMST, n, st->light_w, st->light_s,
             &st->light_rounds, ORC_MAX_ROUNDS);

  if (use_filter) {
    /* initWL<false>: heavy upper-direction edges whose reps differ. */
    len = 0;
    for (int32_t v = 0; v < n; v++) {
      const int32_t a = uf_find(v, parent);
      for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) {
        const int32_t u = g->nlist[j];
        if (u > v && g->eweight[j] > T) {
          st->heavy_all++;
          const int32_t b = uf_find(u, parent);
          if (a != b) {
            wl[len].x = a; wl[len].y = b; wl[len].key = ecl_key(g->eweight[j], kid ? kid[j] : j); len++;
          }
        }
      }
    }
    st->heavy = (int64_t)len;
    run_rounds(wl, len, parent, minv, inMST, n, st->heavy_w, st->heavy_s,
               &st->heavy_rounds, ORC_MAX_ROUNDS);
  }
  int64_t comps = 0, mst = 0;
  uint64_t wsum = 0;
  for (int32_t v = 0; v < n; v++) if (uf_find(v, parent) == v) comps++;
  /* marked edges and their weights, counted at each edge's upper copy */
  for (int32_t v = 0; v < n; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++)
      if (g->nlist[j] > v && inMST[kid ? kid[j] : j]) { mst++; wsum += (uint64_t)(int64_t)g->eweight[j]; }
  st->components = comps;
  st->mst_edges = mst;
  st->weight = wsum;
  free(parent); free(minv); free(wl);
  return 0;
}

int orc_ecl_schedule(const orc_graph* g, int use_filter, int32_t threshold,
                     uint8_t* inMST, orc_sched* st)
{
  return schedule_impl(g, use_filter, threshold, NULL, inMST, st);
}

int orc_ecl_schedule_keyed(const orc_graph* g, int use_filter, int32_t threshold, const int32_t* kid,
                           uint8_t* inMST, orc_sched* st)
{
  return schedule_impl(g, use_filter, threshold, kid, inMST, st);
}

/* Restates verify (ECL_MST.cu:365-386).  Returns 0 when both edge sets and
 * both 64-bit weight sums agree, 1 otherwise ("results differ"). */
int orc_verify(const orc_graph* g, const uint8_t* cpu, const uint8_t* gpu,
               uint64_t* cpu_weight, uint64_t* gpu_weight, int64_t* only_cpu, int64_t* only_gpu)
{
  uint64_t cw = 0, gw = 0;
  int64_t oc = 0, og = 0;
  for (int32_t j = 0; j < g->edges; j++) {
    const int a = cpu[j] != 0, b = gpu[j] != 0;
    if (!a && b) og++;
    if (a && !b) oc++;
    if (b) gw += (uint64_t)(int64_t)g->eweight[j];
    if (a) cw += (uint64_t)(int64_t)g->eweight[j];
  }
  if (cpu_weight) *cpu_weight = cw;
  if (gpu_weight) *gpu_weight = gw;
  if (only_cpu) *only_cpu = oc;
  if (only_gpu) *only_gpu = og;
  return (cw != gw || oc != 0 || og != 0) ? 1 : 0;
}

/* Size-independent checks of a candidate forest (no reference counterpart;
 * pinned by the unit tests against Kruskal): union-find over the marked
 * edges detects a cycle (not a forest) or a marked non-upper edge; a second
 * union-find over all upper-direction edges counts the components c of the
 * graph.  A spanning forest has exactly n - c edges. */
int orc_forest_check(const orc_graph* g, const uint8_t* inMST, int64_t* marked, int64_t* components,
                     int* is_forest)
{
  const int32_t n = g->nodes;
  int32_t* p = (int32_t*)malloc((size_t)n * 4);
  if (!p) return ORC_E_ALLOC;
  for (int32_t v = 0; v < n; v++) p[v] = v;
  int64_t mk = 0;
  int ok = 1;
  for (int32_t v = 0; v < n; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) {
      if (!inMST[j]) continue;
      mk++;
      const int32_t u = g->nlist[j];
      if (u <= v) { ok = 0; continue; }
      const int32_t a = uf_find(v, p), b = uf_find(u, p);
      if (a == b) ok = 0; else uf_union_min(a, b, p);
    }
  for (int32_t v = 0; v < n; v++) p[v] = v;
  int64_t comps = n;
  for (int32_t v = 0; v < n; v++)
    for (int32_t j = g->nindex[v]; j < g->nindex[v + 1]; j++) {
      const int32_t u = g->nlist[j];
      if (u > v) {
        const int32_t a = uf_find(v, p), b = uf_find(u, p);
        if (a != b) { uf_union_min(a, b, p); comps--; }
      }
    }
  free(p);
  *marked = mk;
  *components = comps;
  *is_forest = ok;
  return 0;
}
