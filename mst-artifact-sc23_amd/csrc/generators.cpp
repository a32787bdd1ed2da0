// generators.cpp -- seeded synthetic graphs in the ECLgraph CSR layout.
//
// Every graph is symmetric, has no self-loops or duplicate edges, keeps each
// adjacency list sorted ascending, and mirrors each undirected edge's weight
// to both directions.  All randomness is counter-based (SplitMix64 of the
// seed, a stream tag and an index), so the bytes do not depend on the
// number of worker threads.

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/emst_gen.h"
#include "host_util.hpp"

namespace {

using emst::bounded64;
using emst::mix64;
using emst::parallel_for;
using emst::stream64;
using emst::unit_double;

enum : uint64_t { TAG_W = 1, TAG_RMAT = 2, TAG_X = 3, TAG_Y = 4, TAG_U = 5, TAG_V = 6 };

struct WeightFn {
  uint64_t seed;
  int32_t lo;
  uint64_t range;
  int32_t operator()(int32_t a, int32_t b) const
  {
    const uint32_t u = (uint32_t)std::min(a, b), v = (uint32_t)std::max(a, b);
    return lo + (int32_t)bounded64(stream64(seed, TAG_W, ((uint64_t)u << 32) | v), range);
  }
};

bool alloc_graph(emst_graph_buf* g, int32_t n, int64_t m)
{
  std::memset(g, 0, sizeof(*g));
  if (m > INT32_MAX) return false;
  g->nodes = n;
  g->edges = (int32_t)m;
  g->nindex = (int32_t*)std::malloc(((size_t)n + 1) * 4);
  g->nlist = (int32_t*)std::malloc((size_t)(m ? m : 1) * 4);
  g->eweight = (int32_t*)std::malloc((size_t)(m ? m : 1) * 4);
  if (!g->nindex || !g->nlist || !g->eweight) { emst_free_graph(g); return false; }
  return true;
}

// Exclusive prefix sum of cnt[0..n) into off[0..n], in parallel.
void prefix_sum(const int64_t* cnt, int64_t* off, int64_t n)
{
  const int W = emst::host_threads();
  std::vector<int64_t> part(W + 1, 0);
  parallel_for(n, [&](int64_t b, int64_t e, int t) {
    int64_t s = 0;
    for (int64_t i = b; i < e; i++) s += cnt[i];
    part[t + 1] = s;
  }, W);
  for (int t = 0; t < W; t++) part[t + 1] += part[t];
  parallel_for(n, [&](int64_t b, int64_t e, int t) {
    int64_t s = part[t];
    for (int64_t i = b; i < e; i++) { off[i] = s; s += cnt[i]; }
  }, W);
  off[n] = part[W];
}

// Builds a symmetric, deduplicated, sorted CSR from undirected samples
// (src[i], dst[i]), dropping self-loops.
int build_csr(int32_t n, const std::vector<int32_t>& src, const std::vector<int32_t>& dst,
              const WeightFn& wf, emst_graph_buf* out)
{
  const int64_t S = (int64_t)src.size();
  std::vector<int64_t> deg((size_t)n, 0), off((size_t)n + 1, 0);
  int64_t* degp = deg.data();
  parallel_for(S, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; i++) {
      const int32_t u = src[i], v = dst[i];
      if (u == v) continue;
      __atomic_fetch_add(&degp[u], 1, __ATOMIC_RELAXED);
      __atomic_fetch_add(&degp[v], 1, __ATOMIC_RELAXED);
    }
  });
  prefix_sum(deg.data(), off.data(), n);
  std::vector<int32_t> adj((size_t)std::max<int64_t>(off[n], 1));
  std::vector<int64_t> cur(off.begin(), off.end() - 1);
  int64_t* curp = cur.data();
  int32_t* adjp = adj.data();
  parallel_for(S, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; i++) {
      const int32_t u = src[i], v = dst[i];
      if (u == v) continue;
      adjp[__atomic_fetch_add(&curp[u], 1, __ATOMIC_RELAXED)] = v;
      adjp[__atomic_fetch_add(&curp[v], 1, __ATOMIC_RELAXED)] = u;
    }
  });
  // Sort and deduplicate each list; the scatter order above is arbitrary,
  // the result after sorting is not.
  parallel_for(n, [&](int64_t b, int64_t e, int) {
    for (int64_t v = b; v < e; v++) {
      int32_t* lo = adjp + off[v];
      int32_t* hi = adjp + off[v + 1];
      std::sort(lo, hi);
      degp[v] = std::unique(lo, hi) - lo;
    }
  });
  std::vector<int64_t> noff((size_t)n + 1);
  prefix_sum(deg.data(), noff.data(), n);
  if (noff[n] > INT32_MAX) return EMST_E_LIMIT;
  if (!alloc_graph(out, n, noff[n])) return EMST_E_NOMEM;
  parallel_for(n, [&](int64_t b, int64_t e, int) {
    for (int64_t v = b; v < e; v++) {
      out->nindex[v] = (int32_t)noff[v];
      const int64_t d = degp[v];
      for (int64_t k = 0; k < d; k++) {
        const int32_t u = adjp[off[v] + k];
        out->nlist[noff[v] + k] = u;
        out->eweight[noff[v] + k] = wf((int32_t)v, u);
      }
    }
  });
  out->nindex[n] = (int32_t)noff[n];
  return EMST_OK;
}

bool weight_range_ok(int32_t wlo, int32_t whi) { return wlo >= 0 && whi > wlo; }

}  // namespace

extern "C" {

int emst_gen_grid(int32_t rows, int32_t cols, uint64_t seed, int32_t wlo, int32_t whi,
                  emst_graph_buf* out)
{
  if (!out || rows < 1 || cols < 1 || !weight_range_ok(wlo, whi)) return EMST_E_ARG;
  const int64_t n = (int64_t)rows * cols;
  const int64_t m = 2 * ((int64_t)rows * (cols - 1) + (int64_t)(rows - 1) * cols);
  if (n > INT32_MAX || m > INT32_MAX) return EMST_E_LIMIT;
  if (!alloc_graph(out, (int32_t)n, m)) return EMST_E_NOMEM;
  const WeightFn wf{seed, wlo, (uint64_t)whi - (uint64_t)wlo};
  std::vector<int64_t> deg((size_t)n), off((size_t)n + 1);
  parallel_for(n, [&](int64_t b, int64_t e, int) {
    for (int64_t id = b; id < e; id++) {
      const int64_t r = id / cols, c = id % cols;
      deg[id] = (r > 0) + (c > 0) + (c < cols - 1) + (r < rows - 1);
    }
  });
  prefix_sum(deg.data(), off.data(), n);
  parallel_for(n, [&](int64_t b, int64_t e, int) {
    for (int64_t id = b; id < e; id++) {
      const int64_t r = id / cols, c = id % cols;
      const int64_t before = off[id];
      out->nindex[id] = (int32_t)before;
      int64_t j = before;
      const int32_t v = (int32_t)id;
      auto put = [&](int64_t u) {
        out->nlist[j] = (int32_t)u;
        out->eweight[j] = wf(v, (int32_t)u);
        j++;
      };
      if (r > 0) put(id - cols);
      if (c > 0) put(id - 1);
      if (c < cols - 1) put(id + 1);
      if (r < rows - 1) put(id + cols);
    }
  });
  out->nindex[n] = (int32_t)m;
  return EMST_OK;
}

int emst_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                  uint64_t seed, int32_t wlo, int32_t whi, emst_graph_buf* out)
{
  if (!out || scale < 1 || scale > 30 || edge_factor < 1 || !weight_range_ok(wlo, whi)) return EMST_E_ARG;
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return EMST_E_ARG;
  const int32_t n = (int32_t)(1u << scale);
  const int64_t S = (int64_t)edge_factor << scale;
  std::vector<int32_t> src((size_t)S), dst((size_t)S);
  const uint64_t base = mix64(seed ^ (TAG_RMAT * 0xd1b54a32d192ed03ull));
  const double ab = a + b, abc = a + b + c;
  parallel_for(S, [&](int64_t lo, int64_t hi, int) {
    for (int64_t i = lo; i < hi; i++) {
      uint32_t u = 0, v = 0;
      const uint64_t ctr = base + (uint64_t)i * (uint64_t)scale;
      for (int l = 0; l < scale; l++) {
        const double r = unit_double(mix64(ctr + (uint64_t)l));
        const uint32_t bit = 1u << (scale - 1 - l);
        if (r < a) {
        } else if (r < ab) {
          v |= bit;
        } else if (r < abc) {
          u |= bit;
        } else {
          u |= bit;
          v |= bit;
        }
      }
      src[i] = (int32_t)u;
      dst[i] = (int32_t)v;
    }
  });
  const WeightFn wf{seed, wlo, (uint64_t)whi - (uint64_t)wlo};
  const int rc = build_csr(n, src, dst, wf, out);
  return rc;
}

int emst_gen_er(int32_t n, double mean_degree, uint64_t seed, int32_t wlo, int32_t whi,
                emst_graph_buf* out)
{
  if (!out || n < 1 || mean_degree < 0 || !weight_range_ok(wlo, whi)) return EMST_E_ARG;
  const int64_t M = (int64_t)std::llround((double)n * mean_degree / 2.0);
  if (2 * M > INT32_MAX) return EMST_E_LIMIT;
  std::vector<int32_t> src((size_t)M), dst((size_t)M);
  parallel_for(M, [&](int64_t lo, int64_t hi, int) {
    for (int64_t i = lo; i < hi; i++) {
      src[i] = (int32_t)bounded64(stream64(seed, TAG_U, (uint64_t)i), (uint64_t)n);
      dst[i] = (int32_t)bounded64(stream64(seed, TAG_V, (uint64_t)i), (uint64_t)n);
    }
  });
  const WeightFn wf{seed, wlo, (uint64_t)whi - (uint64_t)wlo};
  return build_csr(n, src, dst, wf, out);
}

int emst_gen_rgg(int32_t n, double mean_degree, uint64_t seed, int32_t wmax, emst_graph_buf* out)
{
  if (!out || n < 1 || mean_degree <= 0 || wmax < 3) return EMST_E_ARG;
  const double r = std::sqrt(mean_degree / (M_PI * (double)n));
  const double r2 = r * r;
  int64_t G = (int64_t)std::floor(1.0 / r);
  if (G < 1) G = 1;
  if (G > 65536) G = 65536;
  std::vector<double> x((size_t)n), y((size_t)n);
  std::vector<int64_t> cell((size_t)n);
  parallel_for(n, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; i++) {
      x[i] = unit_double(stream64(seed, TAG_X, (uint64_t)i));
      y[i] = unit_double(stream64(seed, TAG_Y, (uint64_t)i));
      const int64_t cx = std::min<int64_t>(G - 1, (int64_t)(x[i] * (double)G));
      const int64_t cy = std::min<int64_t>(G - 1, (int64_t)(y[i] * (double)G));
      cell[i] = cy * G + cx;
    }
  });
  // Counting sort of points by cell (stable: ascending id inside a cell).
  const int64_t C = G * G;
  std::vector<int64_t> ccount((size_t)C, 0), cstart((size_t)C + 1, 0);
  for (int64_t i = 0; i < n; i++) ccount[cell[i]]++;
  prefix_sum(ccount.data(), cstart.data(), C);
  std::vector<int32_t> order((size_t)n);
  {
    std::vector<int64_t> cur(cstart.begin(), cstart.end() - 1);
    for (int64_t i = 0; i < n; i++) order[cur[cell[i]]++] = (int32_t)i;
  }
  auto for_neighbours = [&](int64_t i, auto&& emit) {
    const int64_t cx = cell[i] % G, cy = cell[i] / G;
    for (int64_t dy = -1; dy <= 1; dy++) {
      const int64_t yy = cy + dy;
      if (yy < 0 || yy >= G) continue;
      for (int64_t dx = -1; dx <= 1; dx++) {
        const int64_t xx = cx + dx;
        if (xx < 0 || xx >= G) continue;
        const int64_t cc = yy * G + xx;
        for (int64_t k = cstart[cc]; k < cstart[cc + 1]; k++) {
          const int32_t j = order[k];
          if (j == i) continue;
          // Distance from the (min, max)-ordered pair so both directions
          // compute identical bits.
          const int32_t lo = (int32_t)std::min<int64_t>(i, j), hi = (int32_t)std::max<int64_t>(i, j);
          const double ddx = x[lo] - x[hi], ddy = y[lo] - y[hi];
          const double d2 = ddx * ddx + ddy * ddy;
          if (d2 < r2) emit(j, d2);
        }
      }
    }
  };
  std::vector<int64_t> deg((size_t)n, 0), off((size_t)n + 1, 0);
  parallel_for(n, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; i++) {
      int64_t d = 0;
      for_neighbours(i, [&](int32_t, double) { d++; });
      deg[i] = d;
    }
  });
  prefix_sum(deg.data(), off.data(), n);
  if (off[n] > INT32_MAX) return EMST_E_LIMIT;
  if (!alloc_graph(out, n, off[n])) return EMST_E_NOMEM;
  parallel_for(n, [&](int64_t b, int64_t e, int) {
    std::vector<std::pair<int32_t, int32_t>> nb;
    for (int64_t i = b; i < e; i++) {
      nb.clear();
      for_neighbours(i, [&](int32_t j, double d2) {
        const double d = std::sqrt(d2);
        int64_t w = 1 + (int64_t)std::floor(d / r * (double)(wmax - 2));
        if (w < 1) w = 1;
        if (w > wmax - 1) w = wmax - 1;
        nb.emplace_back(j, (int32_t)w);
      });
      std::sort(nb.begin(), nb.end());
      out->nindex[i] = (int32_t)off[i];
      for (size_t k = 0; k < nb.size(); k++) {
        out->nlist[off[i] + (int64_t)k] = nb[k].first;
        out->eweight[off[i] + (int64_t)k] = nb[k].second;
      }
    }
  });
  out->nindex[n] = (int32_t)off[n];
  return EMST_OK;
}

}  // extern "C"
