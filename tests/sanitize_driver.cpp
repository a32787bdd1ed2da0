// sanitize_driver.cpp -- host-code exerciser for the sanitizer test: built
// with -fsanitize=address,undefined from the library's host sources
// (reader, writers, generators, validation, threshold, Kruskal, verify) and
// the oracle, it runs them on the reader fixtures, on generated graphs, on
// every writer, and on byte-mutated copies of a valid file, comparing the
// library with the oracle where both answer.  Any sanitizer report fails
// the test; so does a mismatch (exit status 1).
//
//   sanitize_driver <fixture dir> <scratch dir>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dirent.h>
#include <string>
#include <vector>

#include "../include/emst.h"
#include "../include/emst_gen.h"
#include "../oracle/emst_oracle.h"

static int failures = 0;

#define CHECK(cond, ...)                       \
  do {                                         \
    if (!(cond)) {                             \
      std::fprintf(stderr, "MISMATCH: ");      \
      std::fprintf(stderr, __VA_ARGS__);       \
      std::fprintf(stderr, "\n");              \
      failures++;                              \
    }                                          \
  } while (0)

static emst_graph view(const emst_graph_buf& b) { return {b.nodes, b.edges, b.nindex, b.nlist, b.eweight}; }
static orc_graph oview(const emst_graph_buf& b) { return {b.nodes, b.edges, b.nindex, b.nlist, b.eweight}; }

// Library and oracle forests on a valid weighted graph.
static void check_forests(const emst_graph_buf& b, const char* what)
{
  const emst_graph g = view(b);
  const orc_graph og = oview(b);
  const size_t m = (size_t)(b.edges > 0 ? b.edges : 1);
  std::vector<uint8_t> lib(m), orc(m), sched(m), bor(m), fk(m);
  CHECK(emst_cpu_mst(&g, lib.data()) == EMST_OK, "%s: cpu_mst failed", what);
  CHECK(orc_kruskal(&og, orc.data()) == 0, "%s: oracle kruskal failed", what);
  CHECK(std::memcmp(lib.data(), orc.data(), (size_t)b.edges) == 0, "%s: kruskal differs", what);
  orc_sched st;
  int use_filter = 0;
  int32_t T = 0;
  CHECK(emst_plain_threshold(&g, &use_filter, &T) == EMST_OK, "%s: plain threshold failed", what);
  CHECK(orc_ecl_schedule(&og, use_filter, T, sched.data(), &st) == 0, "%s: schedule failed", what);
  CHECK(orc_boruvka_omp(&og, bor.data(), 2) == 0 && orc_filter_kruskal_omp(&og, fk.data(), 2) == 0,
        "%s: openmp restatements failed", what);
  uint64_t cw = 0, gw = 0;
  CHECK(emst_verify(&g, lib.data(), lib.data(), &cw, &gw) == 0 && cw == gw, "%s: verify of equal forests", what);
  CHECK(emst_mst_weight(&g, lib.data()) == cw, "%s: mst weight", what);
}

static void exercise_writers(const emst_graph_buf& b, const std::string& dir, const char* what)
{
  const emst_graph g = view(b);
  const std::string base = dir + "/w_" + what;
  CHECK(emst_write_mtx((base + ".mtx").c_str(), &g) == EMST_OK, "%s: mtx", what);
  CHECK(emst_write_pbbs((base + ".pbbs").c_str(), &g) == EMST_OK, "%s: pbbs", what);
  CHECK(emst_write_galois((base + ".gr").c_str(), &g) == EMST_OK, "%s: galois", what);
  CHECK(emst_write_uminho((base + ".um").c_str(), &g) == EMST_OK, "%s: uminho", what);
  CHECK(emst_write_jucele((base + ".jmst").c_str(), &g) == EMST_OK, "%s: jucele", what);
  CHECK(emst_write_egr((base + ".egr").c_str(), &g) == EMST_OK, "%s: egr", what);
  emst_graph_buf r{};
  CHECK(emst_read_egr((base + ".egr").c_str(), &r) == EMST_OK, "%s: read back", what);
  CHECK(r.nodes == b.nodes && r.edges == b.edges, "%s: read back sizes", what);
  if (r.nodes == b.nodes && r.edges == b.edges) {
    CHECK(std::memcmp(r.nindex, b.nindex, 4 * ((size_t)b.nodes + 1)) == 0, "%s: nindex", what);
    CHECK(b.edges == 0 || std::memcmp(r.nlist, b.nlist, 4 * (size_t)b.edges) == 0, "%s: nlist", what);
  }
  emst_free_graph(&r);
}

static std::vector<unsigned char> slurp(const std::string& p)
{
  std::vector<unsigned char> v;
  FILE* f = std::fopen(p.c_str(), "rb");
  if (!f) return v;
  unsigned char buf[4096];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) v.insert(v.end(), buf, buf + k);
  std::fclose(f);
  return v;
}

int main(int argc, char** argv)
{
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <fixture dir> <scratch dir>\n", argv[0]);
    return 2;
  }
  const std::string fixtures = argv[1], scratch = argv[2];
  emst_gen_set_threads(2);

  // 1. every reader fixture through both readers
  int files = 0;
  if (DIR* d = opendir(fixtures.c_str())) {
    while (dirent* e = readdir(d)) {
      const std::string name = e->d_name;
      if (name.size() < 5 || name.substr(name.size() - 4) != ".egr") continue;
      const std::string p = fixtures + "/" + name;
      emst_graph_buf b{};
      orc_graph_buf o{};
      const int rc = emst_read_egr(p.c_str(), &b), orc = orc_read_egr(p.c_str(), &o);
      CHECK((rc == EMST_OK) == (orc == 0), "%s: reader outcome %d vs oracle %d", name.c_str(), rc, orc);
      const emst_graph v = view(b);
      if (rc == EMST_OK && emst_validate(&v) == EMST_OK) {
        if (!b.eweight) CHECK(emst_synth_weights(&b) == EMST_OK, "%s: synth weights", name.c_str());
        check_forests(b, name.c_str());
      }
      emst_free_graph(&b);
      orc_free_graph(&o);
      files++;
    }
    closedir(d);
  }
  CHECK(files > 0, "no fixtures read from %s", fixtures.c_str());

  // 2. generated graphs: forests, writers, round trips
  struct Gen {
    const char* name;
    int rc;
    emst_graph_buf b;
  } gens[4] = {{"grid", 0, {}}, {"rmat", 0, {}}, {"rgg", 0, {}}, {"er", 0, {}}};
  gens[0].rc = emst_gen_grid(24, 31, 1, 1, 50, &gens[0].b);
  gens[1].rc = emst_gen_rmat(10, 8, 0.57, 0.19, 0.19, 2, 1, 40, &gens[1].b);
  gens[2].rc = emst_gen_rgg(3000, 6.0, 3, 1 << 20, &gens[2].b);
  gens[3].rc = emst_gen_er(2000, 1.5, 4, 1, 9, &gens[3].b);
  for (auto& gg : gens) {
    CHECK(gg.rc == EMST_OK, "%s: generator failed", gg.name);
    if (gg.rc != EMST_OK) continue;
    const emst_graph v = view(gg.b);
    CHECK(emst_validate(&v) == EMST_OK, "%s: generated graph invalid", gg.name);
    check_forests(gg.b, gg.name);
    exercise_writers(gg.b, scratch, gg.name);
  }

  // 3. byte-mutated copies of a valid file: the reader and validation must
  // reject or accept without touching memory they do not own, and an
  // accepted graph must give the oracle's forest
  const std::string src = scratch + "/w_rmat.egr";
  const std::vector<unsigned char> orig = slurp(src);
  CHECK(orig.size() > 64, "mutation source missing");
  unsigned long long x = 88172645463325252ull;
  auto rnd = [&x]() {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return x;
  };
  int accepted = 0;
  for (int t = 0; t < 300 && orig.size() > 64; t++) {
    std::vector<unsigned char> buf = orig;
    const int muts = 1 + (int)(rnd() % 4);
    for (int k = 0; k < muts; k++) {
      // past the 8-byte header, so the declared sizes stay small
      const size_t word = 2 + rnd() % ((buf.size() - 8) / 4);
      int32_t val = (int32_t)(rnd() % 2100) - 50;
      std::memcpy(&buf[4 * word], &val, 4);
    }
    if (t % 10 == 0) buf.resize(8 + rnd() % (buf.size() - 8));  // truncations too
    const std::string p = scratch + "/mut.egr";
    FILE* f = std::fopen(p.c_str(), "wb");
    std::fwrite(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    emst_graph_buf b{};
    if (emst_read_egr(p.c_str(), &b) == EMST_OK) {
      const emst_graph v = view(b);
      if (emst_validate(&v) == EMST_OK) {
        if (!b.eweight) emst_synth_weights(&b);
        check_forests(b, "mutated");
        accepted++;
      }
    }
    emst_free_graph(&b);
  }
  for (auto& gg : gens)
    if (gg.rc == EMST_OK) emst_free_graph(&gg.b);
  std::printf("fixtures %d, mutated copies accepted %d, failures %d\n", files, accepted, failures);
  return failures ? 1 : 0;
}
