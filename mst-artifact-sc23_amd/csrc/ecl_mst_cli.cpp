// ecl_mst_cli.cpp -- command-line driver with the reference's stdout
// protocol (ECL-MST/ECL_MST.cu:398-455):
//   input: <path>
//   9 lines of "%12.9f" seconds, one per device-resident solve
//   "ERROR: results differ!" (plus a blank line) when the last GPU result
//   does not match the host Kruskal.
// Reader failures print the reference's messages (ECLgraph.h:42-59) and
// exit with -1, as the reference does.  A JSON line with the solver's
// statistics goes to stderr, leaving stdout in the reference's format.
// EMST_FILTER=learned in the environment replaces the reference's threshold
// rule by the trained network (filter_mlp.f32 next to the bin/ folder).
// EMST_VARIANT=<suffixes> runs a rung of the de-optimization ladder, named
// by the suffixes of the reference's variant binaries
// (Deoptimizations/run_all_ECL-MST_deopts.sh: ecl_mst_NAG_NW_NF_...):
// REF (the reference's kernel sequence, no removal), NAG, NW, NF, NIPC, NOD,
// NT, NDD, NEB, joined by '_'.

#include <unistd.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/emst.h"

// The trained threshold network shipped beside the executable's bin/ folder.
static bool load_network(std::vector<float>& p)
{
  char exe[4096];
  const ssize_t n = readlink("/proc/self/exe", exe, sizeof(exe) - 1);
  if (n <= 0) return false;
  exe[n] = 0;
  std::string dir(exe);
  dir = dir.substr(0, dir.rfind('/'));   // .../bin
  dir = dir.substr(0, dir.rfind('/'));   // package folder
  FILE* f = std::fopen((dir + "/filter_mlp.f32").c_str(), "rb");
  if (!f) return false;
  p.resize(EMST_MLP_PARAMS);
  const bool ok = std::fread(p.data(), 4, p.size(), f) == p.size();
  std::fclose(f);
  return ok;
}

// Flags and filter switch of a ladder variant name; false on an unknown suffix.
static bool parse_variant(const char* name, int* flags, bool* no_filter)
{
  static const struct {
    const char* tok;
    int flag;
  } table[] = {{"REF", 0},
               {"NAG", EMST_F_NO_GUARD},
               {"NW", EMST_F_NO_WARP},
               {"NF", 0},
               {"NIPC", EMST_F_NO_IMPLICIT_PC},
               {"NOD", EMST_F_NO_ONE_DIRECTION},
               {"NT", EMST_F_NO_TUPLES},
               {"NDD", EMST_F_NO_DATA_DRIVEN},
               {"NEB", EMST_F_NO_EDGE_BASED}};
  *flags = EMST_F_LADDER;
  *no_filter = false;
  std::string s(name);
  size_t pos = 0;
  while (pos <= s.size()) {
    const size_t end = std::min(s.find('_', pos), s.size());
    const std::string tok = s.substr(pos, end - pos);
    bool found = false;
    for (const auto& t : table)
      if (tok == t.tok) {
        *flags |= t.flag;
        if (tok == "NF") *no_filter = true;
        found = true;
      }
    if (!found) return false;
    pos = end + 1;
  }
  return true;
}

static const char* reader_message(int rc)
{
  switch (rc) {
    case EMST_E_OPEN: return "could not open file";
    case EMST_E_NODES: return "failed to read nodes";
    case EMST_E_EDGES: return "failed to read edges";
    case EMST_E_COUNTS: return "node or edge count too low";
    case EMST_E_NOMEM: return "memory allocation failed";
    case EMST_E_NINDEX: return "failed to read neighbor index list";
    case EMST_E_NLIST: return "failed to read neighbor list";
    case EMST_E_WEIGHTS: return "failed to read edge weights";
    default: return "unknown reader error";
  }
}

int main(int argc, char* argv[])
{
  if (argc != 2) {
    printf("USAGE: %s input_graph\n", argv[0]);
    exit(-1);
  }
  emst_graph_buf gb;
  int rc = emst_read_egr(argv[1], &gb);
  if (rc) {
    if (rc == EMST_E_OPEN)
      fprintf(stderr, "ERROR: could not open file %s\n\n", argv[1]);
    else
      fprintf(stderr, "ERROR: %s\n\n", reader_message(rc));
    exit(-1);
  }
  printf("input: %s\n", argv[1]);
  if (!gb.eweight && emst_synth_weights(&gb) != EMST_OK) {
    fprintf(stderr, "ERROR: memory allocation failed\n\n");
    exit(-1);
  }
  const emst_graph g{gb.nodes, gb.edges, gb.nindex, gb.nlist, gb.eweight};
  const char* mode = std::getenv("EMST_FILTER");
  const bool learned = mode && std::strcmp(mode, "learned") == 0;
  const char* variant = std::getenv("EMST_VARIANT");
  bool no_filter = false;
  emst_config vcfg{};
  if (variant && *variant && (learned || !parse_variant(variant, &vcfg.flags, &no_filter))) {
    fprintf(stderr, "ERROR: bad EMST_VARIANT %s (learned filter not allowed)\n\n", variant);
    exit(-1);
  }
  emst_ctx* ctx = nullptr;
  if (emst_create(0, &ctx) != EMST_OK) {
    fprintf(stderr, "ERROR: there is no HIP capable device\n\n");
    exit(-1);
  }
  if ((rc = emst_load(ctx, &g)) != EMST_OK) {
    fprintf(stderr, "ERROR %d: %s\n", rc, emst_last_error(ctx));
    exit(-1);
  }
  if (learned) {
    std::vector<float> p;
    if (!load_network(p) || emst_set_filter_model(ctx, p.data()) != EMST_OK) {
      fprintf(stderr, "ERROR: could not load filter_mlp.f32\n\n");
      exit(-1);
    }
  }
  if (variant && *variant) {
    if ((rc = emst_set_config(ctx, &vcfg)) != EMST_OK) {
      fprintf(stderr, "ERROR %d: %s\n", rc, emst_last_error(ctx));
      exit(-1);
    }
  }
  const int runs = 9;  // ECL_MST.cu:49
  emst_stats st{};
  for (int i = 0; i < runs; i++) {
    int use_filter = EMST_FILTER_LEARNED;
    int32_t threshold = INT_MAX;
    if (no_filter)
      use_filter = EMST_FILTER_OFF;  // gpuMST<false>(g, INT_MAX), as the NoFilter variants call it
    else if (!learned)
      emst_plain_threshold(&g, &use_filter, &threshold);
    if ((rc = emst_solve(ctx, use_filter, threshold, &st)) != EMST_OK) {
      fprintf(stderr, "ERROR %d: %s\n", rc, emst_last_error(ctx));
      exit(-1);
    }
    printf("%12.9f\n", st.ms / 1000.0);
  }
  std::vector<uint8_t> gpu((size_t)(g.edges ? g.edges : 1)), cpu((size_t)(g.edges ? g.edges : 1));
  if ((rc = emst_fetch(ctx, gpu.data())) != EMST_OK) {
    fprintf(stderr, "ERROR %d: %s\n", rc, emst_last_error(ctx));
    exit(-1);
  }
  emst_cpu_mst(&g, cpu.data());
  uint64_t cw = 0, gw = 0;
  if (emst_verify(&g, cpu.data(), gpu.data(), &cw, &gw) != 0) printf("ERROR: results differ!\n\n");
  fprintf(stderr,
          "{\"nodes\": %d, \"edges\": %d, \"light\": %lld, \"heavy\": %lld, \"light_rounds\": %d, "
          "\"heavy_rounds\": %d, \"tail_rounds\": %d, \"weight\": %llu, \"last_ms\": %.6f}\n",
          g.nodes, g.edges, (long long)st.light, (long long)st.heavy, st.light_rounds, st.heavy_rounds,
          st.tail_rounds, (unsigned long long)gw, st.ms);
  emst_destroy(ctx);
  emst_free_graph(&gb);
  return 0;
}
