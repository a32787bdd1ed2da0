// ecl_convert -- converts an ECLgraph .egr file into the input format of one
// of the reference's external baselines (the job of its per-baseline
// */inputs/ecl2*_genweights.cpp programs).  A weightless input first gets
// the reference's synthetic weights (ECL_MST.cu:410-419).
#include <cstdio>
#include <string>

#include "../../include/emst.h"

int main(int argc, char* argv[])
{
  if (argc != 4) {
    std::fprintf(stderr, "USAGE: %s <input.egr> <mtx|pbbs|lsg|uminho|jucele|egr> <output>\n", argv[0]);
    return 1;
  }
  emst_graph_buf g{};
  int rc = emst_read_egr(argv[1], &g);
  if (rc != EMST_OK) {
    std::fprintf(stderr, "ERROR: could not read %s (%d)\n", argv[1], rc);
    return 1;
  }
  std::printf("%s\t#name\n%d\t#nodes\n%d\t#edges\n", argv[1], g.nodes, g.edges);
  std::printf("%s\t#weights\n", g.eweight ? "yes" : "no");
  if (!g.eweight) {
    rc = emst_synth_weights(&g);
    if (rc != EMST_OK) {
      std::fprintf(stderr, "ERROR: could not add weights (%d)\n", rc);
      emst_free_graph(&g);
      return 1;
    }
    std::printf("added missing weights using ECL-MST method\n");
  }
  const emst_graph v{g.nodes, g.edges, g.nindex, g.nlist, g.eweight};
  const std::string fmt = argv[2];
  if (fmt == "mtx") rc = emst_write_mtx(argv[3], &v);
  else if (fmt == "pbbs") rc = emst_write_pbbs(argv[3], &v);
  else if (fmt == "lsg") rc = emst_write_galois(argv[3], &v);
  else if (fmt == "uminho") rc = emst_write_uminho(argv[3], &v);
  else if (fmt == "jucele") rc = emst_write_jucele(argv[3], &v);
  else if (fmt == "egr") rc = emst_write_egr(argv[3], &v);
  else {
    std::fprintf(stderr, "ERROR: unknown format %s\n", fmt.c_str());
    emst_free_graph(&g);
    return 1;
  }
  emst_free_graph(&g);
  if (rc != EMST_OK) {
    std::fprintf(stderr, "ERROR: could not write %s (%d)\n", argv[3], rc);
    return 1;
  }
  return 0;
}
