// edgelist_mst_cli.cpp -- the second related-work baseline (edge-list
// Boruvka with label propagation, the Jucele GPU algorithm) behind the same
// command line and stdout protocol as bin/ecl_mst (ECL-MST/ECL_MST.cu:
// 398-455): "input: <path>", then 9 lines of "%12.9f" seconds, then "ERROR:
// results differ!" when the forest differs from the host Kruskal's edge by
// edge or in total weight (the verify of ECL_MST.cu:365-386).  Ties go to the
// smaller CSR position here too, so the edge sets are comparable.

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/emst.h"

int main(int argc, char* argv[])
{
  if (argc != 2) {
    printf("USAGE: %s input_graph\n", argv[0]);
    exit(-1);
  }
  emst_graph_buf gb;
  int rc = emst_read_egr(argv[1], &gb);
  if (rc) {
    fprintf(stderr, "ERROR: could not read %s (code %d)\n\n", argv[1], rc);
    exit(-1);
  }
  printf("input: %s\n", argv[1]);
  if (!gb.eweight && emst_synth_weights(&gb) != EMST_OK) {
    fprintf(stderr, "ERROR: memory allocation failed\n\n");
    exit(-1);
  }
  const emst_graph g{gb.nodes, gb.edges, gb.nindex, gb.nlist, gb.eweight};
  emst_ctx* ctx = nullptr;
  if (emst_create(0, &ctx) != EMST_OK) {
    fprintf(stderr, "ERROR: there is no HIP capable device\n\n");
    exit(-1);
  }
  std::vector<uint8_t> gpu((size_t)(g.edges ? g.edges : 1)), cpu((size_t)(g.edges ? g.edges : 1));
  emst_edgelist_stats st{};
  for (int i = 0; i < 9; i++) {
    if ((rc = emst_edgelist_mst(ctx, &g, gpu.data(), &st)) != EMST_OK) {
      fprintf(stderr, "ERROR %d: %s\n", rc, emst_last_error(ctx));
      exit(-1);
    }
    printf("%12.9f\n", st.ms / 1000.0);
  }
  emst_cpu_mst(&g, cpu.data());
  uint64_t cw = 0, gw = 0;
  if (emst_verify(&g, cpu.data(), gpu.data(), &cw, &gw) != 0) printf("ERROR: results differ!\n\n");
  fprintf(stderr, "{\"nodes\": %d, \"edges\": %d, \"iterations\": %lld, \"label_passes\": %lld, \"weight\": %llu, "
          "\"mst_edges\": %lld, \"last_ms\": %.6f}\n",
          g.nodes, g.edges, (long long)st.iterations, (long long)st.label_passes, (unsigned long long)st.weight,
          (long long)st.mst_edges, st.ms);
  emst_destroy(ctx);
  emst_free_graph(&gb);
  return 0;
}
