// ecl_gen -- writes the seeded synthetic inputs as ECLgraph .egr files
// (ECLgraph.h:66-80 layout), so the benchmark inputs can be handed to other
// programs: the reference's CLI protocol works on files.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/emst.h"
#include "../../include/emst_gen.h"

static void usage(const char* argv0)
{
  std::fprintf(stderr,
               "USAGE: %s grid <rows> <cols> <seed> <wlo> <whi> <out.egr>\n"
               "       %s rmat <scale> <edge_factor> <a> <b> <c> <seed> <wlo> <whi> <out.egr>\n"
               "       %s rgg <n> <mean_degree> <seed> <wmax> <out.egr>\n"
               "       %s er <n> <mean_degree> <seed> <wlo> <whi> <out.egr>\n"
               "       %s config <grid512|rmat24|rgg24|rmat26> <out.egr>\n",
               argv0, argv0, argv0, argv0, argv0);
}

int main(int argc, char* argv[])
{
  if (argc < 3) {
    usage(argv[0]);
    return 1;
  }
  const std::string kind = argv[1];
  emst_graph_buf g{};
  int rc = EMST_E_ARG;
  const char* out = argv[argc - 1];
  if (kind == "grid" && argc == 8) {
    rc = emst_gen_grid(std::atoi(argv[2]), std::atoi(argv[3]), std::strtoull(argv[4], nullptr, 10), std::atoi(argv[5]),
                       std::atoi(argv[6]), &g);
  } else if (kind == "rmat" && argc == 11) {
    rc = emst_gen_rmat(std::atoi(argv[2]), std::atoi(argv[3]), std::atof(argv[4]), std::atof(argv[5]),
                       std::atof(argv[6]), std::strtoull(argv[7], nullptr, 10), std::atoi(argv[8]), std::atoi(argv[9]), &g);
  } else if (kind == "rgg" && argc == 7) {
    rc = emst_gen_rgg(std::atoi(argv[2]), std::atof(argv[3]), std::strtoull(argv[4], nullptr, 10), std::atoi(argv[5]), &g);
  } else if (kind == "er" && argc == 8) {
    rc = emst_gen_er(std::atoi(argv[2]), std::atof(argv[3]), std::strtoull(argv[4], nullptr, 10), std::atoi(argv[5]),
                     std::atoi(argv[6]), &g);
  } else if (kind == "config" && argc == 4) {
    const std::string c = argv[2];
    if (c == "grid512") rc = emst_gen_grid(512, 512, 1, 1, 1 << 20, &g);
    else if (c == "rmat24") rc = emst_gen_rmat(24, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24, &g);
    else if (c == "rgg24") rc = emst_gen_rgg(1 << 24, 6.0, 3, 1 << 24, &g);
    else if (c == "rmat26") rc = emst_gen_rmat(26, 16, 0.57, 0.19, 0.19, 5, 1, 65, &g);
  } else {
    usage(argv[0]);
    return 1;
  }
  if (rc != EMST_OK) {
    std::fprintf(stderr, "ERROR: generation failed (%d)\n", rc);
    return 1;
  }
  const emst_graph v{g.nodes, g.edges, g.nindex, g.nlist, g.eweight};
  rc = emst_write_egr(out, &v);
  if (rc != EMST_OK) {
    std::fprintf(stderr, "ERROR: could not write %s (%d)\n", out, rc);
    emst_free_graph(&g);
    return 1;
  }
  std::printf("%s\t#name\n%d\t#nodes\n%d\t#edges\n", out, g.nodes, g.edges);
  emst_free_graph(&g);
  return 0;
}
