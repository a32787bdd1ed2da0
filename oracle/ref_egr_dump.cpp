// ref_egr_dump.cpp -- TEST INFRASTRUCTURE ONLY.  Reads an .egr file with the
// reference's own reader (ECL-MST/ECLgraph.h:37-64, included from the
// reference tree at build time, never copied) and prints what it parsed, so
// tests can compare the product reader against it on the same bytes.
// Output: "nodes edges has_weights fnv(nindex) fnv(nlist) fnv(eweight)".
// On a reader error the reference prints its message and exits with -1.
#include <cstdint>
#include <cstdio>
#include "ECLgraph.h"

static uint64_t fnv(const int* p, long n)
{
  uint64_t h = 1469598103934665603ull;
  const unsigned char* b = (const unsigned char*)p;
  for (long i = 0; i < n * 4; i++) { h ^= b[i]; h *= 1099511628211ull; }
  return h;
}

int main(int argc, char** argv)
{
  if (argc != 2) return 2;
  ECLgraph g = readECLgraph(argv[1]);
  printf("%d %d %d %llu %llu %llu\n", g.nodes, g.edges, g.eweight != NULL,
         (unsigned long long)fnv(g.nindex, g.nodes + 1L), (unsigned long long)fnv(g.nlist, g.edges),
         (unsigned long long)(g.eweight ? fnv(g.eweight, g.edges) : 0));
  freeECLgraph(g);
  return 0;
}
