// ref_egr_roundtrip.cpp -- test driver compiled against the reference's own
// .egr reader/writer (library/ECLgraph.h:49-92, included from the reference
// tree by the Makefile; nothing is copied).  Test infrastructure only.
//
//   egr_roundtrip <in.egr> <out.egr>
// reads <in.egr> with readECLgraph, prints "nodes edges has_weights
// checksum", and writes it back with writeECLgraph to <out.egr>.  The tests
// use it to show that files written by libeclgraph.so are read by the
// reference and that files written by the reference are read back by
// libeclgraph.so byte for byte.
#include <cstdint>
#include <cstdio>

#include "ECLgraph.h"

static uint64_t fnv(uint64_t h, const int* p, long n) {
  for (long i = 0; i < n; i++) {
    h ^= (uint32_t)p[i];
    h *= 1099511628211ULL;
  }
  return h;
}

int main(int argc, char* argv[]) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s in.egr out.egr\n", argv[0]);
    return 2;
  }
  ECLgraph g = readECLgraph(argv[1]);
  uint64_t h = 1469598103934665603ULL;
  h = fnv(h, g.nindex, (long)g.nodes + 1);
  h = fnv(h, g.nlist, g.edges);
  if (g.eweight) h = fnv(h, g.eweight, g.edges);
  std::printf("%d %d %d %llu\n", g.nodes, g.edges, g.eweight != NULL ? 1 : 0,
              (unsigned long long)h);
  writeECLgraph(g, argv[2]);
  freeECLgraph(g);
  return 0;
}
