// mm2ecl_cli.cpp -- `mm2ecl input.mtx output.egr`, the command-line face of
// eclgraph_mm2ecl (include/eclgraph.h).  It stands in for the reference's
// converter (scripts/mm2ecl_real.cpp, run by download_inputs.sh:30,39) and
// writes the same .egr bytes.  Like it, it prints the input name, the node
// and entry counts and the kind of weights, one "<value>\t#<key>" line each
// (:81-83, :94, :117, :146), and reports a failure as "ERROR: ..." on
// stderr with exit status -1.

#include <cstdio>
#include <cstdlib>

#include "../../include/eclgraph.h"

static void die(const char* msg) {
  std::fprintf(stderr, "ERROR: %s\n\n", msg);
  std::exit(-1);
}

int main(int argc, char* argv[]) {
  std::printf("MatrixMarket to ECL Graph Converter (MI355X build, libeclgraph)\n");
  if (argc != 3) {
    std::fprintf(stderr, "USAGE: %s input_file_name output_file_name\n\n", argv[0]);
    std::exit(-1);
  }
  eclgraph_csr* g = nullptr;
  int kind = ECLGRAPH_WEIGHTS_NONE;
  if (eclgraph_mm2ecl(argv[1], &g, &kind) != ECLGRAPH_OK) die(eclgraph_last_error());
  std::printf("%s\t#name\n", argv[1]);
  std::printf("%d\t#nodes\n", eclgraph_nodes(g));
  std::printf("%lld\t#edges\n", (long long)eclgraph_edges(g));
  std::printf("%s\t#weights\n", kind == ECLGRAPH_WEIGHTS_INT    ? "int"
                                : kind == ECLGRAPH_WEIGHTS_REAL ? "float"
                                                                : "no");
  if (eclgraph_write(g, argv[2]) != ECLGRAPH_OK) die(eclgraph_last_error());
  eclgraph_free(g);
  return 0;
}
