/* eclcc_caller.c -- the in-process drop-in exercised from plain C, the way
 * the reference's main would call it once computeCC
 * (src/racefree/egr-input/CC/ECL-CC-atomic.cu:275-325, called at :363) is
 * replaced by the library: read the .egr file, allocate the label array,
 * call eclcc_compute(g.nodes, g.edges, g.nindex, g.nlist, nodestatus), and
 * on failure print "ERROR: <message>" and exit(-1).  Test infrastructure:
 * it writes the labels as raw int32 to <labels.bin> for the test to check.
 *
 *   eclcc_caller <graph.egr> <labels.bin>
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "eclcc.h"

static void die(const char* msg) {
  fprintf(stderr, "ERROR: %s\n\n", msg);
  exit(-1);
}

int main(int argc, char* argv[]) {
  if (argc != 3) {
    fprintf(stderr, "USAGE: %s input_file_name labels_file_name\n\n", argv[0]);
    exit(-1);
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) die("could not open input file");
  int32_t nodes = 0, edges = 0;
  if (fread(&nodes, 4, 1, f) != 1 || fread(&edges, 4, 1, f) != 1) die("failed to read header");
  if (nodes < 1 || edges < 0) die("node or edge count too low");
  int32_t* nindex = (int32_t*)malloc(4 * ((size_t)nodes + 1));
  int32_t* nlist = (int32_t*)malloc(4 * ((size_t)edges + 1));
  int32_t* nodestatus = (int32_t*)malloc(4 * (size_t)nodes);
  if (!nindex || !nlist || !nodestatus) die("memory allocation failed");
  if (fread(nindex, 4, (size_t)nodes + 1, f) != (size_t)nodes + 1) die("failed to read neighbor index list");
  if (fread(nlist, 4, (size_t)edges, f) != (size_t)edges) die("failed to read neighbor list");
  fclose(f);

  /* the one call that replaces computeCC */
  if (eclcc_compute(nodes, edges, nindex, nlist, nodestatus) != ECLCC_OK) die(eclcc_last_error());

  FILE* o = fopen(argv[2], "wb");
  if (!o || fwrite(nodestatus, 4, (size_t)nodes, o) != (size_t)nodes) die("could not write labels");
  fclose(o);
  printf("eclcc_compute: %d nodes, %d edges, labels written to %s\n", nodes, edges, argv[2]);
  free(nindex);
  free(nlist);
  free(nodestatus);
  return 0;
}
