// eclcc_cli.cpp -- command-line face of the library, `eclcc <graph.egr>`.
//
// Mirrors the reference's main (src/racefree/egr-input/CC/ECL-CC-atomic.cu
// :338-400) so its harness can drive it unchanged: same single argument,
// same stdout lines (input graph / average, minimum, maximum degree /
// "compute time: %.9f s" / Mnodes/s / Medges/s / number of connected
// components / "verification passed"), errors as "ERROR: ..." on stderr with
// exit status -1.  The reference's harness parses `time: +(\d+\.\d+) s`
// (scripts/analyze_results.py:62).
//
// Verification restates the reference's host checks (:365-395): adjacent
// vertices share a label, no negative labels, one walk per component finds a
// single ID, distinct labels = number of walks.  The walk is iterative, so no
// `ulimit -s unlimited` is needed (all_tests.sh:20).

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/eclcc.h"
#include "../../include/eclgraph.h"

static void die(const char* msg) {
  std::fprintf(stderr, "ERROR: %s\n\n", msg);
  std::exit(-1);
}

int main(int argc, char* argv[]) {
  std::printf("ECL-CC (MI355X race-free build; %s)\n", eclcc_version());
  // optional --check before the file: test the CSR preconditions first
  const bool check = argc == 3 && std::strcmp(argv[1], "--check") == 0;
  if (argc != 2 && !check) {
    std::fprintf(stderr, "USAGE: %s [--check] input_file_name\n\n", argv[0]);
    std::exit(-1);
  }
  const char* path = argv[argc - 1];
  eclgraph_csr* g = nullptr;
  if (eclgraph_read(path, &g) != 0) die(eclgraph_last_error());
  const int32_t n = eclgraph_nodes(g);
  const int32_t m = (int32_t)eclgraph_edges(g);
  const int32_t* nidx = eclgraph_nindex(g);
  const int32_t* nlist = eclgraph_nlist(g);
  // nindex must start at 0, never decrease and end at the entry count, and
  // every entry must name a vertex: the kernels index nlist and the labels
  // with these values, so a file that breaks them never reaches the GPU
  // (the reference does not check them).  --check adds the symmetry test.
  if (eclgraph_check(n, m, nidx, nlist, check ? ECLGRAPH_CHECK_SYMMETRY : 0, nullptr) != 0)
    die(eclgraph_last_error());

  std::printf("input graph: %d nodes and %d edges (%s)\n", n, m, path);
  std::printf("average degree: %.2f edges per node\n", 1.0 * m / n);
  int mindeg = n, maxdeg = 0;
  for (int32_t v = 0; v < n; v++) {
    const int deg = nidx[v + 1] - nidx[v];
    if (deg < mindeg) mindeg = deg;
    if (deg > maxdeg) maxdeg = deg;
  }
  std::printf("minimum degree: %d edges\n", mindeg);
  std::printf("maximum degree: %d edges\n", maxdeg);

  std::vector<int32_t> nstat((size_t)n);
  double secs = 0;
  if (eclcc_compute_timed(n, m, nidx, nlist, nstat.data(), nullptr, &secs) != ECLCC_OK)
    die(eclcc_last_error());
  std::printf("compute time: %.9f s\n", secs);
  std::printf("throughput: %.3f Mnodes/s\n", n * 0.000001 / secs);
  std::printf("throughput: %.3f Medges/s\n", m * 0.000001 / secs);

  std::vector<unsigned char> seen((size_t)n, 0);
  long ncomp = 0;
  for (int32_t v = 0; v < n; v++) {
    const int32_t l = nstat[v];
    if (l < 0) die("found negative component number");
    if (l >= n) die("found component number that is not a vertex ID");
    if (!seen[l]) {
      seen[l] = 1;
      ncomp++;
    }
  }
  std::printf("number of connected components: %ld\n", ncomp);

  for (int32_t v = 0; v < n; v++)
    for (int32_t i = nidx[v]; i < nidx[v + 1]; i++)
      if (nstat[nlist[i]] != nstat[v]) die("found adjacent nodes in different components");
  std::vector<int32_t> st(nstat), stack;
  std::vector<unsigned char> seen2((size_t)n, 0);
  long count = 0, s2 = 0;
  for (int32_t v = 0; v < n; v++) {
    if (st[v] < 0) continue;
    const int32_t id = st[v];
    count++;
    if (!seen2[id]) {
      seen2[id] = 1;
      s2++;
    }
    stack.assign(1, v);
    st[v] = -1;
    while (!stack.empty()) {
      const int32_t x = stack.back();
      stack.pop_back();
      for (int32_t i = nidx[x]; i < nidx[x + 1]; i++) {
        const int32_t y = nlist[i];
        if (st[y] >= 0) {
          if (st[y] != id) die("found incorrect ID value");
          st[y] = -1;
          stack.push_back(y);
        }
      }
    }
  }
  if (ncomp != s2) die("number of components do not match");
  if (ncomp != count) die("component IDs are not unique");
  std::printf("verification passed\n\n");
  eclgraph_free(g);
  return 0;
}
