// eclmst_cli.cpp -- command-line face of libeclmst.so, `eclmst <graph.egr>`.
//
// Mirrors the reference's main (src/racefree/egr-input/MST/ECL-MST_10-atomic.cu
// :418-474) so its harness can drive it unchanged: one argument, and the
// same stdout lines -- the "ECL-MST" banner, "input: <file>", the device line,
// "Dev+: %12.9f s" (filtered, average degree >= 4) or "Dev-: ...", the host
// Kruskal's "Host: %12.9f s", then "all good" or "ERROR: results differ!".
// scripts/analyze_results.py:83-84 turns "Dev[-+]" into "runtime" and reads
// the time with `time: +(\d+\.\d+) s`.  Errors go to stderr as "ERROR: ..."
// with exit status -1.
//
// Weights: the file's, or the reference's default weights (:428-437) when the
// file has none.  The host check is the reference's (cpuMST :112-154 and
// verify :385-406): Kruskal over the edges sorted by (weight, entry), then
// the two edge sets and their weights compared.

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/eclgraph.h"
#include "../../include/eclmst.h"

static void die(const char* msg) {
  std::fprintf(stderr, "ERROR: %s\n\n", msg);
  std::exit(-1);
}

static int find_root(std::vector<int32_t>& parent, int32_t x) {
  while (parent[x] != x) {
    parent[x] = parent[parent[x]];
    x = parent[x];
  }
  return x;
}

// Host Kruskal: in_mst[j] for the entries j (i -> k, k > i) of the minimum
// spanning forest under (weight, j).
static std::vector<uint8_t> host_mst(int32_t n, const int32_t* nidx, const int32_t* nlist,
                                     const int32_t* w) {
  struct E {
    uint64_t key;
    int32_t a, b;
  };
  std::vector<E> es;
  for (int32_t i = 0; i < n; i++)
    for (int32_t j = nidx[i]; j < nidx[i + 1]; j++)
      if (nlist[j] > i) es.push_back(E{((uint64_t)(uint32_t)w[j] << 32) | (uint32_t)j, i, nlist[j]});
  std::sort(es.begin(), es.end(), [](const E& x, const E& y) { return x.key < y.key; });
  std::vector<int32_t> parent((size_t)n);
  for (int32_t v = 0; v < n; v++) parent[v] = v;
  std::vector<uint8_t> in((size_t)nidx[n], 0);
  int32_t left = n - 1;
  for (const E& e : es) {
    if (left == 0) break;
    const int32_t ra = find_root(parent, e.a), rb = find_root(parent, e.b);
    if (ra == rb) continue;
    in[(uint32_t)e.key] = 1;
    parent[std::max(ra, rb)] = std::min(ra, rb);
    left--;
  }
  return in;
}

int main(int argc, char* argv[]) {
  std::printf("ECL-MST (MI355X race-free build; %s)\n\n", eclmst_version());
  std::fflush(stdout);
  if (argc != 2) {
    std::fprintf(stderr, "USAGE: %s input_graph\n\n", argv[0]);
    std::exit(-1);
  }
  eclgraph_csr* g = nullptr;
  if (eclgraph_read(argv[1], &g) != 0) die(eclgraph_last_error());
  std::printf("input: %s\n", argv[1]);
  const int32_t n = eclgraph_nodes(g);
  const int32_t m = (int32_t)eclgraph_edges(g);
  const int32_t* nidx = eclgraph_nindex(g);
  const int32_t* nlist = eclgraph_nlist(g);
  // the kernels index nlist and the per-vertex arrays with nindex and the
  // entries: a decreasing nindex or an entry outside [0, n) stops here, not
  // on the GPU (the reference does not check them)
  if (eclgraph_check(n, m, nidx, nlist, 0, nullptr) != 0) die(eclgraph_last_error());
  const int32_t* fw = eclgraph_eweight(g);
  std::vector<int32_t> dw;
  if (!fw) {  // assign weights if needed
    dw.resize((size_t)std::max(m, 1));
    if (eclgraph_mst_weights(n, nidx, nlist, dw.data()) != 0) die(eclgraph_last_error());
    fw = dw.data();
  }
  std::printf("GPU: %s\n\n", eclmst_version());
  std::fflush(stdout);

  std::vector<uint8_t> dev((size_t)std::max(m, 1), 0);
  double secs = 0;
  if (eclmst_compute_timed(n, m, nidx, nlist, fw, dev.data(), &secs) != ECLMST_OK)
    die(eclmst_last_error());
  const bool filtered = m / n >= 4;
  std::printf("Dev%c: %12.9f s\n", filtered ? '+' : '-', secs);

  const auto t0 = std::chrono::steady_clock::now();
  const std::vector<uint8_t> host = host_mst(n, nidx, nlist, fw);
  const double hsecs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("Host: %12.9f s\n", hsecs);

  uint64_t wd = 0, wh = 0;
  int64_t only_dev = 0, only_host = 0;
  for (int32_t j = 0; j < m; j++) {
    if (dev[j] && !host[j]) only_dev++;
    if (host[j] && !dev[j]) only_host++;
    if (dev[j]) wd += (uint32_t)fw[j];
    if (host[j]) wh += (uint32_t)fw[j];
  }
  if (wd != wh || only_dev != 0 || only_host != 0) std::printf("ERROR: results differ!\n\n");
  else std::printf("all good\n\n");
  eclgraph_free(g);
  return 0;
}
