// eclgraph.cpp -- host library behind include/eclgraph.h: ECLgraph .egr I/O,
// edge-list -> CSR construction and the seeded synthetic generators
// (Erdos-Renyi G(n,M), R-MAT, 3-D 7-point grid).
//
// Format follows library/ECLgraph.h:41-92 of the reference; the generators
// have no counterpart there (the reference downloads inputs,
// download_inputs.sh:11-27).
//
// Determinism: every random number is splitmix64 output number `counter` of a
// stream started at `seed` (a counter-based generator), so results do not
// depend on how work is split over threads.  Adjacency lists are sorted.

#include "../../include/eclgraph.h"

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

struct eclgraph_csr {
  int32_t nodes = 0;
  int64_t edges = 0;
  std::unique_ptr<int32_t[]> nindex;
  std::unique_ptr<int32_t[]> nlist;
  std::unique_ptr<int32_t[]> eweight;  // null when absent
};

namespace {

thread_local std::string g_err;
std::atomic<int> g_threads{0};
thread_local int t_threads = 0;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Host exceptions never cross the C ABI: every entry point that can meet
// one (a host allocation that fails, a thread that cannot start) turns it
// into a status here.
int host_exception() {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    return fail(ECLGRAPH_ERR_ALLOC, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(ECLGRAPH_ERR_ALLOC, std::string("host resources: ") + e.what());
  } catch (...) {
    return fail(ECLGRAPH_ERR_ALLOC, "host error");
  }
}

int worker_count() {
  if (t_threads > 0) return t_threads;
  int t = g_threads.load();
  if (t > 0) return t;
  if (const char* e = std::getenv("ECLGRAPH_THREADS")) {
    int v = std::atoi(e);
    if (v > 0) return v;
  }
  unsigned hc = std::thread::hardware_concurrency();
  if (hc == 0) hc = 1;
  return (int)std::min(hc, 16u);
}

// Runs fn(lo, hi) over [0, n) split into dynamically claimed chunks.
template <class Fn>
void parallel_for(int64_t n, int64_t chunk, Fn&& fn) {
  if (n <= 0) return;
  int nt = worker_count();
  if (nt <= 1 || n <= chunk) {
    fn((int64_t)0, n);
    return;
  }
  std::atomic<int64_t> next{0};
  auto body = [&]() {
    for (;;) {
      int64_t lo = next.fetch_add(chunk);
      if (lo >= n) break;
      fn(lo, std::min(n, lo + chunk));
    }
  };
  std::vector<std::thread> pool;
  int use = (int)std::min<int64_t>(nt, (n + chunk - 1) / chunk);
  pool.reserve(use - 1);
  for (int i = 1; i < use; i++) pool.emplace_back(body);
  body();
  for (auto& t : pool) t.join();
}

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Output number `ctr` (0-based) of splitmix64 started at state `seed`.
inline uint64_t rand64(uint64_t seed, uint64_t ctr) {
  return mix64(seed + (ctr + 1) * 0x9E3779B97F4A7C15ULL);
}

// Maps 32 random bits onto [0, bound) by a multiply-high.
inline uint32_t bounded32(uint32_t r, uint32_t bound) {
  return (uint32_t)(((uint64_t)r * bound) >> 32);
}

template <class T>
std::unique_ptr<T[]> alloc_array(int64_t n) {
  return std::unique_ptr<T[]>(new (std::nothrow) T[(size_t)std::max<int64_t>(n, 1)]);
}

// Fisher-Yates shuffle of the identity on stream (seed, i).
std::unique_ptr<int32_t[]> random_permutation(int32_t n, uint64_t seed) {
  auto p = alloc_array<int32_t>(n);
  if (!p) return p;
  for (int32_t i = 0; i < n; i++) p[i] = i;
  for (int32_t i = n - 1; i > 0; i--) {
    uint32_t j = bounded32((uint32_t)(rand64(seed, (uint64_t)i) >> 32), (uint32_t)i + 1);
    std::swap(p[i], p[j]);
  }
  return p;
}

// Edge source for the CSR builder: pair(i) -> (u, v).
struct PairArrays {
  const int32_t* src;
  const int32_t* dst;
  void get(int64_t i, int32_t& u, int32_t& v) const { u = src[i]; v = dst[i]; }
};

// Degree counters: relaxed atomics when several threads build, plain
// integers when one does (the batch generator builds many graphs at once,
// one thread each).
struct Counts {
  bool atomic;
  std::unique_ptr<std::atomic<int32_t>[]> a;
  std::unique_ptr<int32_t[]> p;
  explicit Counts(int32_t n, bool at) : atomic(at) {
    if (at) a.reset(new (std::nothrow) std::atomic<int32_t>[(size_t)n]);
    else p.reset(new (std::nothrow) int32_t[(size_t)n]);
  }
  bool ok() const { return atomic ? (bool)a : (bool)p; }
  void set(int64_t x, int32_t v) {
    if (atomic) a[x].store(v, std::memory_order_relaxed);
    else p[x] = v;
  }
  int32_t get(int64_t x) const { return atomic ? a[x].load(std::memory_order_relaxed) : p[x]; }
  int32_t inc(int64_t x) { return atomic ? a[x].fetch_add(1, std::memory_order_relaxed) : p[x]++; }
};

// Builds a CSR from npairs directed pairs. flags as in eclgraph_from_edges.
// Counting sort by source, then per-list sort (+ unique) and compaction.
template <class Src>
int build_csr(int32_t n, int64_t npairs, const Src& pairs, int flags, eclgraph_csr** out) {
  const bool sym = flags & ECLGRAPH_SYMMETRIZE;
  const bool noself = flags & ECLGRAPH_DROP_SELF;
  const bool dedupe = flags & ECLGRAPH_DEDUPE;

  Counts cnt(n, worker_count() > 1);
  if (!cnt.ok()) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory (degree counts)");
  parallel_for(n, 1 << 16, [&](int64_t lo, int64_t hi) {
    for (int64_t x = lo; x < hi; x++) cnt.set(x, 0);
  });
  std::atomic<int> bad{0};
  parallel_for(npairs, 1 << 16, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; i++) {
      int32_t u, v;
      pairs.get(i, u, v);
      if ((uint32_t)u >= (uint32_t)n || (uint32_t)v >= (uint32_t)n) { bad.store(1); continue; }
      if (noself && u == v) continue;
      cnt.inc(u);
      if (sym && u != v) cnt.inc(v);
    }
  });
  if (bad.load()) return fail(ECLGRAPH_ERR_ARG, "edge endpoint out of range [0, nodes)");

  auto off = alloc_array<int64_t>((int64_t)n + 1);
  if (!off) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory (offsets)");
  off[0] = 0;
  for (int32_t x = 0; x < n; x++) off[x + 1] = off[x] + cnt.get(x);
  const int64_t total = off[n];

  auto tmp = alloc_array<int32_t>(total);
  if (!tmp) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory (edge buffer)");
  parallel_for(n, 1 << 16, [&](int64_t lo, int64_t hi) {
    for (int64_t x = lo; x < hi; x++) cnt.set(x, 0);
  });
  parallel_for(npairs, 1 << 16, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; i++) {
      int32_t u, v;
      pairs.get(i, u, v);
      if (noself && u == v) continue;
      tmp[off[u] + cnt.inc(u)] = v;
      if (sym && u != v) tmp[off[v] + cnt.inc(v)] = u;
    }
  });

  // sort each list, optionally unique; the new length goes to cnt[x]
  parallel_for(n, 1 << 12, [&](int64_t lo, int64_t hi) {
    for (int64_t x = lo; x < hi; x++) {
      int32_t* b = tmp.get() + off[x];
      int32_t* e = tmp.get() + off[x + 1];
      std::sort(b, e);
      int64_t len = dedupe ? (std::unique(b, e) - b) : (e - b);
      cnt.set(x, (int32_t)len);
    }
  });

  std::unique_ptr<eclgraph_csr> g(new (std::nothrow) eclgraph_csr);
  if (!g) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory");
  g->nodes = n;
  g->nindex = alloc_array<int32_t>((int64_t)n + 1);
  if (!g->nindex) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory (nindex)");
  int64_t acc = 0;
  for (int32_t x = 0; x < n; x++) {
    if (acc > INT32_MAX) return fail(ECLGRAPH_ERR_RANGE, "edge count exceeds int32 CSR format");
    g->nindex[x] = (int32_t)acc;
    acc += cnt.get(x);
  }
  if (acc > INT32_MAX) return fail(ECLGRAPH_ERR_RANGE, "edge count exceeds int32 CSR format");
  g->nindex[n] = (int32_t)acc;
  g->edges = acc;
  if (acc == total) {
    g->nlist = std::move(tmp);
  } else {
    g->nlist = alloc_array<int32_t>(acc);
    if (!g->nlist) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory (nlist)");
    int32_t* dstp = g->nlist.get();
    const int32_t* srcp = tmp.get();
    const int32_t* ni = g->nindex.get();
    parallel_for(n, 1 << 14, [&](int64_t lo, int64_t hi) {
      for (int64_t x = lo; x < hi; x++)
        std::memcpy(dstp + ni[x], srcp + off[x], sizeof(int32_t) * (size_t)(ni[x + 1] - ni[x]));
    });
  }
  *out = g.release();
  return ECLGRAPH_OK;
}

bool read_exact(FILE* f, void* p, size_t elem, size_t cnt) {
  return std::fread(p, elem, cnt, f) == cnt;
}

}  // namespace

extern "C" {

const char* eclgraph_last_error(void) { return g_err.c_str(); }

void eclgraph_set_threads(int nthreads) {
  t_threads = nthreads > 0 ? nthreads : 0;
}

uint64_t eclgraph_rand64(uint64_t seed, uint64_t counter) { return rand64(seed, counter); }

int32_t eclgraph_nodes(const eclgraph_csr* g) { return g ? g->nodes : 0; }
int64_t eclgraph_edges(const eclgraph_csr* g) { return g ? g->edges : 0; }
const int32_t* eclgraph_nindex(const eclgraph_csr* g) { return g ? g->nindex.get() : nullptr; }
const int32_t* eclgraph_nlist(const eclgraph_csr* g) { return g ? g->nlist.get() : nullptr; }
const int32_t* eclgraph_eweight(const eclgraph_csr* g) { return g ? g->eweight.get() : nullptr; }

void eclgraph_free(eclgraph_csr* g) { delete g; }

int eclgraph_read(const char* path, eclgraph_csr** out) try {
  if (!path || !out) return fail(ECLGRAPH_ERR_ARG, "null argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(ECLGRAPH_ERR_IO, std::string("could not open file ") + path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
  int32_t hdr[2];
  if (!read_exact(f, &hdr[0], 4, 1)) return fail(ECLGRAPH_ERR_IO, "failed to read nodes");
  if (!read_exact(f, &hdr[1], 4, 1)) return fail(ECLGRAPH_ERR_IO, "failed to read edges");
  if (hdr[0] < 1 || hdr[1] < 0) return fail(ECLGRAPH_ERR_FORMAT, "node or edge count too low");
  std::unique_ptr<eclgraph_csr> g(new (std::nothrow) eclgraph_csr);
  if (!g) return fail(ECLGRAPH_ERR_ALLOC, "memory allocation failed");
  g->nodes = hdr[0];
  g->edges = hdr[1];
  g->nindex = alloc_array<int32_t>((int64_t)g->nodes + 1);
  g->nlist = alloc_array<int32_t>(g->edges);
  if (!g->nindex || !g->nlist) return fail(ECLGRAPH_ERR_ALLOC, "memory allocation failed");
  if (!read_exact(f, g->nindex.get(), 4, (size_t)g->nodes + 1))
    return fail(ECLGRAPH_ERR_IO, "failed to read neighbor index list");
  if (!read_exact(f, g->nlist.get(), 4, (size_t)g->edges))
    return fail(ECLGRAPH_ERR_IO, "failed to read neighbor list");
  // optional weights: present iff at least one more int32 follows
  if (g->edges > 0) {
    auto w = alloc_array<int32_t>(g->edges);
    if (!w) return fail(ECLGRAPH_ERR_ALLOC, "memory allocation failed");
    size_t got = std::fread(w.get(), 4, (size_t)g->edges, f);
    if (got != 0) {
      if (got != (size_t)g->edges) return fail(ECLGRAPH_ERR_IO, "failed to read edge weights");
      g->eweight = std::move(w);
    }
  }
  if (g->nindex[0] != 0 || g->nindex[g->nodes] != g->edges)
    return fail(ECLGRAPH_ERR_FORMAT, "neighbor index list inconsistent with edge count");
  *out = g.release();
  return ECLGRAPH_OK;
} catch (...) {
  return host_exception();
}

int eclgraph_write(const eclgraph_csr* g, const char* path) try {
  if (!g || !path) return fail(ECLGRAPH_ERR_ARG, "null argument");
  if (g->nodes < 1 || g->edges < 0) return fail(ECLGRAPH_ERR_ARG, "node or edge count too low");
  if (g->edges > INT32_MAX) return fail(ECLGRAPH_ERR_RANGE, "edge count exceeds int32 format");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(ECLGRAPH_ERR_IO, std::string("could not open file ") + path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
  int32_t hdr[2] = {g->nodes, (int32_t)g->edges};
  if (std::fwrite(&hdr[0], 4, 1, f) != 1) return fail(ECLGRAPH_ERR_IO, "failed to write nodes");
  if (std::fwrite(&hdr[1], 4, 1, f) != 1) return fail(ECLGRAPH_ERR_IO, "failed to write edges");
  if (std::fwrite(g->nindex.get(), 4, (size_t)g->nodes + 1, f) != (size_t)g->nodes + 1)
    return fail(ECLGRAPH_ERR_IO, "failed to write neighbor index list");
  if (std::fwrite(g->nlist.get(), 4, (size_t)g->edges, f) != (size_t)g->edges)
    return fail(ECLGRAPH_ERR_IO, "failed to write neighbor list");
  if (g->eweight && std::fwrite(g->eweight.get(), 4, (size_t)g->edges, f) != (size_t)g->edges)
    return fail(ECLGRAPH_ERR_IO, "failed to write edge weights");
  if (std::fflush(f) != 0) return fail(ECLGRAPH_ERR_IO, "failed to flush output file");
  return ECLGRAPH_OK;
} catch (...) {
  return host_exception();
}

int eclgraph_mst_weights(int32_t nodes, const int32_t* nindex, const int32_t* nlist,
                         int32_t* eweight) try {
  if (nodes < 1 || !nindex || (nindex[nodes] > 0 && (!nlist || !eweight)))
    return fail(ECLGRAPH_ERR_ARG, "bad argument");
  parallel_for(nodes, 1 << 14, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; i++) {
      for (int32_t j = nindex[i]; j < nindex[i + 1]; j++) {
        // int product of the reference, formed without signed overflow
        const int32_t prod = (int32_t)((uint32_t)i * (uint32_t)nlist[j]);
        int32_t w = 1 + prod % nodes;
        if (w < 0) w = -w;
        eweight[j] = w;
      }
    }
  });
  return ECLGRAPH_OK;
} catch (...) {
  return host_exception();
}

int eclgraph_check(int32_t nodes, int64_t edges, const int32_t* nindex, const int32_t* nlist,
                   int flags, int64_t* where) try {
  if (where) *where = -1;
  if (!nindex || (edges > 0 && !nlist)) return fail(ECLGRAPH_ERR_ARG, "null argument");
  if (nodes < 1 || edges < 0) return fail(ECLGRAPH_ERR_ARG, "node or edge count too low");
  auto bad = [&](int64_t at, const char* msg) {
    if (where) *where = at;
    return fail(ECLGRAPH_ERR_FORMAT, msg);
  };
  if (nindex[0] != 0) return bad(0, "nindex[0] is not 0");
  for (int32_t v = 0; v < nodes; v++)
    if (nindex[v + 1] < nindex[v]) return bad(v + 1, "nindex decreases");
  if (nindex[nodes] != edges) return bad(nodes, "nindex[nodes] differs from the edge count");
  // first offending entry, found in parallel: the minimum position over chunks
  std::atomic<int64_t> first_range{INT64_MAX};
  parallel_for(edges, 1 << 16, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; i++)
      if (nlist[i] < 0 || nlist[i] >= nodes) {
        int64_t cur = first_range.load();
        while (i < cur && !first_range.compare_exchange_weak(cur, i)) {
        }
        break;
      }
  });
  if (first_range.load() != INT64_MAX) return bad(first_range.load(), "neighbour ID out of range");
  if (flags & ECLGRAPH_CHECK_SYMMETRY) {
    // sorted copy of every list, then a binary search for each reverse entry
    auto sorted = alloc_array<int32_t>(edges);
    if (edges && !sorted) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory (sorted lists)");
    parallel_for(nodes, 1 << 12, [&](int64_t lo, int64_t hi) {
      for (int64_t v = lo; v < hi; v++) {
        std::copy(nlist + nindex[v], nlist + nindex[v + 1], sorted.get() + nindex[v]);
        std::sort(sorted.get() + nindex[v], sorted.get() + nindex[v + 1]);
      }
    });
    std::atomic<int64_t> first_asym{INT64_MAX};
    parallel_for(nodes, 1 << 12, [&](int64_t lo, int64_t hi) {
      for (int64_t v = lo; v < hi; v++)
        for (int64_t i = nindex[v]; i < nindex[v + 1]; i++) {
          const int32_t u = nlist[i];
          const int32_t* b = sorted.get() + nindex[u];
          const int32_t* e = sorted.get() + nindex[u + 1];
          if (!std::binary_search(b, e, (int32_t)v)) {
            int64_t cur = first_asym.load();
            while (i < cur && !first_asym.compare_exchange_weak(cur, i)) {
            }
            return;
          }
        }
    });
    if (first_asym.load() != INT64_MAX)
      return bad(first_asym.load(), "entry without its reverse (graph is not symmetric)");
  }
  return ECLGRAPH_OK;
} catch (...) {
  return host_exception();
}

int eclgraph_from_arrays(int32_t nodes, int64_t edges, const int32_t* nindex,
                         const int32_t* nlist, const int32_t* eweight, eclgraph_csr** out) try {
  if (!out || !nindex || (edges > 0 && !nlist)) return fail(ECLGRAPH_ERR_ARG, "null argument");
  if (nodes < 1 || edges < 0) return fail(ECLGRAPH_ERR_ARG, "node or edge count too low");
  if (edges > INT32_MAX) return fail(ECLGRAPH_ERR_RANGE, "edge count exceeds int32 format");
  std::unique_ptr<eclgraph_csr> g(new (std::nothrow) eclgraph_csr);
  if (!g) return fail(ECLGRAPH_ERR_ALLOC, "memory allocation failed");
  g->nodes = nodes;
  g->edges = edges;
  g->nindex = alloc_array<int32_t>((int64_t)nodes + 1);
  g->nlist = alloc_array<int32_t>(edges);
  if (!g->nindex || !g->nlist) return fail(ECLGRAPH_ERR_ALLOC, "memory allocation failed");
  std::memcpy(g->nindex.get(), nindex, sizeof(int32_t) * ((size_t)nodes + 1));
  if (edges) std::memcpy(g->nlist.get(), nlist, sizeof(int32_t) * (size_t)edges);
  if (eweight) {
    g->eweight = alloc_array<int32_t>(edges);
    if (!g->eweight) return fail(ECLGRAPH_ERR_ALLOC, "memory allocation failed");
    if (edges) std::memcpy(g->eweight.get(), eweight, sizeof(int32_t) * (size_t)edges);
  }
  *out = g.release();
  return ECLGRAPH_OK;
} catch (...) {
  return host_exception();
}

int eclgraph_from_edges(int32_t nodes, int64_t nedges, const int32_t* src, const int32_t* dst,
                        int flags, eclgraph_csr** out) try {
  if (!out || (nedges > 0 && (!src || !dst))) return fail(ECLGRAPH_ERR_ARG, "null argument");
  if (nodes < 1 || nedges < 0) return fail(ECLGRAPH_ERR_ARG, "node or edge count too low");
  *out = nullptr;
  return build_csr(nodes, nedges, PairArrays{src, dst}, flags, out);
} catch (...) {
  return host_exception();
}

// ---- MatrixMarket conversion (scripts/mm2ecl_real.cpp:50-184) ----------------

namespace {

// One entry of a coordinate file: 0-based row and column, weight.  Real
// weights are kept as read (double) until the CSR is written, because the
// reference sorts by the value before scaling it (mm2ecl_real.cpp:153-172).
struct MtxEntry {
  int32_t src, dst;
  int64_t iw;  // integer weight
  double fw;   // real weight
};

// Scanning of the entry section the way the reference's fscanf loop reads
// it (:98, :124, :154): whitespace, newlines included, separates numbers
// anywhere; "%d" is an optional sign and at least one decimal digit, "%lf"
// whatever strtod takes.  The loop stops at the first entry that does not
// scan completely, and the entry count is checked afterwards.
struct Scanner {
  const char* p;
  const char* end;
  void skip_ws() {
    while (p < end && std::isspace((unsigned char)*p)) p++;
  }
  bool int_token(int64_t* v) {
    skip_ws();
    const char* q = p;
    bool neg = false;
    if (q < end && (*q == '+' || *q == '-')) neg = *q++ == '-';
    if (q >= end || !std::isdigit((unsigned char)*q)) return false;
    int64_t x = 0;
    while (q < end && std::isdigit((unsigned char)*q)) {
      if (x < (int64_t)1 << 40) x = x * 10 + (*q - '0');  // clamps far past int32
      q++;
    }
    *v = neg ? -x : x;
    p = q;
    return true;
  }
  bool real_token(double* v) {
    skip_ws();
    if (p >= end) return false;
    char* stop = nullptr;
    *v = std::strtod(p, &stop);  // the buffer ends in a NUL, so strtod stops there
    if (stop == p) return false;
    p = stop;
    return true;
  }
};

}  // namespace

int eclgraph_mm2ecl(const char* path, eclgraph_csr** out, int* weight_kind) try {
  if (!path || !out) return fail(ECLGRAPH_ERR_ARG, "null argument");
  *out = nullptr;
  if (weight_kind) *weight_kind = ECLGRAPH_WEIGHTS_NONE;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(ECLGRAPH_ERR_IO, std::string("could not open input file ") + path);
  std::string text;
  {
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    if (std::ferror(f)) return fail(ECLGRAPH_ERR_IO, std::string("could not read ") + path);
  }
  // lines as getline returns them (each with its newline, if it has one)
  size_t pos = 0;
  auto next_line = [&](std::string* line) {
    if (pos >= text.size()) return false;
    const size_t nl = text.find('\n', pos);
    const size_t stop = nl == std::string::npos ? text.size() : nl + 1;
    line->assign(text, pos, stop - pos);
    pos = stop;
    return true;
  };
  auto has = [](const std::string& line, const char* word) {
    return line.find(word) != std::string::npos;
  };
  std::string line;
  // banner (:62-72): at least 30 characters, and these words, case-sensitive
  if (!next_line(&line) || line.size() < 30) return fail(ECLGRAPH_ERR_FORMAT, "could not read first line");
  if (!has(line, "%%MatrixMarket"))
    return fail(ECLGRAPH_ERR_FORMAT, "first line does not contain \"%%MatrixMarket\"");
  if (!has(line, "matrix")) return fail(ECLGRAPH_ERR_FORMAT, "first line does not contain \"matrix\"");
  if (!has(line, "coordinate"))
    return fail(ECLGRAPH_ERR_FORMAT, "first line does not contain \"coordinate\"");
  if (!has(line, "general") && !has(line, "symmetric"))
    return fail(ECLGRAPH_ERR_FORMAT, "first line does not contain \"general\" or \"symmetric\"");
  if (!has(line, "integer") && !has(line, "pattern") && !has(line, "real"))
    return fail(ECLGRAPH_ERR_FORMAT,
                "first line does not contain \"integer\" or \"pattern\" or \"real\"");
  // integer takes precedence over real, as in the reference (:70-72, :93-145)
  const int kind = has(line, "integer") ? ECLGRAPH_WEIGHTS_INT
                   : has(line, "real")  ? ECLGRAPH_WEIGHTS_REAL
                                        : ECLGRAPH_WEIGHTS_NONE;
  // comment lines: every line holding a '%' anywhere (:74); then the size line
  bool got = false;
  while ((got = next_line(&line)) && has(line, "%")) {
  }
  if (!got || line.size() < 3) return fail(ECLGRAPH_ERR_FORMAT, "could not find non-comment line");
  long long rows = 0, cols = 0, nnz = 0;
  if (std::sscanf(line.c_str(), "%lld %lld %lld", &rows, &cols, &nnz) != 3 || rows < 1 || nnz < 0 ||
      rows != cols)
    return fail(ECLGRAPH_ERR_FORMAT, "failed to parse first data line");
  if (rows > INT32_MAX || nnz > INT32_MAX)
    return fail(ECLGRAPH_ERR_RANGE, "matrix size exceeds the int32 format");
  const int32_t nodes = (int32_t)rows;
  // entries, scanned from the rest of the text
  Scanner sc{text.c_str() + pos, text.c_str() + text.size()};
  std::vector<MtxEntry> ent;
  ent.reserve((size_t)nnz);
  for (;;) {
    int64_t i, j, w = 0;
    double x = 0.0;
    if (!sc.int_token(&i) || !sc.int_token(&j)) break;
    if (kind == ECLGRAPH_WEIGHTS_INT && !sc.int_token(&w)) break;
    if (kind == ECLGRAPH_WEIGHTS_REAL && !sc.real_token(&x)) break;
    if (i < 1 || i > nodes || j < 1 || j > nodes)
      return fail(ECLGRAPH_ERR_FORMAT, "source out of range");
    if (kind == ECLGRAPH_WEIGHTS_INT && (w < INT32_MIN || w > INT32_MAX))
      return fail(ECLGRAPH_ERR_RANGE, "integer weight exceeds the int32 format");
    if (kind == ECLGRAPH_WEIGHTS_REAL && !(std::fabs(x * 100.0) < 2147483648.0))
      return fail(ECLGRAPH_ERR_RANGE, "real weight x 100 exceeds the int32 format");
    if ((long long)ent.size() >= nnz)
      return fail(ECLGRAPH_ERR_FORMAT, "failed to read correct number of edges");
    ent.push_back(MtxEntry{(int32_t)(i - 1), (int32_t)(j - 1), w, x});
  }
  if ((long long)ent.size() != nnz) return fail(ECLGRAPH_ERR_FORMAT, "failed to read correct number of edges");
  // order by (row, column, weight) as the reference's tuple sort does
  std::sort(ent.begin(), ent.end(), [kind](const MtxEntry& a, const MtxEntry& b) {
    if (a.src != b.src) return a.src < b.src;
    if (a.dst != b.dst) return a.dst < b.dst;
    return kind == ECLGRAPH_WEIGHTS_REAL ? a.fw < b.fw : a.iw < b.iw;
  });
  std::unique_ptr<eclgraph_csr> g(new (std::nothrow) eclgraph_csr);
  if (!g) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory");
  g->nodes = nodes;
  g->edges = (int64_t)ent.size();
  g->nindex = alloc_array<int32_t>((int64_t)nodes + 1);
  g->nlist = alloc_array<int32_t>(g->edges);
  if (!g->nindex || !g->nlist) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory (CSR)");
  if (kind != ECLGRAPH_WEIGHTS_NONE) {
    g->eweight = alloc_array<int32_t>(g->edges);
    if (!g->eweight) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory (weights)");
  }
  // nindex[v + 1] = entries whose row is at most v
  std::fill(g->nindex.get(), g->nindex.get() + nodes + 1, 0);
  for (const MtxEntry& e : ent) g->nindex[e.src + 1]++;
  for (int32_t v = 0; v < nodes; v++) g->nindex[v + 1] += g->nindex[v];
  for (size_t k = 0; k < ent.size(); k++) {
    g->nlist[k] = ent[k].dst;
    if (kind == ECLGRAPH_WEIGHTS_INT) g->eweight[k] = (int32_t)ent[k].iw;
    // scaled and truncated toward zero (:172)
    if (kind == ECLGRAPH_WEIGHTS_REAL) g->eweight[k] = (int32_t)(ent[k].fw * 100.0);
  }
  if (weight_kind) *weight_kind = kind;
  *out = g.release();
  return ECLGRAPH_OK;
} catch (...) {
  return host_exception();
}

int eclgraph_read_mtx(const char* path, int flags, eclgraph_csr** out) try {
  if (!path || !out) return fail(ECLGRAPH_ERR_ARG, "null argument");
  *out = nullptr;
  eclgraph_csr* raw = nullptr;
  int rc = eclgraph_mm2ecl(path, &raw, nullptr);
  if (rc) return rc;
  std::unique_ptr<eclgraph_csr> g(raw);
  if (!(flags & (ECLGRAPH_SYMMETRIZE | ECLGRAPH_DROP_SELF | ECLGRAPH_DEDUPE))) {
    *out = g.release();
    return ECLGRAPH_OK;
  }
  // (row, column, weight) triples; mirrored, filtered, sorted, deduplicated
  struct T {
    int32_t u, v, w;
  };
  const bool wt = (bool)g->eweight;
  std::vector<T> t;
  t.reserve((size_t)g->edges * ((flags & ECLGRAPH_SYMMETRIZE) ? 2 : 1));
  for (int32_t u = 0; u < g->nodes; u++)
    for (int32_t k = g->nindex[u]; k < g->nindex[u + 1]; k++) {
      const int32_t v = g->nlist[k], w = wt ? g->eweight[k] : 0;
      if ((flags & ECLGRAPH_DROP_SELF) && u == v) continue;
      t.push_back(T{u, v, w});
      if ((flags & ECLGRAPH_SYMMETRIZE) && u != v) t.push_back(T{v, u, w});
    }
  std::sort(t.begin(), t.end(), [](const T& a, const T& b) {
    return a.u != b.u ? a.u < b.u : a.v != b.v ? a.v < b.v : a.w < b.w;
  });
  if (flags & ECLGRAPH_DEDUPE)  // keeps the lightest of equal (row, column)
    t.erase(std::unique(t.begin(), t.end(), [](const T& a, const T& b) { return a.u == b.u && a.v == b.v; }),
            t.end());
  if ((int64_t)t.size() > INT32_MAX) return fail(ECLGRAPH_ERR_RANGE, "edge count exceeds int32 CSR format");
  g->edges = (int64_t)t.size();
  g->nlist = alloc_array<int32_t>(g->edges);
  if (wt) g->eweight = alloc_array<int32_t>(g->edges);
  if (!g->nlist || (wt && !g->eweight)) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory (CSR)");
  std::fill(g->nindex.get(), g->nindex.get() + g->nodes + 1, 0);
  for (size_t k = 0; k < t.size(); k++) {
    g->nindex[t[k].u + 1]++;
    g->nlist[k] = t[k].v;
    if (wt) g->eweight[k] = t[k].w;
  }
  for (int32_t v = 0; v < g->nodes; v++) g->nindex[v + 1] += g->nindex[v];
  *out = g.release();
  return ECLGRAPH_OK;
} catch (...) {
  return host_exception();
}

int eclgraph_gen_er(int32_t n, int64_t pairs, uint64_t seed, eclgraph_csr** out) try {
  if (!out) return fail(ECLGRAPH_ERR_ARG, "null argument");
  *out = nullptr;
  if (n < 1 || pairs < 0) return fail(ECLGRAPH_ERR_ARG, "n must be >= 1 and pairs >= 0");
  const double maxp = 0.5 * (double)n * (double)(n - 1);
  if ((double)pairs > 0.5 * maxp) return fail(ECLGRAPH_ERR_ARG, "too many pairs for G(n,M) sampling (> half of all pairs)");
  if (2 * pairs > INT32_MAX) return fail(ECLGRAPH_ERR_RANGE, "edge count exceeds int32 CSR format");
  // open-addressing set of normalised keys (u << 32 | v), u < v
  uint64_t cap = 16;
  while (cap < (uint64_t)pairs * 2 + 16) cap <<= 1;
  std::vector<uint64_t> table(cap, ~0ULL);
  std::vector<int32_t> su((size_t)pairs), sv((size_t)pairs);
  int64_t have = 0;
  for (uint64_t ctr = 0; have < pairs; ctr++) {
    const uint64_t r = rand64(seed, ctr);
    int32_t u = (int32_t)bounded32((uint32_t)(r >> 32), (uint32_t)n);
    int32_t v = (int32_t)bounded32((uint32_t)r, (uint32_t)n);
    if (u == v) continue;
    if (u > v) std::swap(u, v);
    const uint64_t key = ((uint64_t)(uint32_t)u << 32) | (uint32_t)v;
    uint64_t h = mix64(key) & (cap - 1);
    bool fresh = true;
    while (table[h] != ~0ULL) {
      if (table[h] == key) { fresh = false; break; }
      h = (h + 1) & (cap - 1);
    }
    if (!fresh) continue;
    table[h] = key;
    su[(size_t)have] = u;
    sv[(size_t)have] = v;
    have++;
  }
  return build_csr(n, pairs, PairArrays{su.data(), sv.data()}, ECLGRAPH_SYMMETRIZE, out);
} catch (...) {
  return host_exception();
}

int eclgraph_gen_rmat(int scale, int32_t n, int64_t samples, double a, double b, double c,
                      uint64_t seed, int permute, eclgraph_csr** out) try {
  if (!out) return fail(ECLGRAPH_ERR_ARG, "null argument");
  *out = nullptr;
  if (scale < 1 || scale > 30) return fail(ECLGRAPH_ERR_ARG, "scale must be in [1, 30]");
  const int64_t full = (int64_t)1 << scale;
  if (n <= 0) n = (int32_t)full;
  if ((int64_t)n > full) return fail(ECLGRAPH_ERR_ARG, "n exceeds 2^scale");
  if (samples < 0) return fail(ECLGRAPH_ERR_ARG, "samples must be >= 0");
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return fail(ECLGRAPH_ERR_ARG, "bad quadrant probabilities");
  // quadrant thresholds on 21-bit draws; one 64-bit draw serves three levels
  const double two21 = 2097152.0;
  const uint32_t ta = (uint32_t)std::llround(a * two21);
  const uint32_t tab = (uint32_t)std::llround((a + b) * two21);
  const uint32_t tabc = (uint32_t)std::llround((a + b + c) * two21);
  const int draws = (scale + 2) / 3;  // 64-bit draws per sample
  const uint64_t eseed = seed;
  const uint64_t pseed = mix64(seed ^ 0x5045524D55544531ULL);  // permutation stream

  std::vector<int32_t> src((size_t)samples), dst((size_t)samples);
  auto sample = [&](uint64_t trial, int32_t& su, int32_t& sv) {
    uint32_t x = 0, y = 0;
    uint64_t r64 = 0;
    for (int l = 0; l < scale; l++) {
      if (l % 3 == 0) r64 = rand64(eseed, trial * (uint64_t)draws + (uint64_t)(l / 3));
      const uint32_t r = (uint32_t)(r64 >> (21 * (l % 3))) & 0x1FFFFFu;
      const uint32_t bx = r >= tab ? 1u : 0u;
      const uint32_t by = (r >= ta && r < tab) || r >= tabc ? 1u : 0u;
      x = (x << 1) | bx;
      y = (y << 1) | by;
    }
    su = (int32_t)x;
    sv = (int32_t)y;
  };
  if ((int64_t)n == full) {
    parallel_for(samples, 1 << 16, [&](int64_t lo, int64_t hi) {
      for (int64_t k = lo; k < hi; k++) sample((uint64_t)k, src[(size_t)k], dst[(size_t)k]);
    });
  } else {
    uint64_t trial = 0;
    for (int64_t k = 0; k < samples;) {
      int32_t u, v;
      sample(trial++, u, v);
      if (u >= n || v >= n) continue;
      src[(size_t)k] = u;
      dst[(size_t)k] = v;
      k++;
    }
  }
  if (permute) {
    auto p = random_permutation(n, pseed);
    if (!p) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory (permutation)");
    parallel_for(samples, 1 << 16, [&](int64_t lo, int64_t hi) {
      for (int64_t k = lo; k < hi; k++) {
        src[(size_t)k] = p[src[(size_t)k]];
        dst[(size_t)k] = p[dst[(size_t)k]];
      }
    });
  }
  return build_csr(n, samples, PairArrays{src.data(), dst.data()},
                   ECLGRAPH_SYMMETRIZE | ECLGRAPH_DROP_SELF | ECLGRAPH_DEDUPE, out);
} catch (...) {
  return host_exception();
}

int eclgraph_gen_grid3d(int32_t nx, int32_t ny, int32_t nz, double keep_p, uint64_t seed,
                        int permute, eclgraph_csr** out) try {
  if (!out) return fail(ECLGRAPH_ERR_ARG, "null argument");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1) return fail(ECLGRAPH_ERR_ARG, "grid dimensions must be >= 1");
  const int64_t nn = (int64_t)nx * ny * nz;
  if (nn > INT32_MAX) return fail(ECLGRAPH_ERR_RANGE, "too many vertices for int32 IDs");
  const int32_t n = (int32_t)nn;
  // keep edge (v, dir) iff draw < thr; keep_p >= 1 keeps all, <= 0 none
  const bool all = keep_p >= 1.0, none = keep_p <= 0.0;
  const uint64_t thr = (all || none) ? 0 : (uint64_t)std::ldexp(keep_p, 64);
  const uint64_t pseed = mix64(seed ^ 0x5045524D55544531ULL);
  auto kept = [&](int64_t v, int dir) -> bool {
    if (all) return true;
    if (none) return false;
    return rand64(seed, (uint64_t)v * 3 + (uint64_t)dir) < thr;
  };
  const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;
  // neighbours of v in the kept-edge graph (at most 6)
  auto neighbours = [&](int64_t v, int32_t* nb) -> int {
    const int64_t x = v % nx, y = (v / nx) % ny, z = v / sz;
    int k = 0;
    if (x + 1 < nx && kept(v, 0)) nb[k++] = (int32_t)(v + sx);
    if (y + 1 < ny && kept(v, 1)) nb[k++] = (int32_t)(v + sy);
    if (z + 1 < nz && kept(v, 2)) nb[k++] = (int32_t)(v + sz);
    if (x > 0 && kept(v - sx, 0)) nb[k++] = (int32_t)(v - sx);
    if (y > 0 && kept(v - sy, 1)) nb[k++] = (int32_t)(v - sy);
    if (z > 0 && kept(v - sz, 2)) nb[k++] = (int32_t)(v - sz);
    return k;
  };
  std::unique_ptr<int32_t[]> p;
  if (permute) {
    p = random_permutation(n, pseed);
    if (!p) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory (permutation)");
  }
  auto deg = alloc_array<int32_t>(n);
  if (!deg) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory");
  parallel_for(n, 1 << 14, [&](int64_t lo, int64_t hi) {
    int32_t nb[6];
    for (int64_t v = lo; v < hi; v++) deg[p ? p[v] : v] = neighbours(v, nb);
  });
  std::unique_ptr<eclgraph_csr> g(new (std::nothrow) eclgraph_csr);
  if (!g) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory");
  g->nodes = n;
  g->nindex = alloc_array<int32_t>((int64_t)n + 1);
  if (!g->nindex) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory");
  int64_t acc = 0;
  for (int32_t x = 0; x < n; x++) {
    g->nindex[x] = (int32_t)acc;
    acc += deg[x];
  }
  if (acc > INT32_MAX) return fail(ECLGRAPH_ERR_RANGE, "edge count exceeds int32 CSR format");
  g->nindex[n] = (int32_t)acc;
  g->edges = acc;
  g->nlist = alloc_array<int32_t>(acc);
  if (!g->nlist) return fail(ECLGRAPH_ERR_ALLOC, "out of host memory");
  parallel_for(n, 1 << 14, [&](int64_t lo, int64_t hi) {
    int32_t nb[6];
    for (int64_t v = lo; v < hi; v++) {
      const int k = neighbours(v, nb);
      const int32_t id = p ? p[v] : (int32_t)v;
      int32_t* dstp = g->nlist.get() + g->nindex[id];
      for (int i = 0; i < k; i++) dstp[i] = p ? p[nb[i]] : nb[i];
      std::sort(dstp, dstp + k);
    }
  });
  *out = g.release();
  return ECLGRAPH_OK;
} catch (...) {
  return host_exception();
}

}  // extern "C"
