/*
 * eclmst_oracle.c -- CPU oracle for the race-free ECL-MST path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and the cpu_baseline leg of
 * bench.py.  Never linked into the product library.  See oracle.h for what
 * each function restates and how it is pinned.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* Default weights of the reference's main
   (src/racefree/egr-input/MST/ECL-MST_10-atomic.cu:428-437): the entry j of
   vertex i with neighbour k weighs 1 + ((i * k) % nodes), made positive if
   negative.  i * k is an int product there, so it wraps in two's
   complement; the product is formed in unsigned arithmetic here to get the
   same bits without the undefined signed overflow. */
void oracle_mst_weights(int32_t n, const int32_t* nindex, const int32_t* nlist, int32_t* w) {
  for (int32_t i = 0; i < n; i++) {
    for (int32_t j = nindex[i]; j < nindex[i + 1]; j++) {
      const int32_t k = nlist[j];
      const int32_t prod = (int32_t)((uint32_t)i * (uint32_t)k);
      int32_t x = 1 + prod % n;
      if (x < 0) x = -x;
      w[j] = x;
    }
  }
}

typedef struct {
  uint64_t key; /* weight << 32 | entry index: a strict total order */
  int32_t a, b;
} mst_edge;

static int cmp_edge(const void* p, const void* q) {
  const uint64_t x = ((const mst_edge*)p)->key, y = ((const mst_edge*)q)->key;
  return x < y ? -1 : x > y;
}

static int32_t mst_find(int32_t* parent, int32_t x) {
  while (parent[x] != x) {
    parent[x] = parent[parent[x]]; /* path halving */
    x = parent[x];
  }
  return x;
}

/* Kruskal (ECL-MST_10-atomic.cu:112-154): the edges i-k with k > i (one
   direction, entry j) sorted by (weight, j), each joined if its ends are in
   different trees.  Weights must be non-negative (the reference's are >= 0
   by construction); the key keeps the weight in the high 32 bits. */
int oracle_mst(int32_t n, const int32_t* nindex, const int32_t* nlist, const int32_t* w,
               uint8_t* in_mst) {
  if (n < 1 || !nindex || !in_mst) return -1;
  const int64_t m = nindex[n];
  memset(in_mst, 0, (size_t)m);
  int64_t cnt = 0;
  for (int32_t i = 0; i < n; i++)
    for (int32_t j = nindex[i]; j < nindex[i + 1]; j++) {
      if (nlist[j] < 0 || nlist[j] >= n || w[j] < 0) return -2;
      cnt += nlist[j] > i;
    }
  mst_edge* e = (mst_edge*)malloc(sizeof(mst_edge) * (size_t)(cnt ? cnt : 1));
  int32_t* parent = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  if (!e || !parent) {
    free(e);
    free(parent);
    return -3;
  }
  int64_t c = 0;
  for (int32_t i = 0; i < n; i++)
    for (int32_t j = nindex[i]; j < nindex[i + 1]; j++)
      if (nlist[j] > i) {
        e[c].key = ((uint64_t)(uint32_t)w[j] << 32) | (uint32_t)j;
        e[c].a = i;
        e[c].b = nlist[j];
        c++;
      }
  qsort(e, (size_t)cnt, sizeof(mst_edge), cmp_edge);
  for (int32_t v = 0; v < n; v++) parent[v] = v;
  int32_t left = n - 1;
  for (int64_t t = 0; t < cnt && left > 0; t++) {
    const int32_t ra = mst_find(parent, e[t].a), rb = mst_find(parent, e[t].b);
    if (ra == rb) continue;
    in_mst[(uint32_t)e[t].key] = 1;
    if (ra < rb) parent[rb] = ra;
    else parent[ra] = rb;
    left--;
  }
  free(e);
  free(parent);
  return 0;
}
