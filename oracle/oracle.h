/*
 * oracle.h -- CPU oracle for the race-free ECL-CC connected-components path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker
 * (or, for the CPU baseline, as the thing timed beside the GPU).  The product
 * library (libeclcc.so) never links or calls it and has no CPU fallback.
 *
 * Reference algorithm: src/racefree/egr-input/CC/ECL-CC-atomic.cu (init
 * :69-88, representative :92-104, compute1 :108-157, compute2 :161-199,
 * compute3 :203-246, flatten :250-264, host verifier :327-336 and :365-395).
 *
 * What the output must be: label[v] = the minimum vertex ID in v's connected
 * component, for every interleaving of the GPU threads (the reference's
 * pointers only ever move to smaller IDs and every edge is hooked, so each
 * component ends as one tree rooted at its minimum).
 *
 * How each function is pinned (tests/test_oracle.py):
 *   oracle_cc_unionfind  - committed golden vectors whose labels were computed
 *                          by scipy.sparse.csgraph (tests/golden/make_golden.py),
 *                          brute-force BFS labels, closed forms (empty graph,
 *                          path, star, clique, disjoint cliques, full grid).
 *   oracle_cc_bfs        - same fixtures; an independent algorithm.
 *   oracle_init          - closed form on hand-built lists (first smaller
 *                          neighbour in list order, else v), incl. unsorted lists.
 *   oracle_verify        - accepts the golden labels; rejects each planted
 *                          mistake (split component, merged components,
 *                          non-minimum ID, negative / out-of-range label).
 *   oracle_eclcc_serial / oracle_eclcc_threads
 *                        - equal to the golden labels and to union-find on
 *                          every fixture and on random graphs.
 */
#ifndef ECLCC_ORACLE_H
#define ECLCC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Serial union-find with min-linking (the larger root is hung under the
   smaller) and path halving; writes label[v] = min ID of v's component.
   Pattern of the reference's serial union-find in
   src/racefree/egr-input/MST/ECL-MST_10-atomic.cu:85-109, with the link
   direction reversed (that code hangs the smaller root under the larger). */
int oracle_cc_unionfind(int32_t n, const int32_t* nindex, const int32_t* nlist, int32_t* label);

/* Breadth-first labelling: vertices are visited in increasing ID order and
   each new BFS labels its component with its start vertex, which is the
   component minimum.  Independent of union-find; O(n + m). */
int oracle_cc_bfs(int32_t n, const int32_t* nindex, const int32_t* nlist, int32_t* label);

/* init (ECL-CC-atomic.cu:69-88): nstat[v] = first neighbour smaller than v
   in list order, else v. */
void oracle_init(int32_t n, const int32_t* nindex, const int32_t* nlist, int32_t* nstat);

/* Restated host verifier (ECL-CC-atomic.cu:365-395, verify :327-336), made
   iterative and run on a private copy of the labels, plus the minimum-ID
   check the reference leaves out.  Returns 0 if the labels are correct, else
   one of the ORACLE_VERIFY_* codes.  *ncomp (if not NULL) receives the
   number of components (distinct labels). */
enum {
  ORACLE_VERIFY_OK = 0,
  ORACLE_VERIFY_ADJACENT = -1,   /* "found adjacent nodes in different components" */
  ORACLE_VERIFY_NEGATIVE = -2,   /* "found negative component number" */
  ORACLE_VERIFY_WRONG_ID = -3,   /* "found incorrect ID value" (DFS) */
  ORACLE_VERIFY_COUNT = -4,      /* "number of components do not match" */
  ORACLE_VERIFY_NOT_UNIQUE = -5, /* "component IDs are not unique" */
  ORACLE_VERIFY_RANGE = -6,      /* label >= n (not a vertex ID) */
  ORACLE_VERIFY_NOT_MIN = -7,    /* label is not the component minimum */
  ORACLE_VERIFY_ALLOC = -8
};
int oracle_verify(int32_t n, const int32_t* nindex, const int32_t* nlist, const int32_t* label,
                  int64_t* ncomp);

/* CPU-1 baseline: the ECL-CC algorithm restated serially -- init, degree
   binning into the double-ended worklist with the reference's thresholds
   (16, 352), hooking through representative() with intermediate pointer
   jumping and min-linking, then flatten.  Output = final nstat. */
int oracle_eclcc_serial(int32_t n, const int32_t* nindex, const int32_t* nlist, int32_t* nstat);

/* CPU-N baseline: the same phases on nthreads POSIX threads (0 = all online
   cores), relaxed __atomic loads/stores and CAS on nstat, worklist counters
   by fetch-add, high-degree lists split into edge chunks across threads.
   Returns the thread count used (> 0) or a negative status. */
int oracle_eclcc_threads(int32_t n, const int32_t* nindex, const int32_t* nlist, int32_t* nstat,
                         int nthreads);

/* Number of distinct values among label[0..n) that lie in [0, n). */
int64_t oracle_count_components(int32_t n, const int32_t* label);

/* ---- ECL-MST (src/racefree/egr-input/MST/ECL-MST_10-atomic.cu) ----------
   What the output must be: in_mst[j] = 1 for the entry j (vertex i, neighbour
   k = nlist[j] > i) of every edge of the minimum spanning forest under the
   strict order (weight, j); 0 for every other entry, including the k < i
   direction.  The reference's GPU code keys each edge by weight << 32 | j
   (:251) and its CPU Kruskal sorts by the tuple (weight, j, i, k)
   (:122-131), so both pick this one forest.
   Pinned (tests/test_mst.py): the forest's total weight equals
   scipy.sparse.csgraph.minimum_spanning_tree's on random graphs; the edge
   set equals the brute-force minimum over all spanning forests of the keys
   weight * 2^32 + j on every graph of 5 vertices with small weights; it has
   n - (number of components) edges and joins each component; closed forms
   for a tree (every edge) and for a cycle (all but the largest key). */

/* Default weights (ECL-MST_10-atomic.cu:428-437). */
void oracle_mst_weights(int32_t n, const int32_t* nindex, const int32_t* nlist, int32_t* w);

/* Kruskal (ECL-MST_10-atomic.cu:112-154).  Weights must be >= 0.  Returns 0,
   -1 on bad arguments, -2 on an ID out of range or a negative weight, -3 if
   out of memory. */
int oracle_mst(int32_t n, const int32_t* nindex, const int32_t* nlist, const int32_t* w,
               uint8_t* in_mst);

#ifdef __cplusplus
}
#endif
#endif
