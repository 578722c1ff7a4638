/*
 * eclgraph.h -- host-side C ABI for the ECLgraph int32 CSR format and for the
 * seeded synthetic inputs the connected-components path is measured on.
 *
 * Library: libeclgraph.so (host C++ only, no GPU needed).
 *
 * Format (byte-compatible with the reference reader/writer,
 * library/ECLgraph.h:41-92): little-endian
 *     int32 nodes; int32 edges; int32 nindex[nodes+1]; int32 nlist[edges];
 *     [int32 eweight[edges]]            -- optional, present iff the file has it
 * "edges" counts CSR entries, so an undirected edge is stored twice.
 *
 * Every function returns 0 on success or a negative status; on failure
 * eclgraph_last_error() returns a message (thread-local).  Nothing here calls
 * exit(): the reference's reader exits the process on any error
 * (ECLgraph.h:53-72); this library reports the same conditions as statuses.
 * No C++ exception leaves the library: a host allocation that fails, or
 * worker threads that cannot start, is ECLGRAPH_ERR_ALLOC.
 */
#ifndef ECLGRAPH_ABI_H
#define ECLGRAPH_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ECLGRAPH_OK = 0,
  ECLGRAPH_ERR_ARG = -1,     /* bad argument (nodes < 1, edges < 0, NULL, ...) */
  ECLGRAPH_ERR_IO = -2,      /* open / short read / short write */
  ECLGRAPH_ERR_ALLOC = -3,   /* host allocation failed */
  ECLGRAPH_ERR_FORMAT = -4,  /* header or index inconsistent */
  ECLGRAPH_ERR_RANGE = -5    /* a count does not fit the int32 format */
};

/* Flags of eclgraph_check. */
#define ECLGRAPH_CHECK_SYMMETRY 1

/* Checks the preconditions of the CC path that the reference leaves
   unchecked (ECL-CC-atomic.cu:132 hooks each edge only from its larger end,
   so an entry whose reverse is missing may be lost): nindex[0] == 0,
   nindex non-decreasing, nindex[nodes] == edges, every nlist entry in
   [0, nodes), and with ECLGRAPH_CHECK_SYMMETRY every entry v->u has an entry
   u->v.  Returns 0, or ECLGRAPH_ERR_FORMAT with *where (if not NULL) = the
   first offending nindex position (index problems) or nlist position
   (entries), and a message. */
int eclgraph_check(int32_t nodes, int64_t edges, const int32_t* nindex, const int32_t* nlist,
                   int flags, int64_t* where);

/* Default edge weights of the reference's ECL-MST main
   (src/racefree/egr-input/MST/ECL-MST_10-atomic.cu:428-437): the entry j of
   vertex i with neighbour k weighs 1 + ((i * k) % nodes), negated if
   negative, with i * k wrapping in 32-bit two's complement as the
   reference's int product does.  eweight gets nindex[nodes] entries. */
int eclgraph_mst_weights(int32_t nodes, const int32_t* nindex, const int32_t* nlist,
                         int32_t* eweight);

/* Host CSR owned by the library; arrays are read through the accessors. */
typedef struct eclgraph_csr eclgraph_csr;

const char* eclgraph_last_error(void);

/* Replaces readECLgraph (library/ECLgraph.h:49-76): same checks
   (nodes >= 1, edges >= 0, full nindex and nlist, eweight optional). */
int eclgraph_read(const char* path, eclgraph_csr** out);

/* Replaces writeECLgraph (library/ECLgraph.h:78-92). */
int eclgraph_write(const eclgraph_csr* g, const char* path);

/* Replaces freeECLgraph (library/ECLgraph.h:94-102). NULL is accepted. */
void eclgraph_free(eclgraph_csr* g);

/* Copies caller arrays into a new library-owned CSR (eweight may be NULL). */
int eclgraph_from_arrays(int32_t nodes, int64_t edges, const int32_t* nindex,
                         const int32_t* nlist, const int32_t* eweight,
                         eclgraph_csr** out);

/* Accessors (struct ECLgraph fields, library/ECLgraph.h:41-47). */
int32_t eclgraph_nodes(const eclgraph_csr* g);
int64_t eclgraph_edges(const eclgraph_csr* g);
const int32_t* eclgraph_nindex(const eclgraph_csr* g);
const int32_t* eclgraph_nlist(const eclgraph_csr* g);
const int32_t* eclgraph_eweight(const eclgraph_csr* g); /* NULL if absent */

/* Builds a CSR from a directed edge list (src[i] -> dst[i]).
   flags: bit0 symmetrise (add dst->src), bit1 drop self-loops,
          bit2 remove duplicate entries. Adjacency lists come out sorted.
   Used for MatrixMarket-style inputs (scripts/mm2ecl_real.cpp:107-115 sorts
   pairs the same way but neither symmetrises nor dedupes). */
enum { ECLGRAPH_SYMMETRIZE = 1, ECLGRAPH_DROP_SELF = 2, ECLGRAPH_DEDUPE = 4 };
int eclgraph_from_edges(int32_t nodes, int64_t nedges, const int32_t* src,
                        const int32_t* dst, int flags, eclgraph_csr** out);

/* Weights of a converted MatrixMarket file. */
enum { ECLGRAPH_WEIGHTS_NONE = 0, ECLGRAPH_WEIGHTS_INT = 1, ECLGRAPH_WEIGHTS_REAL = 2 };

/* Replaces scripts/mm2ecl_real.cpp (:50-184): the MatrixMarket coordinate
   file at `path` becomes the CSR the reference's converter writes, byte for
   byte once written with eclgraph_write:
     - the banner line holds at least 30 characters and the words
       "%%MatrixMarket", "matrix", "coordinate", "general" or "symmetric",
       and "integer", "pattern" or "real" (case-sensitive);
     - every following line that holds a '%' is a comment; the size line
       must be square (rows == columns >= 1, entries >= 0);
     - entries are kept as given: a "symmetric" file is NOT mirrored,
       duplicates and self-loops stay; indices must lie in [1, rows] and
       their count must equal the size line's;
     - entries are ordered by (row, column, weight);
     - "integer" files keep their weights, "real" files store
       (int)(weight * 100) (truncated toward zero), "pattern" files have
       none.  *weight_kind (optional) receives ECLGRAPH_WEIGHTS_*.
   Failures the reference reports with exit(-1) are ECLGRAPH_ERR_FORMAT /
   ERR_IO; values it leaves undefined (an int32 overflow) are
   ECLGRAPH_ERR_RANGE. */
int eclgraph_mm2ecl(const char* path, eclgraph_csr** out, int* weight_kind);

/* eclgraph_mm2ecl followed by the flags of eclgraph_from_edges on the
   (row, column, weight) entries: ECLGRAPH_SYMMETRIZE adds each reverse
   entry with the same weight, ECLGRAPH_DROP_SELF drops self-loops,
   ECLGRAPH_DEDUPE keeps the lightest of equal (row, column) entries.
   flags = 0 is the reference's conversion; ECL-CC needs the symmetric
   CSR that SYMMETRIZE | DROP_SELF | DEDUPE gives.  Weights are kept. */
int eclgraph_read_mtx(const char* path, int flags, eclgraph_csr** out);

/* ---- seeded synthetic inputs (the reference downloads its inputs,
   download_inputs.sh:11-27; these generators replace that step) ----
   All randomness is splitmix64 on (seed, counter), so the output is a pure
   function of the arguments and independent of the thread count. */

/* Erdos-Renyi G(n, M): exactly `pairs` distinct undirected pairs {u,v},
   u != v, taken in stream order from the seeded candidate stream; CSR is
   symmetric with sorted adjacency. Vertex IDs are not permuted. */
int eclgraph_gen_er(int32_t n, int64_t pairs, uint64_t seed, eclgraph_csr** out);

/* R-MAT / Kronecker: `samples` directed samples over [0,2^scale)^2 with
   quadrant probabilities (a,b,c,1-a-b-c), no noise. If n < 2^scale, samples
   with an endpoint >= n are redrawn. Then symmetrised, self-loops dropped,
   duplicates removed, IDs randomly permuted when `permute` != 0, adjacency
   sorted. n <= 0 means n = 2^scale. */
int eclgraph_gen_rmat(int scale, int32_t n, int64_t samples, double a, double b,
                      double c, uint64_t seed, int permute, eclgraph_csr** out);

/* 3-D 7-point grid nx*ny*nz (non-periodic); each of the
   (nx-1)ny nz + nx(ny-1)nz + nx ny(nz-1) undirected edges is kept with
   probability keep_p; IDs randomly permuted when `permute` != 0. */
int eclgraph_gen_grid3d(int32_t nx, int32_t ny, int32_t nz, double keep_p,
                        uint64_t seed, int permute, eclgraph_csr** out);

/* splitmix64 counter stream: output number `counter` of the generator
   started at state `seed` (exposed so tests and scripts can draw the same
   numbers as the generators). */
uint64_t eclgraph_rand64(uint64_t seed, uint64_t counter);

/* Number of worker threads the generators use (0 = automatic: the hardware
   thread count capped at 16, or ECLGRAPH_THREADS if set). */
void eclgraph_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
