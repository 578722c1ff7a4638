/*
 * eclmst.h -- C ABI of the MI355X (gfx950) race-free ECL-MST minimum
 * spanning forest library, libeclmst.so.
 *
 * The reference's face for this path is
 *     template <bool filter> static bool* gpuMST(const ECLgraph& g, const int threshold)
 * (src/racefree/egr-input/MST/ECL-MST_10-atomic.cu:285-382), file-local and
 * called from main after it picks the light-edge threshold (:450-462).
 * eclmst_compute() below is both with a status return instead of exit(-1),
 * and with the output written into the caller's array instead of a new[]ed
 * one.  eclmst_run_device() is added for callers that keep the graph in HBM.
 *
 * Data contract (same as the reference):
 *   nindex[nodes+1], nlist[edges], eweight[edges]: int32 CSR of a SYMMETRIC
 *   graph with weights >= 0 (the reference's default weights, :428-437, are;
 *   eclgraph_mst_weights() in eclgraph.h computes them).  IDs in [0, nodes).
 *   Output in_mst[edges]: 1 for the entry j of vertex i with nlist[j] > i of
 *   every edge of the minimum spanning forest, else 0 (the k < i direction
 *   of an edge is always 0).  Ties are broken by j: the forest is the unique
 *   minimum under the order (weight, j), as in the reference, whose GPU code
 *   keys each edge by weight << 32 | j (:251) and whose CPU Kruskal sorts
 *   by (weight, j, ...) (:122-131).
 *
 * Conventions as in eclcc.h: 0 or a negative ECLMST_ERR_* status, never an
 * exit or an escaping C++ exception (a failed host allocation is
 * ECLMST_ERR_ALLOC); eclmst_last_error() gives the message (per host thread).  Calls
 * from several host threads are safe and run one at a time (they share the
 * small pinned buffer through which each round's list size is read).
 */
#ifndef ECLMST_ABI_H
#define ECLMST_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ECLMST_OK = 0,
  ECLMST_ERR_ARG = -1,      /* invalid argument */
  ECLMST_ERR_ALLOC = -2,    /* device or host allocation failed */
  ECLMST_ERR_COPY = -3,     /* host<->device copy failed */
  ECLMST_ERR_LAUNCH = -4,   /* kernel launch or execution error */
  ECLMST_ERR_DEVICE = -5,   /* no usable HIP device */
  ECLMST_ERR_FORMAT = -6,   /* negative weight, or an edge list longer than a
                               symmetric graph allows (the worklists hold
                               edges / 2 entries, as the reference's :299, :305) */
  ECLMST_ERR_WORKSPACE = -7 /* workspace too small */
};

const char* eclmst_last_error(void);
const char* eclmst_version(void);

/* Light-edge threshold that main picks for graphs of average degree >= 4
   (ECL-MST_10-atomic.cu:450-462: the weight below which about 3 * nodes of
   the entries lie), estimated from 256 evenly spaced entries; INT32_MAX
   (one phase, no filter) for sparser graphs.  The forest does not depend
   on it. */
int32_t eclmst_threshold(int32_t nodes, int32_t edges, const int32_t* eweight);

/* Replaces gpuMST (ECL-MST_10-atomic.cu:285-382) plus main's threshold
   choice: host CSR and weights in, in_mst[edges] out.  Allocates, copies,
   runs, copies back and frees, as the reference does. */
int eclmst_compute(int32_t nodes, int32_t edges, const int32_t* nindex, const int32_t* nlist,
                   const int32_t* eweight, uint8_t* in_mst);

/* As eclmst_compute; *seconds receives the device time of the kernel
   sequence (the GPUTimer window of :321-363). */
int eclmst_compute_timed(int32_t nodes, int32_t edges, const int32_t* nindex,
                         const int32_t* nlist, const int32_t* eweight, uint8_t* in_mst,
                         double* seconds);

/* Device-resident form: all pointers are device memory of the current
   device; d_work holds eclmst_workspace_bytes() bytes.  threshold < 0 picks
   it on the device as eclmst_threshold() does; INT32_MAX runs one phase.
   Runs on `stream` (hipStream_t as void*) and returns when the forest is in
   d_in_mst: each Boruvka round reads its worklist size back to the host, as
   the reference does (:332-357).  *device_ms (if not NULL) receives the
   time from the first kernel to the last. */
size_t eclmst_workspace_bytes(int32_t nodes, int32_t edges);
int eclmst_run_device(int32_t nodes, int32_t edges, const int32_t* d_nindex,
                      const int32_t* d_nlist, const int32_t* d_eweight, uint8_t* d_in_mst,
                      void* d_work, size_t work_bytes, int32_t threshold, void* stream,
                      float* device_ms);

#ifdef __cplusplus
}
#endif

#endif
