/*
 * eclcc.h -- C ABI of the MI355X (gfx950) race-free ECL-CC connected
 * components library, libeclcc.so.
 *
 * The reference's only in-process face for this path is
 *     static void computeCC(const int nodes, const int edges,
 *                           const int* nidx, const int* nlist, int* nstat)
 * (src/racefree/egr-input/CC/ECL-CC-atomic.cu:275-325), file-local and
 * called once from main (:363).  eclcc_compute() below is that function with
 * a status return instead of exit(-1).  Everything else is added around it:
 * device-resident runs for callers that keep the CSR in HBM, a batched form
 * (many graphs, one launch sequence, capturable in a hipGraph) and a serving
 * loop with admission by HBM budget.
 *
 * Data contract (same as the reference):
 *   nindex[nodes+1], nlist[edges]: int32 CSR of a SYMMETRIC graph; IDs in
 *   [0, nodes).  Self-loops and duplicate entries are harmless.
 *   Output nstat[nodes]: nstat[v] = the minimum vertex ID in v's connected
 *   component (the reference's final state; identical for every thread
 *   interleaving).
 *
 * Conventions: every function returns ECLCC_OK (0) or a negative status and
 * never exits the process; eclcc_last_error() gives the message (per host
 * thread).  No C++ exception leaves the library: a host allocation that
 * fails is ECLCC_ERR_ALLOC.  `stream` arguments are hipStream_t passed as void* (NULL = the
 * null stream).  All device pointers belong to the current HIP device.
 */
#ifndef ECLCC_ABI_H
#define ECLCC_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ECLCC_OK = 0,
  ECLCC_ERR_ARG = -1,      /* invalid argument */
  ECLCC_ERR_ALLOC = -2,    /* device or host allocation failed ("could not allocate ...") */
  ECLCC_ERR_COPY = -3,     /* host<->device copy failed ("copying to/from device failed") */
  ECLCC_ERR_LAUNCH = -4,   /* kernel launch or execution error (CheckCuda) */
  ECLCC_ERR_DEVICE = -5,   /* no usable HIP device */
  ECLCC_ERR_BUDGET = -6,   /* request larger than the HBM budget */
  ECLCC_ERR_STATE = -7,    /* object in the wrong state / unknown id */
  ECLCC_ERR_OVERFLOW = -8, /* internal worklist capacity exceeded */
  ECLCC_ERR_TIMEOUT = -9   /* wait timed out */
};

const char* eclcc_last_error(void);
const char* eclcc_status_string(int status);
/* Library version string and the offload target it was built for. */
const char* eclcc_version(void);

/* ---- tuning parameters (degree thresholds and launch shapes) ---------- */
typedef struct eclcc_params {
  int32_t thread_max_deg; /* degree <= this: one lane per vertex (reference: 16, :120) */
  int32_t wave_max_deg;   /* degree <= this: one 64-lane wavefront per vertex
                             (reference: 352 = 11 x 32-lane warps, :122) */
  int32_t hub_chunk;      /* edges per block task for degree > wave_max_deg */
  int32_t hub_blocks;     /* blocks of the block-level kernel (0 = as many as fit at once) */
  int32_t block_threads;  /* threads per block of the block-level kernel (256..1024) */
  int32_t giant_filter;   /* after the sampled hooks, mark the vertices joined to each
                             graph's largest tree (exact: an edge with both ends marked
                             joins one tree already).  0: off, every edge hooked from its
                             larger end as in the reference; 1: every list read, edges
                             with both ends marked skipped after a bitmap lookup; 2
                             (default): lists of marked vertices not read at all, each
                             unmarked vertex hooks all of its entries */
  int32_t path_mode;      /* shortening of find paths: 0 = re-point every visited vertex
                             to its grandparent (the reference's intermediate pointer
                             jumping, ECL-CC-atomic.cu:92-104); 1 = re-point only the
                             start vertex, to the root (one store per walk); 2 = as 1,
                             and the sampled hooks' meeting walks also halve the path
                             they climb (a store per two levels).  The sampled hooks
                             walk both paths together to where they meet in every
                             mode */
  int32_t sample_shift;   /* the sampled hook of a vertex that has a smaller neighbour
                             (its second entry) is taken by 1 in 2^sample_shift of
                             them, picked by a hash of the ID; 0 (default): by all.
                             Vertices without a smaller neighbour always take theirs.
                             The labels do not depend on it (0..4) */
  int32_t sample_min_deg; /* and only those whose list has at least this many entries
                             (0, the default: every length).  A short list left
                             outside the largest tree costs compute one thread; a
                             long one costs a wavefront, so on skewed graphs the
                             sampled hooks are best spent on long lists (R-MAT-24
                             with 16: 5 % less time per pass).  Labels do not
                             depend on it (>= 0) */
} eclcc_params;

/* Fills the built-in defaults (thresholds re-derived for 64-lane waves). */
void eclcc_params_default(eclcc_params* p);

/* Fills *p with the defaults tuned for this graph's class, chosen on the
   host from its degree statistics.  Graphs of fewer than 2^20 vertices
   keep the built-in defaults; for larger ones, graphs whose lists all have
   about the
   same length (3-D grids, Erdos-Renyi: the largest of 4096 evenly spaced
   degrees is at most 4 x the average + 8) take the sampled second-entry
   hooks of half of their vertices (sample_shift 1; grid 256^3: 1.09 -> 0.96
   ms per pass on MI355X), and from 2^23 vertices on also halve the paths
   of their meeting walks (path_mode 2; Erdos-Renyi n = 2^24, degree 6:
   1.11 -> 0.90 ms); skewed degree distributions (R-MAT, power-law)
   keep them all for the vertices with at least 16 entries and halve the
   paths their meeting walks climb (sample_shift 0, sample_min_deg 16,
   path_mode 2; R-MAT-24: 0.500 -> 0.491 ms with the halving, 0.475 ->
   0.448 ms with the length limit, 0.59 with sample_shift 1).  nindex is host memory; at most 8192 of its entries are
   read.  The labels do not depend on the choice. */
int eclcc_params_for_graph(int32_t nodes, int64_t edges, const int32_t* nindex, eclcc_params* p);

/* ---- plain drop-in ----------------------------------------------------- */

/* Replaces computeCC (ECL-CC-atomic.cu:275-325): host CSR in, host labels
   out.  Allocates, copies, runs the kernel sequence (init ... flatten),
   copies back and frees, like the reference.  edges is the CSR length. */
int eclcc_compute(int32_t nodes, int32_t edges, const int32_t* nindex, const int32_t* nlist,
                  int32_t* nstat);

/* Same, and *seconds = device time of the kernel sequence only (the window
   of the reference's GPUTimer, ECL-CC-atomic.cu:305-312).  params NULL:
   eclcc_params_for_graph's choice (eclcc_compute passes NULL). */
int eclcc_compute_timed(int32_t nodes, int32_t edges, const int32_t* nindex, const int32_t* nlist,
                        int32_t* nstat, const eclcc_params* params, double* seconds);

/* ---- device-resident (no allocation, no copies, asynchronous) ---------- */

/* Device workspace the run needs (worklist bitmap, high-degree task list,
   counters), in bytes. */
size_t eclcc_workspace_bytes(int32_t nodes, int64_t edges, const eclcc_params* params);

/* Enqueues the kernel sequence on `stream` over device-resident CSR:
   d_nstat[nodes] receives the labels; d_work is workspace of at least
   eclcc_workspace_bytes() bytes.  Returns after launching (no sync). */
int eclcc_run_device(int32_t nodes, int32_t edges, const int32_t* d_nindex,
                     const int32_t* d_nlist, int32_t* d_nstat, void* d_work, size_t work_bytes,
                     const eclcc_params* params, void* stream);

/* No counterpart in the reference (it allocates on every call,
   ECL-CC-atomic.cu:290-293).  For callers that cannot allocate device
   memory: code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc.  Same as eclcc_run_device, but the workspace is a fixed
   device array of ECLCC_NOALLOC_SCRATCH_BYTES inside the library (one per
   GPU, held until the process ends); the call allocates nothing.  Graphs
   whose workspace (eclcc_workspace_bytes) exceeds it are refused with
   ECLCC_ERR_BUDGET (with the default parameters: about 4.9 bytes per vertex
   and 0.012 per entry, so about 13 million vertices).  Calls that overlap
   on one GPU share the array and race: keep them on one stream, or wait
   for one before starting the next. */
#define ECLCC_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int eclcc_run_device_noalloc(int32_t nodes, int32_t edges, const int32_t* d_nindex,
                             const int32_t* d_nlist, int32_t* d_nstat,
                             const eclcc_params* params, void* stream);

/* Kernels of one run, in launch order: init, sample, pick, mark, hubscan,
   compute1, compute2, compute3, flatten. */
#define ECLCC_NUM_KERNELS 9
const char* eclcc_kernel_name(int k);

/* As eclcc_run_device, but records HIP events around each kernel, waits,
   and writes their durations (ms) to kernel_ms[ECLCC_NUM_KERNELS] and the
   whole window to *total_ms. */
int eclcc_run_device_profiled(int32_t nodes, int32_t edges, const int32_t* d_nindex,
                              const int32_t* d_nlist, int32_t* d_nstat, void* d_work,
                              size_t work_bytes, const eclcc_params* params, void* stream,
                              float* kernel_ms, float* total_ms);

/* Test hook: runs only the init kernel (ECL-CC-atomic.cu:69-88) so its
   closed form can be checked: d_nstat[v] = first neighbour < v in list
   order, else v.  (The full pass with giant_filter 2 looks at the first two
   entries only and leaves v where neither is smaller; it runs init as
   here with giant_filter 0 and 1.) */
int eclcc_run_init_only(int32_t nodes, int32_t edges, const int32_t* d_nindex,
                        const int32_t* d_nlist, int32_t* d_nstat, void* d_work, size_t work_bytes,
                        void* stream);

/* Reads the worklist counters of the last run in d_work (synchronises
   `stream`): number of high-degree vertices whose lists that run hooked in
   edge tasks (with giant_filter 2 only those outside the sampled largest
   tree), and number of those tasks. */
int eclcc_workspace_stats(const void* d_work, size_t work_bytes, void* stream, int64_t* n_hubs,
                          int64_t* n_hub_tasks);

/* ---- batched: many graphs, one launch sequence ------------------------- */

/* A batch packs G graphs into one device allocation:
     graph g's vertices occupy "slots" [slot0_g, slot0_g + n_g) of one label
     array (slot0 padded to a multiple of 256), its nindex sits at
     nindex_off_g in one packed int32 index array (n_g+1 local offsets), its
     nlist at nlist_off_g (int64 offsets: the batch may exceed 2^31 entries
     in total; each graph's own CSR stays int32).
   The kernel sequence runs once over all graphs.  Each graph's labels are the
   same as eclcc_compute() on that graph alone. */
typedef struct eclcc_batch eclcc_batch;

/* Creates a batch on the current device and uploads the graphs
   (host arrays; nindex[g] has nodes[g]+1 entries, nlist[g] edges[g]). */
int eclcc_batch_create(int32_t ngraphs, const int32_t* nodes, const int32_t* edges,
                       const int32_t* const* nindex, const int32_t* const* nlist,
                       const eclcc_params* params, eclcc_batch** out);

/* Enqueues the batched run on `stream`; when use_graph != 0 the launch
   sequence is captured once into a hipGraph and replayed afterwards. */
int eclcc_batch_run(eclcc_batch* b, void* stream, int use_graph);

/* Replaces the batch's tuning parameters (used by the autotuner on resident
   data).  Reallocates the workspace if its size changes and drops the
   captured hipGraph; the next eclcc_batch_run recaptures it. */
int eclcc_batch_set_params(eclcc_batch* b, const eclcc_params* params);

/* Copies graph g's labels (nodes[g] ints) to host memory (synchronous). */
int eclcc_batch_labels(eclcc_batch* b, int32_t g, int32_t* nstat);

/* Device pointer to graph g's labels (valid until eclcc_batch_free). */
int32_t* eclcc_batch_device_labels(eclcc_batch* b, int32_t g);

/* Total device bytes held by the batch. */
size_t eclcc_batch_device_bytes(const eclcc_batch* b);

void eclcc_batch_free(eclcc_batch* b);

/* As eclcc_batch_run without the hipGraph, recording HIP events around each
   kernel; waits and writes kernel_ms[ECLCC_NUM_KERNELS] and *total_ms. */
int eclcc_batch_run_profiled(eclcc_batch* b, void* stream, float* kernel_ms, float* total_ms);

/* Runs graph g of a batch ALONE through the plain device path (its own
   kernel sequence) on `stream`: the one-graph-per-launch baseline
   the batched run is compared with, on the same resident data. */
int eclcc_batch_run_single(eclcc_batch* b, int32_t g, void* stream);

/* ---- served: continuous batching with admission by HBM budget ---------- */
typedef struct eclcc_server eclcc_server;

typedef struct eclcc_server_config {
  size_t hbm_budget_bytes; /* device bytes the server may hold (0 = 80% of free HBM) */
  int32_t max_inflight;    /* requests resident on the device at once (0 = 4096) */
  int32_t streams;         /* concurrent batch streams (0 = 4) */
  int32_t max_batch;       /* most requests fused into one launch (0 = 1024) */
  int32_t batch_wait_us;   /* while another batch is on the GPU, a batch smaller than
                              max_batch waits until its oldest request has waited
                              this long (0 = launch whatever is waiting at once); an
                              idle GPU always takes the waiting requests at once */
  int32_t reserved[4];
} eclcc_server_config;

/* Starts the serving thread on the current device. params may be NULL. */
int eclcc_server_create(const eclcc_server_config* cfg, const eclcc_params* params,
                        eclcc_server** out);

/* Submits one CC request.  The host arrays must stay valid and unchanged
   until the request completes; labels_out[nodes] receives the result.
   Returns a request id (>= 0) or a negative status (ECLCC_ERR_BUDGET if the
   graph alone exceeds the budget). */
int64_t eclcc_server_submit(eclcc_server* s, int32_t nodes, int32_t edges, const int32_t* nindex,
                            const int32_t* nlist, int32_t* labels_out);

/* Submits one CC request on a graph that is already in device memory (a
   graph store in HBM, or the output of an earlier GPU stage): d_nindex,
   d_nlist and d_labels are device pointers on the server's device and must
   stay valid until the request completes; the labels are written to
   d_labels[nodes] and nothing crosses PCIe.  The budget covers only the
   batch workspace.  Returns a request id or a negative status. */
int64_t eclcc_server_submit_device(eclcc_server* s, int32_t nodes, int32_t edges,
                                   const int32_t* d_nindex, const int32_t* d_nlist,
                                   int32_t* d_labels);

/* Submits `count` device-resident requests at once (arrays of their sizes
   and device pointers, as for eclcc_server_submit_device); ids[count]
   receives their ids.  All are checked before any is queued, and they
   arrive together, so the worker can pack them into full batches.  Returns
   0 or a negative status. */
int eclcc_server_submit_device_many(eclcc_server* s, int32_t count, const int32_t* nodes,
                                    const int32_t* edges, const int32_t* const* d_nindex,
                                    const int32_t* const* d_nlist, int32_t* const* d_labels,
                                    int64_t* ids);

/* Waits for request `id` (timeout_ms < 0: forever).  *latency_ms (optional)
   = submit -> labels in host memory (device memory for a device request). */
int eclcc_server_wait(eclcc_server* s, int64_t id, int32_t timeout_ms, double* latency_ms);

/* Server counters: completed requests, launched batches, peak device bytes,
   largest batch. Any pointer may be NULL. */
int eclcc_server_stats(eclcc_server* s, int64_t* completed, int64_t* batches,
                       size_t* peak_bytes, int32_t* max_batch_seen);

/* Drains outstanding requests and stops the server. */
void eclcc_server_free(eclcc_server* s);

#ifdef __cplusplus
}
#endif
#endif
