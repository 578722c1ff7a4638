// eclcc_internal.h -- types shared by the kernels and the host side of
// libeclcc.so.  Not installed; the public surface is include/eclcc.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/eclcc.h"

namespace eclcc {

// Vertex slots of a batch are padded per graph to this multiple, so every
// 256-slot chunk (and every 64-slot wavefront) belongs to exactly one graph.
constexpr int kSlotAlign = 256;
// Largest hub_chunk (edges per block task of the block-level kernel).
constexpr int kMaxHubChunk = 65536;
// Adjacency entries each thread of the block-level kernel loads per round.
constexpr int kHubPerThread = 8;
// 64-slot words per block of the hub-list scan.
constexpr int kScanBlock = 1024;
// Vertices sampled per graph to find its largest component.
constexpr int kPickSamples = 256;
constexpr int kNumKernels = ECLCC_NUM_KERNELS;
// HIP devices the per-device tables of the library cover.
constexpr int kMaxDevices = 64;

// One graph of a batch.  Its arrays may sit anywhere in device memory (packed
// into one allocation by eclcc_batch, placed in the server's arena, or owned
// by the caller for device-resident requests); 64-bit pointers, so a batch
// may hold any total of vertices and entries.  Its vertices occupy the slots
// vbase .. vbase + n - 1 of the batch's slot space, which indexes the
// workspace bit arrays.
struct GraphDesc {
  const int* nidx;   // n+1 index entries (offsets into its nlist)
  const int* nlist;  // m adjacency entries
  int* nstat;        // n labels
  long long vbase;   // first slot (multiple of kSlotAlign)
  int n;             // vertices
  int m;             // CSR entries
};

// High-degree task list entry: vertex slot, index of its first edge task and
// the bounds of its adjacency list (local nindex values).
struct HubEntry {
  long long slot;
  long long first_task;
  int beg;
  int end;
};

// Everything the kernels read; passed by value.
struct KParams {
  const int* nidx;   // single graph: nindex (batches: per graph in desc)
  const int* nlist;  // single graph: nlist
  int* nstat;        // single graph: labels
  int n;             // single: vertex count (unused for batches)
  int ngraphs;       // 1 for a single graph
  long long nslots;  // slots covered by the vertex-parallel kernels
  const GraphDesc* desc;   // batch: graph table
  const int* chunk_graph;  // batch: graph of each 256-slot chunk
  // one 64-bit word per 64 slots
  unsigned long long* actbits;   // vertices with a smaller neighbour (init)
  unsigned long long* smpbits;   // vertices with a sampled hook (init -> sample)
  unsigned long long* lowbits;   // low-degree vertices to process (init, final after mark -> compute1)
  unsigned long long* midbits;   // medium-degree vertices to process (-> compute2)
  unsigned long long* hubbits;   // high-degree vertices to process (-> compute1 emits their tasks)
  unsigned long long* hubinfo;   // hubs | tasks << 32 of the word (mark), then its prefix in its block
  unsigned long long* gbits;     // members of the sampled largest component (mark -> compute2/3)
  unsigned long long* blocktot;  // per kScanBlock words: totals, then exclusive prefix
  int* giant;                    // per graph: sampled root of the largest component, or -1
  int* second;                   // the entry sample hooks (init -> sample): per 64-slot
                                 // word, the sampled slots' entries packed at its front
                                 // by rank
  HubEntry* hubs;                // high-degree vertices with their first task number
  unsigned long long* hubctr;    // low 32 bits: hubs; high 32: edge tasks
  unsigned int* flags;           // bit 0: hub list overflow
  long long hub_cap;
  int thread_max;
  int wave_max;
  int hub_chunk;
  int filter;     // 0: off; 1: skip edges with both ends in the sampled largest
                  // component; 2: also leave the lists of its members unread
  int path_mode;  // 0: pointer jumping at every step (reference); 1: root jump
  int sample_shift;  // second-entry hooks taken by 1 in 2^sample_shift vertices
  int sample_min_deg;  // ... whose lists hold at least this many entries
};

// Workspace carved out of one device buffer.
struct Workspace {
  unsigned long long* hubctr;
  unsigned int* flags;
  unsigned long long* actbits;
  unsigned long long* smpbits;
  unsigned long long* lowbits;
  unsigned long long* midbits;
  unsigned long long* hubbits;
  unsigned long long* hubinfo;
  unsigned long long* gbits;
  unsigned long long* blocktot;
  int* giant;
  int* second;
  HubEntry* hubs;
  long long hub_cap;
};

struct Shape {
  int hub_blocks;     // grid of the block-level kernel
  int block_threads;  // 256, 512 or 1024
};

// Resolved, validated parameters.
struct Tuning {
  int thread_max, wave_max, hub_chunk, hub_blocks, block_threads, filter, path_mode, sample_shift,
      sample_min_deg;
};

size_t workspace_bytes(long long nslots, long long hub_cap, long long ngraphs);
bool carve_workspace(void* base, size_t bytes, long long nslots, long long hub_cap,
                     long long ngraphs, Workspace* w);
long long hub_capacity(long long edges, int wave_max, long long ngraphs);
void fill_params(KParams* p, const Workspace& w, const Tuning& t);

// Enqueues the kernel sequence on `s`.  If ev is not null, ev[0..kNumKernels]
// are recorded before, between and after the kernels.
hipError_t launch_pipeline(const KParams& p, bool batch, const Shape& shape, hipStream_t s,
                           hipEvent_t* ev);
hipError_t launch_init_only(const KParams& p, hipStream_t s);
// Blocks of the block-level kernel one CU holds at once (occupancy API).
int hub_block_residency(int block_threads);


// Status helpers (thread-local message).
int set_error(int code, const std::string& msg);
// Inside a catch handler: the status for the exception in flight (host
// exceptions never cross the C ABI).
int host_exception();
int check_hip(hipError_t e, int code, const char* what);
int resolve_tuning(const eclcc_params* in, Tuning* out);
int device_cu_count();

// Block order of the batched walk kernels (a permutation of gridDim.x work
// items; the kernels take item order_item(o, blockIdx.x, gridDim.x)).  The
// items may be cut into windows of 2^wlog consecutive items, each permuted
// on its own and taken one after the other; the last, partial window (or
// the whole grid, with wlog 32) is permuted by a stride coprime with its
// size.
struct BlockOrder {
  unsigned wlog;     // log2 of the items per window; 32: one window
  unsigned wstride;  // odd
  unsigned tstride;  // coprime with the last window's size
};

__host__ __device__ inline unsigned order_item(const BlockOrder o, unsigned b, unsigned grid) {
  const unsigned full = o.wlog < 32u ? (grid >> o.wlog) << o.wlog : 0u;
  if (b < full) {
    const unsigned m = (1u << o.wlog) - 1u;
    // b * wstride wraps mod 2^32, a multiple of 2^wlog: the window's base
    // contributes nothing to the offset
    return (b & ~m) | ((b * o.wstride) & m);
  }
  const unsigned r = grid - full;
  return full + (unsigned)(((unsigned long long)(b - full) * o.tstride) % r);
}

// A stride coprime with `blocks` near its golden-ratio fraction.
inline unsigned spread_stride(unsigned blocks) {
  if (blocks <= 2) return 1;
  unsigned x = (unsigned)((double)blocks * 0.6180339887) | 1u;
  for (;; x += 2) {
    unsigned a = x, b = blocks;
    while (b) {
      const unsigned t = a % b;
      a = b;
      b = t;
    }
    if (a == 1) return x;
  }
}

// The order of `blocks` work items of `slots` slots each, in windows whose
// labels take at most `window_bytes` (0: one window over the whole grid).
inline BlockOrder block_order(unsigned blocks, long long slots, long long window_bytes) {
  BlockOrder o{32u, 1u, 1u};
  const long long items = window_bytes / (4 * slots);
  unsigned lg = 0;
  while (lg < 30 && (2ll << lg) <= items) lg++;
  if (items >= 2 && blocks > (1u << lg)) {
    o.wlog = lg;
    o.wstride = (unsigned)((double)(1u << lg) * 0.6180339887) | 1u;
    const unsigned full = (blocks >> lg) << lg;
    o.tstride = blocks > full ? spread_stride(blocks - full) : 1u;
  } else {
    o.tstride = spread_stride(blocks);
  }
  return o;
}

}  // namespace eclcc
