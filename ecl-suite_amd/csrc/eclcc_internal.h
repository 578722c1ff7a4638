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
// Largest hub_chunk the LDS staging buffer of the block-level kernel holds.
constexpr int kMaxHubChunk = 8192;

// One graph inside packed batch arrays.
struct GraphDesc {
  long long vbase;  // first label slot (multiple of kSlotAlign)
  long long ibase;  // first entry of its nindex (n+1 local offsets)
  long long ebase;  // first entry of its nlist
  int n;            // vertices
  int m;            // CSR entries
};

// High-degree task list entry: vertex slot and the index of its first edge task.
struct HubEntry {
  long long slot;
  long long first_task;
};

// Everything the five kernels read; passed by value.
struct KParams {
  const int* nidx;   // single: nindex; batch: packed nindex
  const int* nlist;  // single: nlist; batch: packed nlist
  int* nstat;        // single: labels; batch: packed label slots
  int n;             // single: vertex count (unused for batches)
  long long nslots;  // slots covered by the vertex-parallel kernels
  const GraphDesc* desc;    // batch: graph table
  const int* chunk_graph;   // batch: graph of each 256-slot chunk
  unsigned long long* midbits;  // one bit per slot: medium-degree vertex to process
  HubEntry* hubs;               // high-degree vertices with their task offsets
  unsigned long long* hubctr;   // low 32 bits: hubs pushed; high 32: edge tasks
  unsigned int* flags;          // bit 0: hub list overflow
  long long hub_cap;
  int thread_max;
  int wave_max;
  int hub_chunk;
};

// Workspace carved out of one device buffer.
struct Workspace {
  unsigned long long* hubctr;
  unsigned int* flags;
  unsigned long long* midbits;
  HubEntry* hubs;
  long long hub_cap;
};

struct Shape {
  int hub_blocks;     // grid of the block-level kernel
  int block_threads;  // 256, 512 or 1024
};

// Resolved, validated parameters.
struct Tuning {
  int thread_max, wave_max, hub_chunk, hub_blocks, block_threads;
};

size_t workspace_bytes(long long nslots, long long hub_cap);
bool carve_workspace(void* base, size_t bytes, long long nslots, long long hub_cap, Workspace* w);
long long hub_capacity(long long edges, int wave_max, long long ngraphs);

// Enqueues init, compute1, compute2, compute3, flatten on `s`.  If ev is not
// null, ev[0..5] are recorded before, between and after the kernels.
hipError_t launch_pipeline(const KParams& p, bool batch, const Shape& shape, hipStream_t s,
                           hipEvent_t* ev);
hipError_t launch_init_only(const KParams& p, hipStream_t s);

// Status helpers (thread-local message).
int set_error(int code, const std::string& msg);
int check_hip(hipError_t e, int code, const char* what);
int resolve_tuning(const eclcc_params* in, Tuning* out);
int device_cu_count();

}  // namespace eclcc
