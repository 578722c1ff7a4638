// eclcc_kernels.hip -- the five race-free ECL-CC kernels for gfx950.
//
// Reference algorithm: src/racefree/egr-input/CC/ECL-CC-atomic.cu
//   init :69-88, representative :92-104, compute1 :108-157,
//   compute2 :161-199, compute3 :203-246, flatten :250-264.
//
// Race freedom by construction: every access to the label array that another
// thread may touch concurrently goes through lab_ld / lab_st / lab_cas below
// (relaxed, agent scope).  The only plain label store is in init, where no
// other thread reads labels.
//
// MI355X-specific structure (64-lane wavefronts):
//   * compute1 processes degree <= thread_max at one lane per vertex and, in
//     the same pass, records medium-degree vertices as one 64-bit ballot word
//     per wavefront (a bitmap worklist: no atomics, no ordering), and pushes
//     high-degree vertices with one wave-aggregated 64-bit atomic that
//     reserves both list slots and edge-task numbers.
//   * compute2 gives each medium-degree vertex one whole wavefront.
//   * compute3 splits every high-degree adjacency list into hub_chunk-edge
//     tasks, so one hub is spread over many blocks; each task is read with
//     coalesced loads, filtered (only neighbours < v are hooked) and
//     compacted into LDS, then hooked with all lanes busy.
//   * Batches: the same kernels run over packed graphs; each 64-lane
//     wavefront resolves its graph once from a chunk->graph table.

#include <climits>

#include "eclcc_internal.h"

namespace eclcc {

// ---- label access: relaxed, agent-scope atomics ---------------------------

__device__ __forceinline__ int lab_ld(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lab_st(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Returns the value found at *p (== expect iff the swap happened).
__device__ __forceinline__ int lab_cas(int* p, int expect, int desired) {
  __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expect;
}

// ---- wavefront helpers ------------------------------------------------------

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// number of set bits of `mask` in lanes below this one
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// ---- graph resolution -------------------------------------------------------

struct GraphRef {
  const int* nidx;
  const int* nlist;
  int* nstat;
  int n;
  long long slot0;
};

template <bool BATCH>
__device__ __forceinline__ GraphRef resolve(const KParams& p, long long slot) {
  if constexpr (!BATCH) {
    return GraphRef{p.nidx, p.nlist, p.nstat, p.n, 0};
  } else {
    const int g = p.chunk_graph[slot >> 8];
    const GraphDesc d = p.desc[g];
    return GraphRef{p.nidx + d.ibase, p.nlist + d.ebase, p.nstat + d.vbase, d.n, d.vbase};
  }
}

// first slot of this thread's wavefront, made wave-uniform
__device__ __forceinline__ long long wave_slot() {
  const int w = uniform((int)(threadIdx.x >> 6));
  return ((long long)blockIdx.x * blockDim.x) + (long long)w * 64;
}

// ---- find and hook ----------------------------------------------------------

// Root of idx given curr = label[idx] already loaded; on the way every
// visited vertex is re-pointed to its grandparent ("intermediate pointer
// jumping", ECL-CC-atomic.cu:92-104).  Writes only move a pointer to an
// ancestor, i.e. to a smaller ID.
__device__ __forceinline__ int find_from(int idx, int curr, int* nstat) {
  if (curr != idx) {
    int prev = idx, next;
    while (curr > (next = lab_ld(&nstat[curr]))) {
      lab_st(&nstat[prev], next);
      prev = curr;
      curr = next;
    }
  }
  return curr;
}

__device__ __forceinline__ int find_root(int idx, int* nstat) {
  return find_from(idx, lab_ld(&nstat[idx]), nstat);
}

// Min-linking of two root estimates (ECL-CC-atomic.cu:134-151): the larger
// root is swung to the smaller by CAS; a failed CAS returns the value that
// is there now, which is strictly smaller, and the walk continues from it.
// Returns the caller's updated root estimate.
__device__ __forceinline__ int hook(int vstat, int ostat, int* nstat) {
  while (vstat != ostat) {
    if (vstat < ostat) {
      const int seen = lab_cas(&nstat[ostat], ostat, vstat);
      if (seen == ostat) break;
      ostat = seen;
    } else {
      const int seen = lab_cas(&nstat[vstat], vstat, ostat);
      if (seen == vstat) {
        vstat = ostat;  // v's tree now hangs below ostat
        break;
      }
      vstat = seen;
    }
  }
  return vstat;
}

// Hooks the edge v-u (caller checked v > u); s = label[u] already loaded.
// When u's parent already is v's root estimate the two are joined and
// nothing is written.
__device__ __forceinline__ int hook_edge(int vstat, int u, int s, int* nstat) {
  if (s == vstat) return vstat;
  return hook(vstat, find_from(u, s, nstat), nstat);
}

// ---- init (ECL-CC-atomic.cu:69-88) --------------------------------------------
// label[v] = first neighbour smaller than v in list order, else v.  Each lane
// checks the first four entries of its list; lists that need a longer scan
// are finished cooperatively by the whole wavefront, 64 entries per step, the
// first smaller entry being the lowest set lane of a ballot.

template <bool BATCH>
__global__ __launch_bounds__(256) void k_init(KParams p) {
  const long long ws = wave_slot();
  if (ws >= p.nslots) return;
  const GraphRef G = resolve<BATCH>(p, ws);
  const int lane = lane_id();
  const long long slot = ws + lane;
  const int v = (int)(slot - G.slot0);
  const bool live = slot < p.nslots && v < G.n;
  int m = v, beg = 0, end = 0;
  bool longscan = false;
  if (live) {
    beg = G.nidx[v];
    end = G.nidx[v + 1];
    const int x0 = beg + 0 < end ? G.nlist[beg + 0] : INT_MAX;
    const int x1 = beg + 1 < end ? G.nlist[beg + 1] : INT_MAX;
    const int x2 = beg + 2 < end ? G.nlist[beg + 2] : INT_MAX;
    const int x3 = beg + 3 < end ? G.nlist[beg + 3] : INT_MAX;
    m = x0 < v ? x0 : x1 < v ? x1 : x2 < v ? x2 : x3 < v ? x3 : v;
    longscan = (m == v) && (end - beg > 4);
  }
  unsigned long long pending = __ballot(longscan);
  while (pending) {
    const int src = __ffsll((long long)pending) - 1;
    pending &= pending - 1;
    const int sv = __shfl(v, src);
    const int sb = __shfl(beg, src) + 4;
    const int se = __shfl(end, src);
    int found = sv;
    for (int base = sb; base < se; base += 64) {
      const int i = base + lane;
      const int x = i < se ? G.nlist[i] : INT_MAX;
      const unsigned long long hit = __ballot(x < sv);
      if (hit) {
        found = __shfl(x, __ffsll((long long)hit) - 1);
        break;
      }
    }
    if (lane == src) m = found;
  }
  if (live) G.nstat[v] = m;  // no other thread reads labels during init
  if (slot == 0) {
    *p.hubctr = 0ull;
    *p.flags = 0u;
  }
}

// ---- compute1 (ECL-CC-atomic.cu:108-157) --------------------------------------

// Wave-aggregated push of high-degree vertices: one 64-bit atomic reserves
// list slots (low half) and edge-task numbers (high half) for the whole wave.
__device__ __forceinline__ void push_hubs(const KParams& p, bool hub, long long slot, int deg,
                                          unsigned long long mask) {
  const int lane = lane_id();
  const unsigned tasks = hub ? (unsigned)((deg + p.hub_chunk - 1) / p.hub_chunk) : 0u;
  unsigned incl = tasks;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned t = __shfl_up(incl, d);
    if (lane >= d) incl += t;
  }
  const unsigned total = __shfl(incl, 63);
  unsigned lo = 0, hi = 0;
  if (lane == 0) {
    const unsigned long long add =
        (unsigned long long)__popcll(mask) | ((unsigned long long)total << 32);
    const unsigned long long old =
        __hip_atomic_fetch_add(p.hubctr, add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    lo = (unsigned)old;
    hi = (unsigned)(old >> 32);
  }
  lo = __shfl(lo, 0);
  hi = __shfl(hi, 0);
  if (hub) {
    const long long idx = (long long)lo + lanes_below(mask);
    if (idx < p.hub_cap) {
      p.hubs[idx] = HubEntry{slot, (long long)hi + (long long)(incl - tasks)};
    } else {
      __hip_atomic_fetch_or(p.flags, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

template <bool BATCH>
__global__ __launch_bounds__(256) void k_hook_thread(KParams p) {
  const long long ws = wave_slot();
  if (ws >= p.nslots) return;
  const GraphRef G = resolve<BATCH>(p, ws);
  const int lane = lane_id();
  const long long slot = ws + lane;
  const int v = (int)(slot - G.slot0);
  const bool live = slot < p.nslots && v < G.n;
  bool active = false;
  int beg = 0, end = 0;
  if (live) {
    active = lab_ld(&G.nstat[v]) != v;  // roots after init have no smaller neighbour
    if (active) {
      beg = G.nidx[v];
      end = G.nidx[v + 1];
    }
  }
  const int deg = end - beg;
  const bool mid = active && deg > p.thread_max && deg <= p.wave_max;
  const bool hub = active && deg > p.wave_max;
  const unsigned long long midmask = __ballot(mid);
  if (lane == 0) p.midbits[ws >> 6] = midmask;
  const unsigned long long hubmask = __ballot(hub);
  if (hubmask) push_hubs(p, hub, slot, deg, hubmask);

  if (active && deg <= p.thread_max) {
    int* const ns = G.nstat;
    const int* const nl = G.nlist;
    int vstat = find_root(v, ns);
    // four neighbours per step: their label loads are issued together
    for (int i = beg; i < end; i += 4) {
      const int u0 = nl[i];
      const int u1 = i + 1 < end ? nl[i + 1] : INT_MAX;
      const int u2 = i + 2 < end ? nl[i + 2] : INT_MAX;
      const int u3 = i + 3 < end ? nl[i + 3] : INT_MAX;
      const int s0 = v > u0 ? lab_ld(&ns[u0]) : 0;
      const int s1 = v > u1 ? lab_ld(&ns[u1]) : 0;
      const int s2 = v > u2 ? lab_ld(&ns[u2]) : 0;
      const int s3 = v > u3 ? lab_ld(&ns[u3]) : 0;
      if (v > u0) vstat = hook_edge(vstat, u0, s0, ns);
      if (v > u1) vstat = hook_edge(vstat, u1, s1, ns);
      if (v > u2) vstat = hook_edge(vstat, u2, s2, ns);
      if (v > u3) vstat = hook_edge(vstat, u3, s3, ns);
    }
  }
}

// ---- compute2 (ECL-CC-atomic.cu:161-199) --------------------------------------
// One wavefront per medium-degree vertex.  Wavefront w owns bitmap word w,
// i.e. the 64 slots compute1's wavefront w looked at.

template <bool BATCH>
__global__ __launch_bounds__(256) void k_hook_wave(KParams p) {
  const long long ws = wave_slot();
  if (ws >= p.nslots) return;
  unsigned long long bits = p.midbits[ws >> 6];
  if (!bits) return;
  const GraphRef G = resolve<BATCH>(p, ws);
  const int lane = lane_id();
  int* const ns = G.nstat;
  const int* const nl = G.nlist;
  while (bits) {
    const int b = __ffsll((long long)bits) - 1;
    bits &= bits - 1;
    const int v = (int)(ws + b - G.slot0);
    const int beg = G.nidx[v];
    const int end = G.nidx[v + 1];
    int vstat = find_root(v, ns);
    int i = beg + lane;
    for (; i + 64 < end; i += 128) {
      const int u0 = nl[i];
      const int u1 = nl[i + 64];
      const int s0 = v > u0 ? lab_ld(&ns[u0]) : 0;
      const int s1 = v > u1 ? lab_ld(&ns[u1]) : 0;
      if (v > u0) vstat = hook_edge(vstat, u0, s0, ns);
      if (v > u1) vstat = hook_edge(vstat, u1, s1, ns);
    }
    if (i < end) {
      const int u = nl[i];
      if (v > u) vstat = hook_edge(vstat, u, lab_ld(&ns[u]), ns);
    }
  }
}

// ---- compute3 (ECL-CC-atomic.cu:203-246) --------------------------------------
// Blocks stride over the edge tasks of all high-degree vertices.  A task is
// hub_chunk consecutive entries of one adjacency list: read coalesced,
// neighbours < v compacted into LDS, then hooked by every thread.

template <bool BATCH, int BT>
__global__ __launch_bounds__(BT) void k_hook_block(KParams p) {
  __shared__ int stage[kMaxHubChunk];
  __shared__ int s_cnt;
  const unsigned long long ctr = *p.hubctr;
  const long long pushed = (long long)(unsigned)ctr;
  const long long nhubs = pushed < p.hub_cap ? pushed : p.hub_cap;
  const long long ntasks = (long long)(ctr >> 32);
  const int tid = threadIdx.x;
  const int lane = lane_id();
  for (long long t = blockIdx.x; t < ntasks; t += gridDim.x) {
    // last hub whose first task is <= t
    long long lo = 0, hi = nhubs - 1;
    while (lo < hi) {
      const long long mid = (lo + hi + 1) >> 1;
      if (p.hubs[mid].first_task <= t) lo = mid;
      else hi = mid - 1;
    }
    const HubEntry h = p.hubs[lo];
    const GraphRef G = resolve<BATCH>(p, h.slot);
    const int v = (int)(h.slot - G.slot0);
    const long long cb = (long long)G.nidx[v] + (t - h.first_task) * (long long)p.hub_chunk;
    const long long ce_full = G.nidx[v + 1];
    const int cbeg = (int)cb;
    const int cend = (int)(cb + p.hub_chunk < ce_full ? cb + p.hub_chunk : ce_full);
    if (tid == 0) s_cnt = 0;
    __syncthreads();
    for (int base = cbeg; base < cend; base += BT) {
      const int i = base + tid;
      const int u = i < cend ? G.nlist[i] : INT_MAX;
      const bool keep = u < v;
      const unsigned long long mask = __ballot(keep);
      if (mask) {
        int off = 0;
        if (lane == 0) off = atomicAdd(&s_cnt, __popcll(mask));
        off = __shfl(off, 0);
        if (keep) stage[off + lanes_below(mask)] = u;
      }
    }
    __syncthreads();
    const int cnt = s_cnt;
    if (cnt > 0) {
      int* const ns = G.nstat;
      int vstat = find_root(v, ns);
      int j = tid;
      for (; j + BT < cnt; j += 2 * BT) {
        const int u0 = stage[j];
        const int u1 = stage[j + BT];
        const int s0 = lab_ld(&ns[u0]);
        const int s1 = lab_ld(&ns[u1]);
        vstat = hook_edge(vstat, u0, s0, ns);
        vstat = hook_edge(vstat, u1, s1, ns);
      }
      if (j < cnt) {
        const int u = stage[j];
        vstat = hook_edge(vstat, u, lab_ld(&ns[u]), ns);
      }
    }
    __syncthreads();
  }
}

// ---- flatten (ECL-CC-atomic.cu:250-264) ---------------------------------------

template <bool BATCH>
__global__ __launch_bounds__(256) void k_flatten(KParams p) {
  const long long ws = wave_slot();
  if (ws >= p.nslots) return;
  const GraphRef G = resolve<BATCH>(p, ws);
  const long long slot = ws + lane_id();
  const int v = (int)(slot - G.slot0);
  if (slot < p.nslots && v < G.n) {
    int* const ns = G.nstat;
    int s = lab_ld(&ns[v]);
    const int old = s;
    int next;
    while (s > (next = lab_ld(&ns[s]))) s = next;
    if (s != old) lab_st(&ns[v], s);
  }
}

// ---- host-side launch -----------------------------------------------------------

template <bool BATCH>
static hipError_t launch_all(const KParams& p, const Shape& shape, hipStream_t s, hipEvent_t* ev) {
  const long long vb = (p.nslots + 255) / 256;
  if (vb <= 0) return hipSuccess;
  const dim3 grid((unsigned)vb), block(256);
  if (ev) (void)hipEventRecord(ev[0], s);
  k_init<BATCH><<<grid, block, 0, s>>>(p);
  if (ev) (void)hipEventRecord(ev[1], s);
  k_hook_thread<BATCH><<<grid, block, 0, s>>>(p);
  if (ev) (void)hipEventRecord(ev[2], s);
  k_hook_wave<BATCH><<<grid, block, 0, s>>>(p);
  if (ev) (void)hipEventRecord(ev[3], s);
  const dim3 hgrid((unsigned)shape.hub_blocks);
  switch (shape.block_threads) {
    case 1024: k_hook_block<BATCH, 1024><<<hgrid, dim3(1024), 0, s>>>(p); break;
    case 512: k_hook_block<BATCH, 512><<<hgrid, dim3(512), 0, s>>>(p); break;
    default: k_hook_block<BATCH, 256><<<hgrid, dim3(256), 0, s>>>(p); break;
  }
  if (ev) (void)hipEventRecord(ev[4], s);
  k_flatten<BATCH><<<grid, block, 0, s>>>(p);
  if (ev) (void)hipEventRecord(ev[5], s);
  return hipGetLastError();
}

hipError_t launch_pipeline(const KParams& p, bool batch, const Shape& shape, hipStream_t s,
                           hipEvent_t* ev) {
  return batch ? launch_all<true>(p, shape, s, ev) : launch_all<false>(p, shape, s, ev);
}

hipError_t launch_init_only(const KParams& p, hipStream_t s) {
  const long long vb = (p.nslots + 255) / 256;
  if (vb <= 0) return hipSuccess;
  k_init<false><<<dim3((unsigned)vb), dim3(256), 0, s>>>(p);
  return hipGetLastError();
}

}  // namespace eclcc
