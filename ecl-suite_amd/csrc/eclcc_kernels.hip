// eclcc_kernels.hip -- race-free ECL-CC kernels for gfx950 (MI355X).
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
// Kernel sequence (one launch each, no host synchronisation in between):
//   init      labels := first smaller neighbour (reference init); degree
//             binning into 64-bit ballot words (no atomics)
//   sample    one sampled hook per vertex
//   pick      per graph: the most frequent root among 256 sampled vertices
//   mark      bitmap of the vertices already joined to that root; final work
//             lists; high-degree counts per word
//   hubscan   exclusive prefix of the high-degree counts per 1024-word block;
//             the blocks' totals are summed where needed (compute1, compute3)
//   compute1  writes the high-degree task list; low degree: thread granularity
//   compute2  medium degree: wavefront granularity
//   compute3  high degree: block granularity, edge tasks staged through LDS
//   flatten   labels := root (reference flatten)
// With filter 2 hubscan is not launched: mark lists the few high-degree
// vertices left outside the largest tree with their edge tasks itself, and
// compute3 takes those tasks.  A batch runs the same sequence
// over the slots of all its graphs (BATCH = true): sample and mark then
// work per block list and per 256-slot chunk, with their blocks taken in a
// spread order over the batch (spread_block).
//
// Why the bitmap (MI355X measurements): random 4-byte gathers from the
// 64 MB label array of a 2^24-vertex graph run at ~58 G/s (memory side),
// from a 2 MB table that stays in L2 at ~265 G/s.  After init and the
// sampled hooks almost every remaining edge of random, grid and power-law
// graphs joins two vertices that are already in the largest tree; one bit
// per vertex (2 MB for 2^24) tells that without touching the labels.
// Skipping an edge whose ends are both connected to the sampled root is
// exact: they are in one component already.  Filter 1 reads every entry and
// skips such edges after a lookup; filter 2 does not read the lists of the
// tree's members at all: every edge with an end outside the tree is hooked
// from that end, which hooks all of its entries.

#include <climits>

#include "eclcc_internal.h"

namespace eclcc {

// ---- label access: relaxed, agent-scope atomics ---------------------------
// ECLCC_PLAIN_LABELS builds a research copy (lib/libeclcc_plain.so) whose
// label loads and stores are plain memory accesses, as in the reference's
// racy baseline (src/baseline/egr-input/CC/ECL-CC_11.cu:81-93): it measures
// what race freedom costs on this GPU.  That copy has data races; nothing
// in the product loads it.  The CAS is atomic in both.

#ifndef ECLCC_PLAIN_LABELS
__device__ __forceinline__ int lab_ld(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lab_st(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
#else
__device__ __forceinline__ int lab_ld(const int* p) { return *p; }
__device__ __forceinline__ void lab_st(int* p, int v) { *p = v; }
#endif
// Returns the value found at *p (== expect iff the swap happened).
__device__ __forceinline__ int lab_cas(int* p, int expect, int desired) {
  __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expect;
}

// adjacency entries are streamed once: keep them from evicting the bitmap
__device__ __forceinline__ int ld_stream(const int* p) { return __builtin_nontemporal_load(p); }

// ---- wavefront helpers ------------------------------------------------------

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// number of set bits of `mask` in lanes below this one
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// inclusive prefix sum across the wavefront
__device__ __forceinline__ unsigned wave_incl_sum(unsigned x) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned t = __shfl_up(x, d);
    if (lane >= d) x += t;
  }
  return x;
}

// Loads the work-list words w0 .. w0 + span - 1 (span <= 64) of a block
// at once, by its first wavefront, into s_bits, and places each word's
// vertices in the block's LDS list by a prefix sum of their counts (s_off;
// s_cnt = the total).  One global load latency for the whole block instead
// of one per word.  The caller synchronises the block afterwards.
__device__ __forceinline__ unsigned long long gather_words(const unsigned long long* bits,
                                                           long long w0, int span, long long words,
                                                           unsigned long long* s_bits, int* s_off,
                                                           int* s_cnt) {
  const int i = threadIdx.x;
  const long long w = w0 + i;
  const unsigned long long b = (i < span && w < words) ? bits[w] : 0ull;
  const unsigned c = (unsigned)__popcll(b);
  const unsigned incl = wave_incl_sum(c);
  if (i < span) {
    s_bits[i] = b;
    s_off[i] = (int)(incl - c);
  }
  if (i == 63) *s_cnt = (int)incl;
  return b;
}

// Writes the vertices of the gathered words into the LDS list, one word per
// wavefront at a time (after gather_words and a block barrier).
__device__ __forceinline__ void list_words(const unsigned long long* s_bits, const int* s_off,
                                           int span, unsigned short* s_list) {
  const int lane = lane_id();
  for (int i = (int)(threadIdx.x >> 6); i < span; i += (int)(blockDim.x >> 6)) {
    const unsigned long long b = s_bits[i];
    if ((b >> lane) & 1ull) s_list[s_off[i] + lanes_below(b)] = (unsigned short)(i * 64 + lane);
  }
}

// bit `slot` of a per-slot bitmap
__device__ __forceinline__ unsigned slot_bit(const unsigned long long* bits, long long slot) {
  const unsigned* w = reinterpret_cast<const unsigned*>(bits);
  return (w[slot >> 5] >> (slot & 31)) & 1u;
}

// ---- graph resolution -------------------------------------------------------

struct GraphRef {
  const int* nidx;
  const int* nlist;
  int* nstat;
  int n;
  int g;
  long long slot0;
};

template <bool BATCH>
__device__ __forceinline__ GraphRef graph_ref(const KParams& p, int g) {
  if constexpr (!BATCH) {
    return GraphRef{p.nidx, p.nlist, p.nstat, p.n, 0, 0};
  } else {
    const GraphDesc d = p.desc[g];
    return GraphRef{d.nidx, d.nlist, d.nstat, d.n, g, d.vbase};
  }
}

template <bool BATCH>
__device__ __forceinline__ GraphRef resolve(const KParams& p, long long slot) {
  if constexpr (!BATCH) return graph_ref<false>(p, 0);
  else return graph_ref<true>(p, p.chunk_graph[slot >> 8]);
}

// Block order of the batched walk kernels: a permutation of the work items
// that spreads the blocks resident at one time over many graphs rather
// than a few consecutive ones, whose top-of-tree labels and roots every
// walk of the graph reaches (1024-graph batch of the ER / R-MAT mix: sample
// 5.79 -> 5.49 ms; grouping each graph's blocks on one XCD instead: 6.02
// ms).  The items may be cut into windows of 2^wlog consecutive items,
// each permuted on its own (item = window base + (b * wstride) mod 2^wlog,
// wstride odd) and taken one after the other; the last, partial window
// (or the whole grid, with wlog 32) takes (b * tstride) mod its size,
// tstride coprime with that size (BlockOrder, order_item).  A window whose
// labels fit the 256 MB Infinity Cache keeps mark's walks there (see
// launch_all).
__device__ __forceinline__ unsigned spread_block(const BlockOrder o) {
  return order_item(o, blockIdx.x, gridDim.x);
}

// first slot of this thread's wavefront, made wave-uniform
__device__ __forceinline__ long long wave_slot() {
  const int w = uniform((int)(threadIdx.x >> 6));
  return ((long long)blockIdx.x * blockDim.x) + (long long)w * 64;
}

// ---- find and hook ----------------------------------------------------------

// Vertex 0 is a root for good: no ID is smaller, so no CAS can swing it and
// no pointer jump writes to a root.  Walks stop there without reading its
// label, which every thread of a graph whose largest component holds vertex
// 0 would otherwise read.

// Root of idx given curr = label[idx] already loaded.  Two ways of
// shortening the path, chosen by KParams::path_mode:
//   0  intermediate pointer jumping, as the reference's representative
//      (ECL-CC-atomic.cu:92-104): every visited vertex is re-pointed to its
//      grandparent, one label store per step;
//   1  root jump: the walk only reads, and the start vertex is re-pointed to
//      the root at the end, at most one store per walk.  On MI355X every
//      agent-scope store is written through to the memory side; on random
//      and grid graphs the stores of mode 0 cost more than the longer paths
//      they save (measured: ER n=1M 1.33 -> 0.37 ms, 3-D grid 256^3 2.64 ->
//      1.69 ms per pass).
// Either way a store only moves a pointer to an ancestor, i.e. to a smaller
// ID, and the labels that come out are the same.
__device__ __forceinline__ int find_from(int idx, int curr, int* nstat, int mode) {
  if (curr != idx) {
    if (mode == 0) {
      int prev = idx, next;
      while (curr != 0 && curr > (next = lab_ld(&nstat[curr]))) {
        lab_st(&nstat[prev], next);
        prev = curr;
        curr = next;
      }
    } else {
      const int first = curr;
      int next;
      while (curr != 0 && curr > (next = lab_ld(&nstat[curr]))) curr = next;
      if (curr != first) lab_st(&nstat[idx], curr);
    }
  }
  return curr;
}

__device__ __forceinline__ int find_root(int idx, int* nstat, int mode) {
  return find_from(idx, lab_ld(&nstat[idx]), nstat, mode);
}

// Like find_from, but stops as soon as the walk reaches `stop` (a known
// ancestor of the caller's vertex, usually the root of the largest tree):
// the two are then joined, and the label of `stop` -- read by every thread
// otherwise -- is not touched.
__device__ __forceinline__ int find_until(int idx, int curr, int stop, int* nstat, int mode) {
  if (curr != idx) {
    if (mode == 0) {
      int prev = idx, next;
      while (curr != stop && curr != 0 && curr > (next = lab_ld(&nstat[curr]))) {
        lab_st(&nstat[prev], next);
        prev = curr;
        curr = next;
      }
    } else {
      const int first = curr;
      int next;
      while (curr != stop && curr != 0 && curr > (next = lab_ld(&nstat[curr]))) curr = next;
      if (curr != first) lab_st(&nstat[idx], curr);
    }
  }
  return curr;
}

// Min-linking of two root estimates (ECL-CC-atomic.cu:134-151): the larger
// root is swung to the smaller by CAS; a failed CAS returns the value that
// is there now, which is strictly smaller, and the walk continues from it.
// Each CAS is preceded by a plain relaxed load of the same label: on MI355X
// device-scope atomics execute at the memory side and serialise per address,
// and when many lanes hold the same just-merged root (dense random graphs,
// grids) most of them learn from the load that it has moved and never issue
// the atomic.  Returns the caller's updated root estimate.
__device__ __forceinline__ int hook(int vstat, int ostat, int* nstat) {
  while (vstat != ostat) {
    if (vstat < ostat) {
      const int cur = lab_ld(&nstat[ostat]);
      if (cur != ostat) {  // already swung: follow it without an atomic
        ostat = cur;
        continue;
      }
      const int seen = lab_cas(&nstat[ostat], ostat, vstat);
      if (seen == ostat) break;
      ostat = seen;
    } else {
      const int cur = lab_ld(&nstat[vstat]);
      if (cur != vstat) {
        vstat = cur;
        continue;
      }
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

// Joins the trees of two vertices whose labels are a and b by walking both
// paths at once: the larger of the two current vertices always takes the
// step (every label points to a smaller ID), so the walks stop where the
// paths meet, or where the larger one is a root, which is then swung below
// the other by CAS (a failed CAS hands back the new, smaller label and the
// walk goes on from there).  The top of the largest tree, where all paths
// end and whose labels every thread would read, is rarely reached.  Returns
// the meeting point: a common ancestor of both vertices after the call.
__device__ __forceinline__ int meet_link(int a, int b, int* nstat) {
  while (a != b) {
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int na = lab_ld(&nstat[a]);
    if (na != a) {
      a = na;
      continue;
    }
    const int seen = lab_cas(&nstat[a], a, b);
    if (seen == a) break;
    a = seen;
  }
  return b;
}

// The same meeting walk, halving the stepping side's path as it goes
// (path_mode 2): when the larger vertex a is not a root and its parent na
// lies above b, the walk also reads na's parent nna and, if na is not a root
// either, re-points a to nna and jumps there (one store, two levels).  The
// store moves a to an ancestor, as every shortening store does.  Sample's
// links chain roots below other walks' meeting points, and the walks that
// come after climb those chains; halving them as they are climbed left
// 2.1-3.2 label loads per sampled vertex instead of 3.4-5.1, and 0.6-0.7
// instead of 1.0 per vertex in mark (serial model of R-MAT-20/22 and grid
// 128^3 in random vertex order).  When na is a root above b it is linked
// below b at once.
__device__ __forceinline__ int meet_link_halve(int a, int b, int* nstat) {
  while (a != b) {
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int na = lab_ld(&nstat[a]);
    if (na == a) {
      const int seen = lab_cas(&nstat[a], a, b);
      if (seen == a) break;
      a = seen;
      continue;
    }
    if (na <= b) {  // b is na, or b is now the larger: step once
      a = na;
      continue;
    }
    const int nna = lab_ld(&nstat[na]);
    if (nna != na) {
      lab_st(&nstat[a], nna);
      a = nna;
      continue;
    }
    const int seen = lab_cas(&nstat[na], na, b);
    if (seen == na) break;  // a's root now hangs below b
    a = seen;
  }
  return b;
}

// Hooks the edge v-u (caller checked v > u); s = label[u] already loaded.
// When u's path reaches v's root estimate the two are joined and nothing is
// written.
__device__ __forceinline__ int hook_edge(int vstat, int u, int s, int* nstat, int mode) {
  if (s == vstat) return vstat;
  const int r = find_until(u, s, vstat, nstat, mode);
  if (r == vstat) return vstat;
  return hook(vstat, r, nstat);
}

// ---- init (ECL-CC-atomic.cu:69-88) and degree binning ------------------------
// label[v] = first neighbour smaller than v in list order, else v.  Each lane
// checks the first two entries of its list; the lanes that need more scan
// on together, eight entries of their own list per step (in R-MAT-24 the
// vertices with no smaller neighbour, which must read their whole list,
// crowd into the low IDs: some wavefronts hold dozens of them); lists
// still unfinished after kLaneScan entries are finished cooperatively by
// the whole wavefront, 64 entries per step, the first smaller entry being
// the lowest set lane of a ballot.
// With filter 2 the scan past the first two entries is left out (R-MAT-24
// init 0.134 -> 0.118 ms, grid 256^3 0.116 -> 0.099 ms): a vertex whose
// smaller neighbours lie further on keeps its own ID as label, which is a
// valid forest (every label points to an ancestor), and its entries are
// hooked later like any other vertex's.  Its label is then not the
// reference's init value, and the labels that come out are still the
// component minima.  In sorted lists (the usual case) nothing is lost: a
// list whose first entry is larger than its vertex holds no smaller one.
// The same pass bins the vertices by degree (the reference does this in
// compute1, :116-127) into 64-bit ballot words per wavefront: vertices with
// a smaller neighbour, vertices with a sampled hook, and low / medium / high
// degree (no shared worklist: one contended device-scope atomic per push
// costs ~10 ns, serialised, on MI355X).  It also keeps the entry each vertex
// hooks in sample, so that nobody has to fetch the head of every list a
// second time.

// (10, 26 or 58 entries per lane time the same on R-MAT-24 and the grid:
// few lists get that far)
constexpr int kLaneScan = 2 + 3 * 8;

template <bool BATCH>
__global__ __launch_bounds__(256) void k_init(KParams p) {
  const long long ws = wave_slot();
  if (ws >= p.nslots) return;
  const GraphRef G = resolve<BATCH>(p, ws);
  const int lane = lane_id();
  const long long slot = ws + lane;
  const int v = (int)(slot - G.slot0);
  const bool live = slot < p.nslots && v < G.n;
  int m = v, beg = 0, end = 0, x0 = INT_MAX, x1 = INT_MAX;
  if (live) {
    beg = G.nidx[v];
    end = G.nidx[v + 1];
    x0 = end - beg > 0 ? G.nlist[beg + 0] : INT_MAX;
    x1 = end - beg > 1 ? G.nlist[beg + 1] : INT_MAX;
    m = x0 < v ? x0 : x1 < v ? x1 : v;
  }
  // list positions are unsigned: a list may end at entry 2^31 - 1, and
  // the positions one step past it must not wrap
  const unsigned uend = (unsigned)end;
  unsigned pos = (unsigned)beg + 2u;
  // filters 0 and 1 bin the vertices by the smaller neighbour found here,
  // so they need the whole scan; filter 2 bins every vertex with a list
  const bool full = p.filter != 2;
  bool scan = full && live && m == v && pos < uend;
  while (__ballot(scan)) {
    if (scan) {
      int x[8];
#pragma unroll
      for (int k = 0; k < 8; k++) x[k] = pos + k < uend ? G.nlist[pos + k] : INT_MAX;
#pragma unroll
      for (int k = 0; k < 8; k++) m = (m == v && x[k] < v) ? x[k] : m;
      pos += 8;
      scan = m == v && pos < uend && pos < (unsigned)beg + kLaneScan;
    }
  }
  const bool longscan = full && live && m == v && pos < uend;
  unsigned long long pending = __ballot(longscan);
  while (pending) {
    const int src = __ffsll((long long)pending) - 1;
    pending &= pending - 1;
    const int sv = __shfl(v, src);
    const unsigned sb = __shfl(pos, src);
    const unsigned se = __shfl(uend, src);
    int found = sv;
    for (unsigned base = sb; base < se; base += 64) {
      const unsigned i = base + lane;
      const int x = i < se ? G.nlist[i] : INT_MAX;
      const unsigned long long hit = __ballot(x < sv);
      if (hit) {
        found = __shfl(x, __ffsll((long long)hit) - 1);
        break;
      }
    }
    if (lane == src) m = found;
  }
  // the sampled hook: a vertex with a smaller neighbour takes its second
  // entry if that is smaller too (the first smaller one is its label); a
  // vertex without one takes its first entry, which nothing else would
  // join to it before compute1
  const bool active = live && m != v;
  const int hooked = active ? x1 : x0;
  // with sample_shift > 0 only 1 in 2^sample_shift of the vertices with a
  // smaller neighbour (by a multiplicative hash of the ID) take theirs
  // and only those with at least sample_min_deg entries
  const bool pick = (p.sample_shift == 0 ||
                     ((unsigned)v * 0x9E3779B1u) >> (32 - p.sample_shift) == 0u) &&
                    end - beg >= p.sample_min_deg;
  const bool sampled = active ? (x1 < v && pick) : (live && x0 != INT_MAX && x0 != v);
  if (live) G.nstat[v] = m;  // no other thread reads labels during init
  // degree bins: of the vertices that have a smaller neighbour (they hook
  // their entries below them), and with filter 2 of every vertex that has
  // a neighbour (mark then keeps those outside the largest tree)
  const bool binned = p.filter == 2 ? live && end > beg : active;
  const int deg = binned ? end - beg : 0;
  const bool low = binned && deg <= p.thread_max;
  const bool mid = deg > p.thread_max && deg <= p.wave_max;
  const bool hub = deg > p.wave_max;
  // under filter 2 actbits serves one purpose: the work lists of a graph
  // without a sampled largest tree, whose vertices hook the entries below
  // them; a vertex not scanned past its first two entries may have such
  // entries and is listed too (it hooks nothing if it has none)
  const bool maybe = !full && live && m == v && pos < uend;
  const unsigned long long actmask = __ballot(active || maybe);
  const unsigned long long smpmask = __ballot(sampled);
  // the sampled vertices' entries, packed at the front of the word's 64
  // entries by rank: about a quarter of the bytes written here and read by
  // sample (4096-graph mix: 23.97 -> 23.68 ms; R-MAT-24 init -3 us, grid
  // 256^3 0.932 -> 0.927 ms)
  if (sampled) p.second[ws + lanes_below(smpmask)] = hooked;
  const unsigned long long lowmask = __ballot(low);
  const unsigned long long midmask = __ballot(mid);
  const unsigned long long hubmask = __ballot(hub);
  if (lane == 0) {
    const long long w = ws >> 6;
    p.actbits[w] = actmask;
    p.smpbits[w] = smpmask;
    p.lowbits[w] = lowmask;
    p.midbits[w] = midmask;
    p.hubbits[w] = hubmask;
  }
  if (slot == 0) {
    *p.hubctr = 0ull;
    *p.flags = 0u;
  }
}

// ---- sample: sampled hooks (first part of the reference compute1) -------------
// One hook per vertex, chosen by init: a vertex with a smaller neighbour
// hooks its second list entry if that is smaller too (the first smaller one
// is already its label); a vertex without one hooks its first entry, since
// nothing before compute would join it to anything (R-MAT's degree-1
// vertices hanging from a larger hub are the common case).  With init's
// pointers this links random, grid and power-law graphs into one large
// tree, which mark then finds; with filter 2 the vertices left outside it
// are all compute has to do.  The sampled edge is hooked again later like
// every other edge where needed; re-hooking is a no-op.
// The two paths are walked together (meet_link) and a vertex with a smaller
// neighbour then points at the meeting point: at most one label store per
// vertex (a vertex without one is a root; the hook's CAS already moved it).
// Label stores are write-through on MI355X; a second store (for u) cost
// more than the shorter walks it saves later (measured on a 1024-graph
// batch: 13.8 -> 19.3 ms), and walking both paths to their roots
// (find_from + hook_edge) is slower still (21.6 ms).

// One graph: one lane per vertex.
__global__ __launch_bounds__(256) void k_sample_lane(KParams p) {
  const long long ws = wave_slot();
  if (ws >= p.nslots) return;
  const int lane = lane_id();
  int* const ns = p.nstat;
  const int v = (int)(ws + lane);
  // the lane's label is loaded together with the word that says whether it
  // hooks at all (one latency less; the lines are read anyway whenever a
  // lane of the wavefront hooks); its entry sits at its rank among the
  // word's sampled vertices
  const int vi = v < p.n ? v : 0;  // lanes past the end load vertex 0's
  const int s0 = lab_ld(&ns[vi]);
  const unsigned long long act = p.smpbits[ws >> 6];
  if (!((act >> lane) & 1ull)) return;
  const int u = p.second[ws + lanes_below(act)];
  const int su = lab_ld(&ns[u]);
  const int top = p.path_mode == 2 ? meet_link_halve(s0, su, ns) : meet_link(s0, su, ns);
  if (top < s0 && s0 != v) lab_st(&ns[v], top);
}

// Batches: a block gathers the sampled vertices of kSampleWords words into
// an LDS list and its threads work through it, one meeting walk per thread,
// so no lane idles on padding slots or inactive vertices and each thread
// looks up its graph once per vertex.  Measured on a 1024-graph batch of the
// ER / R-MAT mix: 9.7 ms with one lane per slot (lanes refilled within each
// 256-slot chunk), 5.6 ms dense.  On one large graph the per-lane kernel is
// faster (R-MAT-24: 0.20 vs 0.23 ms; 3-D grid 256^3: 0.67 vs 0.73 ms): a
// dense thread walks its vertices one after the other, and the last blocks
// finish late.  Blocks of fewer words are slower on batches (1024 graphs,
// sample with 8 / 16 / 32 words per block: +90 / +55 / +20 % against 64):
// the blocks that run together then cover fewer graphs (see spread_block).
// 128 words per block, their list prefix taken over two wavefronts, gained
// again (4096-graph mix: 8.01 -> 7.53 ms).  A small batch (the server's, of
// a few dozen graphs) takes fewer words per block, `span`, so that the
// launch still has enough blocks to fill the GPU: at 64 words a batch of
// one ER graph with 65,536 vertices had 16 blocks.
constexpr int kSampleWords = 128;

template <bool BATCH>
__global__ __launch_bounds__(256) void k_sample_dense(KParams p, BlockOrder order, int span) {
  static_assert(kSampleWords <= 128, "one word per lane of the first two wavefronts");
  constexpr int kChunks = kSampleWords * 64 / kSlotAlign;
  __shared__ unsigned short s_list[kSampleWords * 64];
  __shared__ unsigned long long s_bits[kSampleWords];
  __shared__ int s_off[kSampleWords];
  __shared__ unsigned s_wtot[2];
  __shared__ int* s_ns[kChunks];
  __shared__ long long s_s0[kChunks];
  const long long words = (p.nslots + 63) / 64;
  const long long w0 = (long long)spread_block(order) * span;
  const int tid = (int)threadIdx.x;
  if (tid < 128) {
    // the block's words, one per thread of the first two wavefronts, and
    // each word's place in the block's list (a prefix sum of their counts)
    const long long w = w0 + tid;
    const unsigned long long b = (tid < span && w < words) ? p.smpbits[w] : 0ull;
    const unsigned c = (unsigned)__popcll(b);
    const unsigned incl = wave_incl_sum(c);
    if (tid < span) {
      s_bits[tid] = b;
      s_off[tid] = (int)(incl - c);
    }
    if (lane_id() == 63) s_wtot[tid >> 6] = incl;
  } else if (tid < 128 + kChunks) {
    // the graph of each 256-slot chunk of the block, resolved once here
    // instead of per vertex (two dependent loads: chunk map, then the graph
    // table; 4096-graph mix: 8.04 -> 8.01 ms, the loads mostly hit L2).
    // The block's slots start at a chunk boundary when span >= 4; a
    // smaller span keeps the block inside one chunk (c = 0 below).
    const int c = tid - 128;
    const long long slot = (w0 << 6) + (long long)c * kSlotAlign;
    if (c * kSlotAlign < span * 64 && slot < p.nslots) {
      const GraphRef G = resolve<BATCH>(p, slot);
      s_ns[c] = G.nstat;
      s_s0[c] = G.slot0;
    }
  }
  __syncthreads();
  if (tid >= 64 && tid < 128 && tid < span) s_off[tid] += (int)s_wtot[0];
  const int cnt = (int)(s_wtot[0] + (span > 64 ? s_wtot[1] : 0u));
  __syncthreads();
  list_words(s_bits, s_off, span, s_list);
  __syncthreads();
  for (int j = tid; j < cnt; j += 256) {
    const int e = s_list[j];
    const long long slot = (w0 << 6) + e;
    const int c = e / kSlotAlign;
    const int v = (int)(slot - s_s0[c]);
    int* const ns = s_ns[c];
    // the entry sits at the vertex's rank among its word's sampled ones
    const int u = p.second[(slot & ~63ll) + __popcll(s_bits[e >> 6] & ((1ull << (e & 63)) - 1ull))];
    const int s0 = lab_ld(&ns[v]);
    const int su = lab_ld(&ns[u]);
    const int top = p.path_mode == 2 ? meet_link_halve(s0, su, ns) : meet_link(s0, su, ns);
    if (top < s0 && s0 != v) lab_st(&ns[v], top);
  }
}

// ---- compute1 (rest of ECL-CC-atomic.cu:108-157) -------------------------------
// Each block first writes the high-degree task entries of its words, in slot
// order, at the place the scan gave them.  Then low degree (<= thread_max):
// one thread per vertex, eight list entries in flight per step.
//  * filter 0/1: the lane hooks the entries smaller than its vertex.  A
//    vertex marked as joined to the sampled root R takes R as its estimate
//    and skips every entry whose other end is marked too, after one bitmap
//    lookup.
//  * filter 2 in a graph with R: the work list holds only vertices outside
//    R's tree; the lane hooks every entry, and a neighbour inside the tree
//    is replaced by R (they are joined, so the hook is the same).  An edge
//    with both ends in the tree is hooked by nobody: its ends are already
//    in one tree.  An edge with one end outside is hooked from that end.

// Tasks of hub_chunk entries that cover a list of deg entries (unsigned, so
// that a list of nearly 2^31 entries does not wrap).
__device__ __forceinline__ unsigned task_count(int deg, int chunk) {
  return ((unsigned)deg + (unsigned)chunk - 1u) / (unsigned)chunk;
}

// Sum of blocktot[0, nb) over the threads of a block (packed hub and task
// counts add without carrying into each other).  All threads call it.
__device__ __forceinline__ unsigned long long sum_block_totals(const KParams& p, long long nb) {
  __shared__ unsigned long long s_part[16];
  unsigned long long x = 0;
  for (long long b = threadIdx.x; b < nb; b += blockDim.x) x += p.blocktot[b];
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d);
  __syncthreads();
  if (lane_id() == 0) s_part[threadIdx.x >> 6] = x;
  __syncthreads();
  unsigned long long t = 0;
  for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += s_part[k];
  return t;
}

// Writes the high-degree task entries of word w (wave-cooperative); `gp` is
// the count of hubs | tasks << 32 in all scan blocks before w's.
template <bool BATCH>
__device__ __forceinline__ void emit_hubs(const KParams& p, long long w, unsigned long long hb,
                                          unsigned long long gp) {
  const long long ws = w << 6;
  const GraphRef G = resolve<BATCH>(p, ws);
  const int lane = lane_id();
  const long long slot = ws + lane;
  const int v = (int)(slot - G.slot0);
  const unsigned long long base = gp + p.hubinfo[w];
  const bool hub = (hb >> lane) & 1ull;
  unsigned tasks = 0;
  int hbeg = 0, hend = 0;
  if (hub) {
    hbeg = G.nidx[v];
    hend = G.nidx[v + 1];
    tasks = task_count(hend - hbeg, p.hub_chunk);
  }
  const unsigned excl = wave_incl_sum(tasks) - tasks;
  if (hub) {
    const long long idx = (long long)(unsigned)base + lanes_below(hb);
    if (idx < p.hub_cap) {
      p.hubs[idx] = HubEntry{slot, (long long)(unsigned)(base >> 32) + excl, hbeg, hend};
    } else {
      __hip_atomic_fetch_or(p.flags, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Hooks the list of one low-degree vertex (see above).
template <bool BATCH>
__device__ __forceinline__ void hook_low(const KParams& p, long long slot) {
  const GraphRef G = resolve<BATCH>(p, slot);
  const int v = (int)(slot - G.slot0);
  int* const ns = G.nstat;
  const int* const nl = G.nlist;
  const int R = p.filter ? p.giant[G.g] : -1;
  const int beg = G.nidx[v];
  const int end = G.nidx[v + 1];
  if (p.filter == 2 && R >= 0) {
    // Every neighbour inside R's tree stands for R, so the edges into the
    // tree all join the same two trees: the first one is hooked, the rest
    // are no-ops and are dropped (they would each cost a label load and a
    // step of the lane's sequential hook chain).
    const int s = lab_ld(&ns[v]);
    int vstat = -1;  // found on first use
    bool joined = false;
    for (unsigned i = beg; i < (unsigned)end; i += 8) {
      int u[8];
#pragma unroll
      for (int k = 0; k < 8; k++) u[k] = i + k < (unsigned)end ? ld_stream(&nl[i + k]) : -1;
      bool need[8];
#pragma unroll
      for (int k = 0; k < 8; k++) need[k] = u[k] >= 0 && u[k] != v;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        if (need[k] && slot_bit(p.gbits, G.slot0 + u[k])) {
          u[k] = R;
          need[k] = !joined;
          joined = true;
        }
      }
      int sv[8];
#pragma unroll
      for (int k = 0; k < 8; k++) sv[k] = need[k] ? lab_ld(&ns[u[k]]) : 0;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        if (need[k]) {
          if (vstat < 0) vstat = find_from(v, s, ns, p.path_mode);
          vstat = hook_edge(vstat, u[k], sv[k], ns, p.path_mode);
        }
      }
    }
    return;
  }
  const bool marked = p.filter && slot_bit(p.gbits, slot);
  int vstat = marked ? R : -1;  // found on first use otherwise
  for (unsigned i = beg; i < (unsigned)end; i += 8) {
    int u[8];
#pragma unroll
    for (int k = 0; k < 8; k++) u[k] = i + k < (unsigned)end ? ld_stream(&nl[i + k]) : INT_MAX;
    bool need[8];
#pragma unroll
    for (int k = 0; k < 8; k++) need[k] = v > u[k];
    if (marked) {
      unsigned in[8];
#pragma unroll
      for (int k = 0; k < 8; k++) in[k] = need[k] ? slot_bit(p.gbits, G.slot0 + u[k]) : 0u;
#pragma unroll
      for (int k = 0; k < 8; k++) need[k] = need[k] && !in[k];
    }
    int sv[8];
#pragma unroll
    for (int k = 0; k < 8; k++) sv[k] = need[k] ? lab_ld(&ns[u[k]]) : 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (need[k]) {
        if (vstat < 0) vstat = find_root(v, ns, p.path_mode);
        vstat = hook_edge(vstat, u[k], sv[k], ns, p.path_mode);
      }
    }
  }
}

// A block covers `span` words (at most kLowWords).  Its wavefronts first
// write the high-degree task entries of those words, then gather the
// block's low-degree vertices into an LDS list, and the 256 threads work
// through the list together: every lane has a vertex, however sparse the
// work list is (with filter 2 a word holds a few vertices outside the
// largest tree; R-MAT-24: 0.19 -> 0.12 ms).  The span shrinks for small
// graphs so that there are still enough blocks to fill the GPU.
constexpr int kLowWords = 64;

template <bool BATCH>
__global__ __launch_bounds__(256) void k_hook_thread(KParams p, int span) {
  static_assert(kLowWords <= 64, "one word per lane of the first wavefront");
  __shared__ unsigned short s_list[kLowWords * 64];
  __shared__ unsigned long long s_bits[kLowWords], s_hub[kLowWords];
  __shared__ int s_off[kLowWords];
  __shared__ int s_cnt, s_hubs;
  const long long words = (p.nslots + 63) / 64;
  const long long w0 = (long long)blockIdx.x * span;
  const int wv = threadIdx.x >> 6;
  if (threadIdx.x < 64) {
    const int i = threadIdx.x;
    const unsigned long long hb = (i < span && w0 + i < words) ? p.hubbits[w0 + i] : 0ull;
    (void)gather_words(p.lowbits, w0, span, words, s_bits, s_off, &s_cnt);
    s_hub[i] = hb;
    const unsigned long long any = __ballot(hb != 0ull);
    if (i == 0) s_hubs = any != 0ull;
  }
  __syncthreads();
  // hub tasks to write? then the counts before this block's scan block
  // (span divides kScanBlock, so the block lies inside one)
  if (s_hubs) {
    const unsigned long long gp = sum_block_totals(p, w0 / kScanBlock);
    for (int i = wv; i < span && w0 + i < words; i += 4)
      if (s_hub[i]) emit_hubs<BATCH>(p, w0 + i, s_hub[i], gp);
  }
  list_words(s_bits, s_off, span, s_list);
  __syncthreads();
  const int cnt = s_cnt;
  for (int j = threadIdx.x; j < cnt; j += 256) hook_low<BATCH>(p, (w0 << 6) + s_list[j]);
}

// ---- high-degree task list: block prefix, grand prefix, emission -----------

// Block-wide exclusive scan of packed (count | tasks << 32) values, one per
// thread of a 1024-thread block; returns the exclusive prefix and sets *tot.
__device__ __forceinline__ unsigned long long block_excl_scan(unsigned long long x,
                                                              unsigned long long* tot) {
  __shared__ unsigned s_lo[16], s_hi[16];
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  const unsigned lo = wave_incl_sum((unsigned)x);
  const unsigned hi = wave_incl_sum((unsigned)(x >> 32));
  if (lane == 63) {
    s_lo[wv] = lo;
    s_hi[wv] = hi;
  }
  __syncthreads();
  unsigned plo = 0, phi = 0, tlo = 0, thi = 0;
  for (int k = 0; k < 16; k++) {
    if (k < wv) {
      plo += s_lo[k];
      phi += s_hi[k];
    }
    tlo += s_lo[k];
    thi += s_hi[k];
  }
  __syncthreads();
  *tot = (unsigned long long)tlo | ((unsigned long long)thi << 32);
  const unsigned elo = plo + lo - (unsigned)x;
  const unsigned ehi = phi + hi - (unsigned)(x >> 32);
  return (unsigned long long)elo | ((unsigned long long)ehi << 32);
}

__global__ __launch_bounds__(1024) void k_hub_scan(KParams p) {
  const long long words = (p.nslots + 63) / 64;
  const long long w = (long long)blockIdx.x * kScanBlock + threadIdx.x;
  const unsigned long long x = w < words ? p.hubinfo[w] : 0ull;
  unsigned long long tot;
  const unsigned long long ex = block_excl_scan(x, &tot);
  if (w < words) p.hubinfo[w] = ex;
  if (threadIdx.x == 0) p.blocktot[blockIdx.x] = tot;
}

// ---- pick: per graph, the most frequent root among sampled vertices -------

template <bool BATCH>
__global__ __launch_bounds__(kPickSamples) void k_pick(KParams p) {
  __shared__ int s_root[kPickSamples];
  __shared__ unsigned long long s_best[kPickSamples / 64];
  const int g = blockIdx.x;
  const GraphRef G = graph_ref<BATCH>(p, g);
  const int t = threadIdx.x;
  int r = -1;
  if (p.filter && G.n >= 2) {
    const int nsamp = G.n < kPickSamples ? G.n : kPickSamples;
    if (t < nsamp) {
      const int v = (int)(((long long)t * G.n) / nsamp);
      r = find_root(v, G.nstat, p.path_mode);
    }
  }
  s_root[t] = r;
  __syncthreads();
  unsigned c = 0;
  if (r >= 0)
    for (int k = 0; k < kPickSamples; k++) c += s_root[k] == r;
  // largest count, ties to the smaller root
  unsigned long long key = r >= 0 ? ((unsigned long long)c << 32) | (unsigned)(INT_MAX - r) : 0ull;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(key, d);
    key = o > key ? o : key;
  }
  if (lane_id() == 0) s_best[t >> 6] = key;
  __syncthreads();
  if (t == 0) {
    unsigned long long best = 0;
    for (int k = 0; k < kPickSamples / 64; k++) best = s_best[k] > best ? s_best[k] : best;
    const unsigned cnt = (unsigned)(best >> 32);
    p.giant[g] = cnt >= 2 ? INT_MAX - (int)(unsigned)best : -1;
  }
}

// ---- mark: membership bitmap and the final work lists ------------------------
// gbits := the vertices joined to the graph's sampled root R (exact: no hook
// runs during mark).  With filter 2 the work lists keep only the vertices
// outside R's tree (each then hooks all of its entries), or, in a graph
// without a sampled root, the vertices that have a smaller neighbour (each
// hooks the entries below it, as the reference does); the high-degree ones
// among them are few, and mark lists them with their edge tasks at once
// (list_hub) for compute3.  Otherwise mark counts each word's high-degree
// vertices and their edge tasks for the scan.

// Filter 2: appends a high-degree vertex to the hub list.  One 64-bit atomic
// add of (1 | tasks << 32) hands out its index and its first task number
// together, so the list is ordered by first task in whatever order the adds
// arrive, and no scan is needed.  Such vertices lie outside the largest tree
// and are rare (none on R-MAT-24, whose hubs all join it; about 40 tasks in
// 1024 graphs of the batch mix), so the counter is not contended.  Before,
// they joined compute2's chunk lists, where one wavefront dealt a whole hub
// list over its lanes: a vertex with a million neighbours outside the
// largest tree took 10.3 ms there.
__device__ __forceinline__ void list_hub(const KParams& p, const GraphRef& G, long long slot,
                                         int v) {
  const int beg = G.nidx[v];
  const int end = G.nidx[v + 1];
  const unsigned tasks = task_count(end - beg, p.hub_chunk);
  const unsigned long long old =
      __hip_atomic_fetch_add(p.hubctr, 1ull | ((unsigned long long)tasks << 32), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  const long long idx = (long long)(unsigned)old;
  if (idx < p.hub_cap) {
    p.hubs[idx] = HubEntry{slot, (long long)(old >> 32), beg, end};
  } else {
    __hip_atomic_fetch_or(p.flags, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// One graph: one lane per vertex (four walks per lane leave too few
// wavefronts in flight on one graph: R-MAT-24 0.108 vs 0.113 ms).
__global__ __launch_bounds__(256) void k_mark_lane(KParams p) {
  const long long ws = wave_slot();
  if (ws >= p.nslots) return;
  const GraphRef G = resolve<false>(p, ws);
  const int lane = lane_id();
  const long long w = ws >> 6;
  const long long slot = ws + lane;
  const int v = (int)(slot - G.slot0);
  const bool live = slot < p.nslots && v < G.n;
  const int R = p.filter ? p.giant[G.g] : -1;
  // the word's lists and the lane's label, loaded before the walks so that
  // their latency hides behind them
  const unsigned long long act = p.actbits[w], low = p.lowbits[w], mid = p.midbits[w],
                           hub = p.hubbits[w];
  const int s = lab_ld(&G.nstat[live ? v : 0]);
  bool member = false;
  if (R >= 0 && live) member = v == R || find_until(v, s, R, G.nstat, p.path_mode) == R;
  const unsigned long long gw = __ballot(member);
  unsigned long long keep = ~0ull;
  if (p.filter == 2) keep = R >= 0 ? ~gw : act;
  // filter 2: the few high-degree vertices left are listed for compute3
  if (p.filter == 2 && ((hub & keep) >> lane) & 1ull) list_hub(p, G, slot, v);
  const unsigned long long hb = p.filter == 2 ? 0ull : hub & keep;
  unsigned long long info = 0;
  if (hb) {
    unsigned tasks = 0;
    if ((hb >> lane) & 1ull) {
      const int deg = G.nidx[v + 1] - G.nidx[v];
      tasks = task_count(deg, p.hub_chunk);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) tasks += __shfl_xor(tasks, d);
    info = (unsigned long long)__popcll(hb) | ((unsigned long long)tasks << 32);
  }
  if (lane == 0) {
    p.gbits[w] = gw;
    p.hubinfo[w] = info;
    if (p.filter == 2) {
      p.lowbits[w] = low & keep;
      p.midbits[w] = mid & keep;
      p.hubbits[w] = 0ull;
    }
  }
}

// Batches: a wavefront takes CPW 256-slot chunks (each of one graph, possibly
// two different ones), 4 * CPW vertices per lane, their walks toward R in
// flight together (the walks are bound by load latency, as in flatten;
// 1024-graph batch: 1.71 -> 1.52 ms at four per lane; 4096-graph mix: 6.14
// -> 5.75 ms at eight).  The launch takes CPW = 2 for large batches only.
template <bool BATCH, int CPW>
__global__ __launch_bounds__(256) void k_mark(KParams p, BlockOrder order) {
  constexpr int H = kSlotAlign / 64;  // words per chunk
  constexpr int K = CPW * H;          // CPW chunks per wavefront
  const long long ws0 =
      ((long long)spread_block(order) * 4 + uniform((int)(threadIdx.x >> 6))) * (CPW * kSlotAlign);
  if (ws0 >= p.nslots) return;
  const long long nwords = (p.nslots + 63) / 64;
  const GraphRef G0 = resolve<BATCH>(p, ws0);
  const GraphRef G1 =
      CPW == 2 && ws0 + kSlotAlign < p.nslots ? resolve<BATCH>(p, ws0 + kSlotAlign) : G0;
  const int lane = lane_id();
  const int R0 = p.filter ? p.giant[G0.g] : -1;
  const int R1 = CPW == 2 && p.filter ? p.giant[G1.g] : R0;
  int v[K], first[K], cur[K], prev[K];
  bool live[K], go[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const GraphRef& G = k < H ? G0 : G1;
    const long long slot = ws0 + k * 64 + lane;
    v[k] = (int)(slot - G.slot0);
    live[k] = slot < p.nslots && v[k] < G.n;
  }
  // the chunk's list words, loaded before the walks so that their latency
  // hides behind them (no other wavefront touches these words during mark)
  unsigned long long actw[K], loww[K], midw[K], hubw[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const long long w = (ws0 >> 6) + k;
    const bool in = w < nwords;
    actw[k] = in ? p.actbits[w] : 0ull;
    loww[k] = in ? p.lowbits[w] : 0ull;
    midw[k] = in ? p.midbits[w] : 0ull;
    hubw[k] = in ? p.hubbits[w] : 0ull;
  }
#pragma unroll
  for (int k = 0; k < K; k++)
    first[k] = ((k < H ? R0 : R1) >= 0 && live[k]) ? lab_ld(&(k < H ? G0 : G1).nstat[v[k]]) : 0;
#pragma unroll
  for (int k = 0; k < K; k++) {
    cur[k] = first[k];
    prev[k] = v[k];
    go[k] = (k < H ? R0 : R1) >= 0 && live[k] && first[k] != v[k];
  }
  // find_until(v, first, R) for the K vertices at once
  bool more = true;
  while (more) {
    int nx[K];
#pragma unroll
    for (int k = 0; k < K; k++)
      nx[k] = (go[k] && cur[k] != (k < H ? R0 : R1) && cur[k] != 0)
                  ? lab_ld(&(k < H ? G0 : G1).nstat[cur[k]])
                  : cur[k];
    more = false;
#pragma unroll
    for (int k = 0; k < K; k++) {
      if (go[k] && cur[k] != (k < H ? R0 : R1) && cur[k] != 0 && cur[k] > nx[k]) {
        if (p.path_mode == 0) lab_st(&(k < H ? G0 : G1).nstat[prev[k]], nx[k]);
        prev[k] = cur[k];
        cur[k] = nx[k];
        more = true;
      } else {
        go[k] = false;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++)
    if (p.path_mode != 0 && (k < H ? R0 : R1) >= 0 && live[k] && first[k] != v[k] && cur[k] != first[k])
      lab_st(&(k < H ? G0 : G1).nstat[v[k]], cur[k]);
#pragma unroll
  for (int k = 0; k < K; k++) {
    const long long w = (ws0 >> 6) + k;
    if (w >= nwords) break;
    const GraphRef& G = k < H ? G0 : G1;
    const int R = k < H ? R0 : R1;
    const bool member = R >= 0 && live[k] && (v[k] == R || (first[k] != v[k] && cur[k] == R));
    const unsigned long long gw = __ballot(member);
    unsigned long long keep = ~0ull;
    if (p.filter == 2) keep = R >= 0 ? ~gw : actw[k];
    // filter 2: the few high-degree vertices left are listed for compute3
    if (p.filter == 2 && ((hubw[k] & keep) >> lane) & 1ull) list_hub(p, G, ws0 + k * 64 + lane, v[k]);
    const unsigned long long hb = p.filter == 2 ? 0ull : hubw[k] & keep;
    unsigned long long info = 0;
    if (hb) {
      unsigned tasks = 0;
      if ((hb >> lane) & 1ull) {
        const int deg = G.nidx[v[k] + 1] - G.nidx[v[k]];
        tasks = task_count(deg, p.hub_chunk);
      }
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) tasks += __shfl_xor(tasks, d);
      info = (unsigned long long)__popcll(hb) | ((unsigned long long)tasks << 32);
    }
    if (lane == 0) {
      p.gbits[w] = gw;
      p.hubinfo[w] = info;
      if (p.filter == 2) {
        p.lowbits[w] = loww[k] & keep;
        p.midbits[w] = midw[k] & keep;
        p.hubbits[w] = 0ull;
      }
    }
  }
}

// ---- compute2 (ECL-CC-atomic.cu:161-199) --------------------------------------
// Wavefront w owns GC consecutive 256-slot chunks (each of one graph; GC is
// 1, 4 or 16, by launch size) and reads their medium-degree bits in one
// load.  For each chunk that lists a vertex it gathers the chunk's medium-
// degree vertices into an LDS list, and processes them in groups of up to
// 64, one per lane: their adjacency lists are laid end to end and dealt out
// evenly over the 64 lanes, four entries per lane per step, so a wavefront
// never idles on a short list.  The owner of each entry comes from a
// per-wavefront LDS table
// (one byte per entry, filled list by list for a window of kWaveWindow
// entries); list bases and root estimates of the 64 lanes sit beside it.
// A vertex already marked as joined to the sampled root uses that root as
// its estimate without walking.  Entries whose two ends are both marked are
// skipped after one bitmap lookup (filter 1); with filter 2 the vertices
// here lie outside the largest tree and every entry is kept, a neighbour
// inside the tree replaced by its root.  The kept entries are staged in LDS
// as (neighbour, root estimate) pairs and hooked together when the stage
// fills and at the end of each chunk, so the streaming steps never wait for
// label chains.

constexpr int kWaveWindow = 2048;
constexpr int kWaveSurvivors = 256;  // staged (neighbour, root estimate) pairs per wavefront

// Hooks the staged pairs of one wavefront, all lanes busy.
__device__ __forceinline__ void flush_wave(const int2* surv, int cnt, int* ns, int mode) {
  const int lane = lane_id();
  for (int j0 = 0; j0 < cnt; j0 += 256) {
    int u[4], vs[4], sv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int j = j0 + k * 64 + lane;
      const int2 e = j < cnt ? surv[j] : make_int2(-1, 0);
      u[k] = e.x;
      vs[k] = e.y;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) sv[k] = u[k] >= 0 ? lab_ld(&ns[u[k]]) : 0;
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (u[k] >= 0) (void)hook_edge(vs[k], u[k], sv[k], ns, mode);
  }
}

template <bool BATCH, int GC>
__global__ __launch_bounds__(256) void k_hook_wave(KParams p) {
  static_assert(GC >= 1 && GC * (kSlotAlign / 64) <= 64, "one medium word per lane");
  constexpr int H = kSlotAlign / 64;  // medium words per chunk
  __shared__ unsigned char s_owner[4][kWaveWindow];
  __shared__ unsigned char s_list[4][kSlotAlign];
  __shared__ int s_base[4][64];
  __shared__ int s_root[4][64];
  __shared__ int s_vtx[4][64];
  __shared__ int2 s_surv[4][kWaveSurvivors];
  // GC consecutive 256-slot chunks (each of one graph) per wavefront; their
  // medium words come in one load, one word per lane, and a chunk that lists
  // no vertex costs nothing more
  const long long c0 = ((long long)blockIdx.x * 4 + uniform((int)(threadIdx.x >> 6))) * GC;
  if (c0 * kSlotAlign >= p.nslots) return;
  const long long nwords = (p.nslots + 63) / 64;
  const int lane = lane_id();
  const int wv = uniform((int)(threadIdx.x >> 6));
  const long long wl = c0 * H + lane;
  const unsigned long long mword = lane < GC * H && wl < nwords ? p.midbits[wl] : 0ull;
  const unsigned long long busy = __ballot(mword != 0ull);
  unsigned char* const list = s_list[wv];
  for (int c = 0; c < GC; c++) {
    if (((busy >> (c * H)) & ((1ull << H) - 1ull)) == 0ull) continue;
    const long long ws0 = (c0 + c) * kSlotAlign;
    int cnt = 0;  // wave-uniform
#pragma unroll
    for (int i = 0; i < H; i++) {
      const unsigned long long bits = __shfl(mword, c * H + i);
      if ((bits >> lane) & 1ull) list[cnt + lanes_below(bits)] = (unsigned char)(i * 64 + lane);
      cnt += __popcll(bits);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const GraphRef G = resolve<BATCH>(p, ws0);
    int* const ns = G.nstat;
    const int* const nl = G.nlist;
    unsigned char* const owner = s_owner[wv];
    int2* const surv = s_surv[wv];
    const int R = p.filter ? p.giant[G.g] : -1;
    const bool skip = p.filter == 2 && R >= 0;  // listed vertices lie outside R's tree
    int nsurv = 0;  // wave-uniform
    unsigned long long intree = 0;  // per lane: owners seen with an entry inside R's tree
    // groups of up to 64 listed vertices, one per lane
    for (int g0 = 0; g0 < cnt; g0 += 64) {
      const int gn = cnt - g0 < 64 ? cnt - g0 : 64;
      const unsigned long long bits = gn == 64 ? ~0ull : ((1ull << gn) - 1ull);
      const bool mine = lane < gn;
      const long long slot = ws0 + (mine ? list[g0 + lane] : 0);
      const int v = (int)(slot - G.slot0);
      // a marked vertex is joined to the sampled root: that root is a valid
      // estimate (the CAS of a hook checks it), no walk needed
      const bool marked = mine && p.filter && !skip && slot_bit(p.gbits, slot);
      const unsigned long long mk = __ballot(marked);
      int beg = 0, deg = 0, root = 0;
      if (mine) {
        root = marked ? R : lab_ld(&ns[v]);
        beg = G.nidx[v];
        deg = G.nidx[v + 1] - beg;
      }
      if (mine && !marked) root = find_from(v, root, ns, p.path_mode);
      const int incl = (int)wave_incl_sum((unsigned)deg);
      const int off = incl - deg;
      const int total = __shfl(incl, 63);
      s_base[wv][lane] = beg - off;  // entry e of lane L's list is nlist[base_L + e]
      s_root[wv][lane] = root;
      s_vtx[wv][lane] = v;
      // window positions are unsigned: the lists of one wavefront may hold
      // nearly 2^31 entries, and a step past the last must not wrap
      const unsigned utotal = (unsigned)total;
      for (unsigned w0 = 0; w0 < utotal; w0 += kWaveWindow) {
        const unsigned wend = w0 + kWaveWindow < utotal ? w0 + kWaveWindow : utotal;
        for (unsigned long long m = bits; m; m &= m - 1) {
          const int L = __ffsll((long long)m) - 1;
          const int oL = __builtin_amdgcn_readlane(off, L);
          const int dL = __builtin_amdgcn_readlane(deg, L);
          const unsigned a = (unsigned)oL > w0 ? (unsigned)oL : w0;
          const unsigned b = (unsigned)(oL + dL) < wend ? (unsigned)(oL + dL) : wend;
          for (unsigned x = a + lane; x < b; x += 64) owner[x - w0] = (unsigned char)L;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (unsigned e0 = w0; e0 < wend; e0 += 256) {
          int u[4], src[4];
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const unsigned e = e0 + k * 64 + lane;
            const bool valid = e < wend;
            src[k] = valid ? owner[e - w0] : 0;
            u[k] = valid ? ld_stream(&nl[s_base[wv][src[k]] + e]) : INT_MAX;
          }
          unsigned keep = 0;
          if (skip) {
            // every entry; a neighbour inside the largest tree stands for R,
            // and all such entries of one vertex join the same two trees: the
            // owner is noted and hooked to R once, after its list
#pragma unroll
            for (int k = 0; k < 4; k++)
              keep |= (u[k] != INT_MAX && u[k] != s_vtx[wv][src[k]]) ? (1u << k) : 0u;
#pragma unroll
            for (int k = 0; k < 4; k++) {
              if (((keep >> k) & 1u) && slot_bit(p.gbits, G.slot0 + u[k])) {
                keep &= ~(1u << k);
                intree |= 1ull << src[k];
              }
            }
          } else {
#pragma unroll
            for (int k = 0; k < 4; k++) keep |= (s_vtx[wv][src[k]] > u[k]) ? (1u << k) : 0u;
            if (mk) {
              unsigned in = 0;
#pragma unroll
              for (int k = 0; k < 4; k++)
                if (((keep >> k) & 1u) && ((mk >> src[k]) & 1ull))
                  in |= slot_bit(p.gbits, G.slot0 + u[k]) << k;
              keep &= ~in;
            }
          }
          // stage the entries that still need a hook; flush when the stage is full
          const unsigned c = __popc(keep);
          const unsigned pos = wave_incl_sum(c);
          const int wtot = (int)__shfl(pos, 63);
          if (wtot) {
            if (nsurv + wtot > kWaveSurvivors) {
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
              __builtin_amdgcn_wave_barrier();
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
              flush_wave(surv, nsurv, ns, p.path_mode);
              nsurv = 0;
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
              __builtin_amdgcn_wave_barrier();
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            int at = nsurv + (int)(pos - c);
#pragma unroll
            for (int k = 0; k < 4; k++)
              if ((keep >> k) & 1u) surv[at++] = make_int2(u[k], s_root[wv][src[k]]);
            nsurv += wtot;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      if (skip) {
        // one hook to R per listed vertex with a neighbour in R's tree
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) intree |= __shfl_xor(intree, d);
        if ((intree >> lane) & 1ull) (void)hook_edge(root, R, lab_ld(&ns[R]), ns, p.path_mode);
        intree = 0;
      }
    }
    if (nsurv) flush_wave(surv, nsurv, ns, p.path_mode);
    // the next chunk rewrites the LDS tables
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// ---- compute3 (ECL-CC-atomic.cu:203-246) --------------------------------------
// Each block takes a contiguous range of the edge tasks of all high-degree
// vertices (one search for its first task, then a forward walk of the task
// list).  A task is hub_chunk consecutive entries of one adjacency list,
// read in rounds of kHubPerThread coalesced loads per thread, all in flight
// together.  The entries that still need a hook (< v, and not both ends in
// the largest tree; with filter 2, every entry of a listed vertex, which
// lies outside that tree) are staged in LDS as (neighbour, root estimate of v)
// pairs -- one LDS atomic per wavefront per round -- and kept there across
// tasks; the block hooks them all together when the stage is about to fill
// and at the end, with every thread busy and kHubPerThread label loads in
// flight.  So the streaming part of a task never waits for label chains.

template <int R, int BT>
__device__ __forceinline__ void flush_stage(const int2* stage, int cnt, int* ns, int mode) {
  for (int j0 = 0; j0 < cnt; j0 += R * BT) {
    int w[R], vs[R], sv[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int j = j0 + r * BT + (int)threadIdx.x;
      const int2 e = j < cnt ? stage[j] : make_int2(-1, 0);
      w[r] = e.x;
      vs[r] = e.y;
    }
#pragma unroll
    for (int r = 0; r < R; r++) sv[r] = w[r] >= 0 ? lab_ld(&ns[w[r]]) : 0;
#pragma unroll
    for (int r = 0; r < R; r++)
      if (w[r] >= 0) (void)hook_edge(vs[r], w[r], sv[r], ns, mode);
  }
}

template <bool BATCH, int BT>
__global__ __launch_bounds__(BT) void k_hook_block(KParams p) {
  constexpr int R = kHubPerThread;
  constexpr int CAP = 2 * R * BT;  // staged pairs; room for one more round when below R*BT
  __shared__ int2 stage[CAP];
  __shared__ int s_cnt;
  // the hubs and tasks: with filter 2 mark's counter (its atomic adds
  // finished with an earlier kernel on this stream, so a plain load sees
  // their sum), otherwise the scan's block totals
  unsigned long long ctr;
  if (p.filter == 2) {
    ctr = *p.hubctr;
  } else {
    const long long nsb = ((p.nslots + 63) / 64 + kScanBlock - 1) / kScanBlock;
    ctr = sum_block_totals(p, nsb);
    if (blockIdx.x == 0 && threadIdx.x == 0) *p.hubctr = ctr;  // for eclcc_workspace_stats
  }
  const long long pushed = (long long)(unsigned)ctr;
  const long long nhubs = pushed < p.hub_cap ? pushed : p.hub_cap;
  const long long ntasks = (long long)(ctr >> 32);
  if (nhubs <= 0) return;
  // each block takes one contiguous range of tasks
  const long long per = (ntasks + gridDim.x - 1) / gridDim.x;
  const long long t0 = (long long)blockIdx.x * per;
  const long long t1 = t0 + per < ntasks ? t0 + per : ntasks;
  if (t0 >= t1) return;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  if (tid == 0) s_cnt = 0;
  long long k = 0, hi = nhubs - 1;  // last hub whose first task is <= t0
  while (k < hi) {
    const long long mid = (k + hi + 1) >> 1;
    if (p.hubs[mid].first_task <= t0) k = mid;
    else hi = mid - 1;
  }
  int* ns = nullptr;
  for (long long t = t0; t < t1; t++) {
    while (k + 1 < nhubs && p.hubs[k + 1].first_task <= t) k++;
    const HubEntry h = p.hubs[k];
    const GraphRef G = resolve<BATCH>(p, h.slot);
    if (G.nstat != ns) {  // staged pairs belong to the previous graph: hook them first
      __syncthreads();
      const int cnt = s_cnt;
      if (cnt > 0) flush_stage<R, BT>(stage, cnt, ns, p.path_mode);
      __syncthreads();
      if (tid == 0) s_cnt = 0;
      ns = G.nstat;
    }
    const int v = (int)(h.slot - G.slot0);
    const int Rg = p.filter ? p.giant[G.g] : -1;
    const bool skip = p.filter == 2 && Rg >= 0;  // v lies outside Rg's tree
    const bool vin = !skip && p.filter && slot_bit(p.gbits, h.slot);
    const long long cb = (long long)h.beg + (t - h.first_task) * (long long)p.hub_chunk;
    const int cbeg = (int)cb;
    const int cend = (int)(cb + p.hub_chunk < (long long)h.end ? cb + p.hub_chunk : (long long)h.end);
    // root estimate of v: the sampled root if v is marked, else found once
    const int vstat = vin ? Rg : find_root(v, G.nstat, p.path_mode);
    for (unsigned base = cbeg; base < (unsigned)cend; base += R * BT) {
      __syncthreads();
      if (s_cnt > CAP - R * BT) {
        const int cnt = s_cnt;
        flush_stage<R, BT>(stage, cnt, ns, p.path_mode);
        __syncthreads();
        if (tid == 0) s_cnt = 0;
        __syncthreads();
      }
      int u[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const unsigned i = base + r * BT + tid;
        u[r] = i < (unsigned)cend ? ld_stream(&G.nlist[i]) : INT_MAX;
      }
      unsigned keep = 0;
      if (skip) {
        // every entry; a neighbour inside the largest tree stands for its root
#pragma unroll
        for (int r = 0; r < R; r++) keep |= (u[r] != INT_MAX && u[r] != v) ? (1u << r) : 0u;
#pragma unroll
        for (int r = 0; r < R; r++)
          if (((keep >> r) & 1u) && slot_bit(p.gbits, G.slot0 + u[r])) u[r] = Rg;
      } else {
        unsigned in[R];
#pragma unroll
        for (int r = 0; r < R; r++) in[r] = (vin && u[r] < v) ? slot_bit(p.gbits, G.slot0 + u[r]) : 0u;
#pragma unroll
        for (int r = 0; r < R; r++) keep |= (u[r] < v && !in[r]) ? (1u << r) : 0u;
      }
      const unsigned c = __popc(keep);
      const unsigned incl = wave_incl_sum(c);
      const unsigned wtot = __shfl(incl, 63);
      if (wtot) {
        int base_w = 0;
        if (lane == 0) base_w = atomicAdd(&s_cnt, (int)wtot);
        base_w = __shfl(base_w, 0);
        int pos = base_w + (int)(incl - c);
#pragma unroll
        for (int r = 0; r < R; r++)
          if (keep & (1u << r)) stage[pos++] = make_int2(u[r], vstat);
      }
    }
  }
  __syncthreads();
  const int cnt = s_cnt;
  if (cnt > 0) flush_stage<R, BT>(stage, cnt, ns, p.path_mode);
}

// ---- flatten (ECL-CC-atomic.cu:250-264) ---------------------------------------

// A wavefront takes CPW 256-slot chunks (each of one graph, possibly two
// different ones): 4 * CPW vertices per lane, 64 slots apart, their walks in
// flight together; the launch takes CPW = 2 for large launches only.  The walks are short (mark and compute leave most labels
// one step from the root), so the pass is bound by load latency and
// independent chains per lane hide it (R-MAT-24: 0.033 -> 0.024 ms at four
// per lane; grid 256^3: 0.027 -> 0.019 ms; eight per lane: the 4096-graph
// mix 0.81 -> 0.65 ms, one large graph unchanged).
template <bool BATCH, int CPW>
__global__ __launch_bounds__(256) void k_flatten(KParams p) {
  constexpr int H = kSlotAlign / 64;  // words per chunk
  constexpr int K = CPW * H;          // CPW chunks per wavefront
  const long long ws = ((long long)blockIdx.x * 4 + uniform((int)(threadIdx.x >> 6))) * (CPW * kSlotAlign);
  if (ws >= p.nslots) return;
  const GraphRef G0 = resolve<BATCH>(p, ws);
  const GraphRef G1 = CPW == 2 && ws + kSlotAlign < p.nslots ? resolve<BATCH>(p, ws + kSlotAlign) : G0;
  const int lane = lane_id();
  int v[K], s[K], old[K];
  bool go[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const GraphRef& G = k < H ? G0 : G1;
    const long long slot = ws + k * 64 + lane;
    v[k] = (int)(slot - G.slot0);
    go[k] = slot < p.nslots && v[k] < G.n;
  }
#pragma unroll
  for (int k = 0; k < K; k++) s[k] = go[k] ? lab_ld(&(k < H ? G0 : G1).nstat[v[k]]) : 0;
#pragma unroll
  for (int k = 0; k < K; k++) old[k] = s[k];
  bool more = true;
  while (more) {
    int nx[K];
#pragma unroll
    for (int k = 0; k < K; k++) nx[k] = (go[k] && s[k] != 0) ? lab_ld(&(k < H ? G0 : G1).nstat[s[k]]) : s[k];
    more = false;
#pragma unroll
    for (int k = 0; k < K; k++) {
      if (go[k] && s[k] != 0 && s[k] > nx[k]) {
        s[k] = nx[k];
        more = true;
      } else {
        go[k] = false;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const GraphRef& G = k < H ? G0 : G1;
    const long long slot = ws + k * 64 + lane;
    if (slot < p.nslots && v[k] < G.n && s[k] != old[k]) lab_st(&G.nstat[v[k]], s[k]);
  }
}

// ---- host-side launch -----------------------------------------------------------

// Labels per window of the batched mark: its walks mostly read, and a
// window of 32 MB of labels keeps them in the Infinity Cache while still
// spreading the resident blocks over hundreds of graphs (4096-graph mix,
// mark ms at windows of 4 / 8 / 16 / 32 / 128 / 512 MB and one window:
// 4.92 / 4.72 / 4.61 / 4.51 / 4.86 / 5.11 / 5.76).  Sample keeps one
// window: its hooks end in a CAS on each graph's few roots, and windows
// crowd them (7.39 ms at one window, 7.48 at 512 MB, 8.05 at 128, 9.27 at
// 32).
constexpr long long kMarkWindowBytes = 32ll << 20;

// Blocks a vertex-parallel launch should have at least (8 per CU), where
// its per-block share of words can shrink.
constexpr long long kMinBlocks = 2048;
// Wavefronts a launch with several chunks per wavefront must have at least:
// two rounds of the GPU's 8192 resident wavefronts (256 CUs x 4 SIMDs x 8).
constexpr long long kTwoChunkWaves = 16384;

template <bool BATCH>
static hipError_t launch_all(const KParams& p, const Shape& shape, hipStream_t s, hipEvent_t* ev) {
  const long long vb = (p.nslots + 255) / 256;
  if (vb <= 0) return hipSuccess;
  if (BATCH && p.nslots % kSlotAlign) return hipErrorInvalidValue;
  const long long words = (p.nslots + 63) / 64;
  const long long sb = (words + kScanBlock - 1) / kScanBlock;
  const dim3 grid((unsigned)vb), block(256);
  // chunks per wavefront of the batched mark and of flatten: two (eight
  // walks per lane) where that still leaves two rounds of resident
  // wavefronts, else one (a small batch would leave CUs idle: a one-graph
  // batch of ER-65,536 took 86.7 against 72.6 us)
  const int cpw = p.nslots / (2 * kSlotAlign) >= kTwoChunkWaves ? 2 : 1;
  int k = 0;
  auto mark_ev = [&]() {
    if (ev) (void)hipEventRecord(ev[k], s);
    k++;
  };
  mark_ev();
  k_init<BATCH><<<grid, block, 0, s>>>(p);
  mark_ev();
  if constexpr (BATCH) {
    // words per sample block: kSampleWords, fewer where that leaves fewer
    // than kMinBlocks blocks
    int sspan = kSampleWords;
    while (sspan > 1 && (words + sspan - 1) / sspan < kMinBlocks) sspan /= 2;
    const unsigned nb = (unsigned)((words + sspan - 1) / sspan);
    k_sample_dense<true><<<dim3(nb), block, 0, s>>>(p, block_order(nb, 64ll * sspan, 0), sspan);
  } else {
    k_sample_lane<<<grid, block, 0, s>>>(p);
  }
  mark_ev();
  k_pick<BATCH><<<dim3((unsigned)p.ngraphs), dim3(kPickSamples), 0, s>>>(p);
  mark_ev();
  if constexpr (BATCH) {
    const unsigned nb = (unsigned)((p.nslots + 4 * cpw * kSlotAlign - 1) / (4 * cpw * kSlotAlign));
    const BlockOrder order = block_order(nb, 4ll * cpw * kSlotAlign, kMarkWindowBytes);
    if (cpw == 2) k_mark<true, 2><<<dim3(nb), block, 0, s>>>(p, order);
    else k_mark<true, 1><<<dim3(nb), block, 0, s>>>(p, order);
  } else {
    k_mark_lane<<<grid, block, 0, s>>>(p);
  }
  mark_ev();
  if (p.filter != 2) k_hub_scan<<<dim3((unsigned)sb), dim3(kScanBlock), 0, s>>>(p);
  mark_ev();
  // words per compute1 block: as many as keep >= kMinBlocks blocks, 4..kLowWords
  int span = 4;
  while (span < kLowWords && words / (2 * span) >= kMinBlocks) span *= 2;
  k_hook_thread<BATCH><<<dim3((unsigned)((words + span - 1) / span)), block, 0, s>>>(p, span);
  mark_ev();
  {
    // chunks per compute2 wavefront: the most of 16, 4, 1 that still leaves
    // two rounds of resident wavefronts.  Most chunks of a batch list no
    // medium vertex (4096-graph mix: 0.834 ms at one chunk, 0.602 at four,
    // 0.500 at sixteen); R-MAT-24's are busier and want the wavefronts
    // (0.0226 ms at one, 0.0186 at four, 0.0366 at sixteen)
    const long long chunks = (p.nslots + kSlotAlign - 1) / kSlotAlign;
    const int gc = chunks / 16 >= kTwoChunkWaves ? 16 : chunks / 4 >= kTwoChunkWaves ? 4 : 1;
    const dim3 cgrid((unsigned)((chunks + 4 * gc - 1) / (4 * gc)));
    if (gc == 16) k_hook_wave<BATCH, 16><<<cgrid, block, 0, s>>>(p);
    else if (gc == 4) k_hook_wave<BATCH, 4><<<cgrid, block, 0, s>>>(p);
    else k_hook_wave<BATCH, 1><<<cgrid, block, 0, s>>>(p);
  }
  mark_ev();
  {
    // filter 2 too: the lists of the high-degree vertices mark left listed
    // (few or none; every block then leaves after one load)
    const dim3 hgrid((unsigned)shape.hub_blocks);
    switch (shape.block_threads) {
      case 1024: k_hook_block<BATCH, 1024><<<hgrid, dim3(1024), 0, s>>>(p); break;
      case 512: k_hook_block<BATCH, 512><<<hgrid, dim3(512), 0, s>>>(p); break;
      default: k_hook_block<BATCH, 256><<<hgrid, dim3(256), 0, s>>>(p); break;
    }
  }
  mark_ev();
  {
    const dim3 fgrid((unsigned)((p.nslots + 4 * cpw * kSlotAlign - 1) / (4 * cpw * kSlotAlign)));
    if (cpw == 2) k_flatten<BATCH, 2><<<fgrid, block, 0, s>>>(p);
    else k_flatten<BATCH, 1><<<fgrid, block, 0, s>>>(p);
  }
  mark_ev();
  return hipGetLastError();
}

hipError_t launch_pipeline(const KParams& p, bool batch, const Shape& shape, hipStream_t s,
                           hipEvent_t* ev) {
  return batch ? launch_all<true>(p, shape, s, ev) : launch_all<false>(p, shape, s, ev);
}

int hub_block_residency(int block_threads) {
  int per_cu = 0;
  hipError_t e;
  switch (block_threads) {
    case 1024:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_hook_block<false, 1024>, 1024, 0);
      break;
    case 512:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_hook_block<false, 512>, 512, 0);
      break;
    default:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_hook_block<false, 256>, 256, 0);
      break;
  }
  return e == hipSuccess && per_cu > 0 ? per_cu : 1;
}

hipError_t launch_init_only(const KParams& p, hipStream_t s) {
  const long long vb = (p.nslots + 255) / 256;
  if (vb <= 0) return hipSuccess;
  k_init<false><<<dim3((unsigned)vb), dim3(256), 0, s>>>(p);
  return hipGetLastError();
}

}  // namespace eclcc
