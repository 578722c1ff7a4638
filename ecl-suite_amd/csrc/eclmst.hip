// eclmst.hip -- race-free ECL-MST minimum spanning forest for gfx950
// (MI355X): kernels and the C ABI of include/eclmst.h (libeclmst.so).
//
// Reference algorithm: src/racefree/egr-input/MST/ECL-MST_10-atomic.cu
//   find / join :157-174, initPM :177-184, initWL :187-237, kernel1 :240-256,
//   kernel2 :259-271, kernel3 :274-282, host loop gpuMST :285-382, threshold
//   choice in main :450-462.
//
// Boruvka over an edge worklist, as the reference: every edge in the list
// proposes itself, keyed weight << 32 | j, to the minimum of both of its
// components (64-bit atomic min); an edge that is the minimum of either
// component is joined (CAS union, larger root below the smaller) and marked;
// the joined edges' component minima are reset; the list is rebuilt from
// the edges whose ends are still in different components, until it is
// empty.  With a light-edge threshold the edges up to it run first and the
// rest after, most of them then dropped at once because their ends are
// already joined.  The keys are distinct, so the forest is the unique
// minimum under (weight, j) whatever order the threads run in.
//
// What is MI355X-specific:
//   * No per-edge or per-wave atomic on a shared list counter: a contended
//     device-scope atomic executes at the memory side and serialises
//     (~10 ns each).  Blocks are persistent and append through a 1024-entry
//     LDS buffer, one counter atomic per buffer flush; 17 KB of LDS per
//     block leaves room for 8 blocks (32 wavefronts) per CU.
//   * The first lists are built edge-parallel over the CSR (no thread owns a
//     hub's whole list): each block takes a contiguous range of entries,
//     finds the vertex of its first entry by a 64-way parallel search of
//     nindex, then maps entries to vertices through an LDS window of nindex.
//   * Between the light and the heavy phase every vertex is pointed at its
//     root, so the heavy phase reads each end's root with one load.
//   * The component minima are reset by one coalesced pass over the array,
//     not per listed edge: the stores of a large tree's edges would all hit
//     its root's entry, and stores to one address serialise.
//   * Race freedom by construction: parent and minv are only touched through
//     relaxed agent-scope atomics; each in_mst byte has one writer.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <exception>
#include <mutex>
#include <new>
#include <string>

#include "../../include/eclmst.h"
#include "eclcc_stage.h"

namespace eclmst {

constexpr int kThreads = 256;
constexpr int kPer = 4;                      // items per thread per step
constexpr int kTile = kThreads * kPer;       // items per block step
constexpr int kBuf = kTile;                  // LDS output buffer (int4 entries)
constexpr int kWin = kThreads - 1;           // vertices in the LDS nindex window (one entry per thread)
constexpr int kSamples = 256;                // weights sampled for the threshold
constexpr unsigned long long kNoEdge = ~0ull;

// flag bits in the control block
constexpr unsigned kFlagOverflow = 1u;
constexpr unsigned kFlagNegative = 2u;

struct Ctrl {
  unsigned count[2];  // worklist sizes
  unsigned flags;
  int threshold;
  int giant;  // root of the largest sampled tree after the light phase, or -1
};

// ---- atomics (relaxed, agent scope) ----------------------------------------

__device__ __forceinline__ int ld(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int cas(int* p, int expect, int desired) {
  __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expect;
}
__device__ __forceinline__ unsigned long long ld64(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st64(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void min64(unsigned long long* p, unsigned long long v) {
  // test first: once a component's minimum is low, most proposals lose
  // without issuing the memory-side atomic
  if (ld64(p) > v) (void)__hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x (ECL-MST_10-atomic.cu:157-164: read-only walk).
__device__ __forceinline__ int find(int x, int* parent) {
  int p;
  while ((p = ld(&parent[x])) != x) x = p;
  return x;
}

// Union of the trees of a and b (:166-174): the larger root is swung below
// the smaller by CAS; a failed CAS returns the larger one's new parent and
// the union goes on from there.
__device__ __forceinline__ void join(int a, int b, int* parent) {
  while (a != b) {
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int seen = cas(&parent[a], a, b);
    if (seen == a) return;
    a = seen;
  }
}

// bit v of a membership bitmap (written by an earlier kernel)
__device__ __forceinline__ bool member(const unsigned long long* gbits, int v) {
  return (gbits[v >> 6] >> (v & 63)) & 1ull;
}

__device__ __forceinline__ unsigned long long key_of(int4 e) {
  return ((unsigned long long)(unsigned)e.z << 32) | (unsigned)e.w;
}

// ---- block-wide append through an LDS buffer ---------------------------------

struct Appender {
  int4 buf[kBuf];
  int fill;
  unsigned base;
  int wsum[kThreads / 64];
};

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ int wave_incl_sum(int x) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int t = __shfl_up(x, d);
    if (lane >= d) x += t;
  }
  return x;
}

__device__ void flush(Appender& A, int4* wl, unsigned* cnt, unsigned cap, unsigned* flags) {
  if (threadIdx.x == 0)
    A.base = A.fill ? __hip_atomic_fetch_add(cnt, (unsigned)A.fill, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)
                    : 0u;
  __syncthreads();
  const int fill = A.fill;
  const unsigned base = A.base;
  for (int i = threadIdx.x; i < fill; i += kThreads) {
    const unsigned long long p = (unsigned long long)base + (unsigned)i;
    if (p < cap) wl[p] = A.buf[i];
    else __hip_atomic_fetch_or(flags, kFlagOverflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  if (threadIdx.x == 0) A.fill = 0;
  __syncthreads();
}

// Appends this thread's items with ok set (all threads of the block call it).
__device__ void append(Appender& A, const int4 (&it)[kPer], const bool (&ok)[kPer], int4* wl,
                       unsigned* cnt, unsigned cap, unsigned* flags) {
  int c = 0;
#pragma unroll
  for (int r = 0; r < kPer; r++) c += ok[r];
  const int incl = wave_incl_sum(c);
  const int wv = threadIdx.x >> 6;
  if (lane_id() == 63) A.wsum[wv] = incl;
  __syncthreads();
  int off = 0, total = 0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; k++) {
    off += k < wv ? A.wsum[k] : 0;
    total += A.wsum[k];
  }
  if (A.fill + total > kBuf) flush(A, wl, cnt, cap, flags);  // block-uniform
  int pos = A.fill + off + incl - c;
#pragma unroll
  for (int r = 0; r < kPer; r++)
    if (ok[r]) A.buf[pos++] = it[r];
  __syncthreads();
  if (threadIdx.x == 0) A.fill += total;
  __syncthreads();
}

// The same buffer with its fill count kept in registers by every thread
// (each computes the same block totals), so an append costs one block
// barrier instead of three.  k_first uses it (with the step mapping below,
// R-MAT-24: 1.89 -> 1.31 ms for the light list); k_link keeps the other
// (with this one its rounds took 4.25 instead of 3.49 ms per forest).
struct RegAppender {
  int4 buf[kBuf];
  unsigned base;
  int wsum[2][kThreads / 64];  // per-wavefront counts, double-buffered by step parity
};

// Writes the buffer's `fill` items (block-uniform) to the list.  All threads
// of the block call it; its first barrier also orders every thread's buffer
// writes before the reads here, its second the reads before later writes.
__device__ void flush_reg(RegAppender& A, int fill, int4* wl, unsigned* cnt, unsigned cap, unsigned* flags) {
  if (fill == 0) return;  // block-uniform
  if (threadIdx.x == 0)
    A.base = __hip_atomic_fetch_add(cnt, (unsigned)fill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const unsigned base = A.base;
  for (int i = threadIdx.x; i < fill; i += kThreads) {
    const unsigned long long p = (unsigned long long)base + (unsigned)i;
    if (p < cap) wl[p] = A.buf[i];
    else __hip_atomic_fetch_or(flags, kFlagOverflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
}

// Appends this thread's items with ok set; all threads of the block call it
// with the same `fill` (the buffer's item count, kept in registers by every
// thread) and `par` (the step's parity: the counts of consecutive steps use
// different slots, so no barrier is needed after reading them).
__device__ void append_reg(RegAppender& A, int& fill, int par, const int4 (&it)[kPer], const bool (&ok)[kPer],
                       int4* wl, unsigned* cnt, unsigned cap, unsigned* flags) {
  int c = 0;
#pragma unroll
  for (int r = 0; r < kPer; r++) c += ok[r];
  const int incl = wave_incl_sum(c);
  const int wv = threadIdx.x >> 6;
  if (lane_id() == 63) A.wsum[par][wv] = incl;
  __syncthreads();
  int off = 0, total = 0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; k++) {
    const int x = A.wsum[par][k];
    off += k < wv ? x : 0;
    total += x;
  }
  if (fill + total > kBuf) {  // block-uniform
    flush_reg(A, fill, wl, cnt, cap, flags);
    fill = 0;
  }
  int pos = fill + off + incl - c;
#pragma unroll
  for (int r = 0; r < kPer; r++)
    if (ok[r]) A.buf[pos++] = it[r];
  fill += total;
}

// ---- kernels -------------------------------------------------------------------

// initPM (:177-184) and the zeroed output.
__global__ __launch_bounds__(kThreads) void k_init(int n, int* parent, unsigned long long* minv,
                                                   long long m, unsigned char* in_mst,
                                                   Ctrl* ctrl) {
  const long long stride = (long long)gridDim.x * kThreads;
  const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
  for (long long v = t; v < n; v += stride) {
    parent[v] = (int)v;  // no other thread reads parent during this kernel
    minv[v] = kNoEdge;
  }
  // in_mst: byte head, 16-byte body, byte tail
  const uintptr_t a = reinterpret_cast<uintptr_t>(in_mst);
  const long long head = std::min<long long>(m, (long long)((16 - (a & 15)) & 15));
  const long long body = (m - head) / 16;
  for (long long i = t; i < head; i += stride) in_mst[i] = 0;
  uint4* w = reinterpret_cast<uint4*>(in_mst + head);
  for (long long i = t; i < body; i += stride) w[i] = make_uint4(0, 0, 0, 0);
  for (long long i = head + body * 16 + t; i < m; i += stride) in_mst[i] = 0;
  if (t == 0) {
    ctrl->count[0] = ctrl->count[1] = 0;
    ctrl->flags = 0;
  }
}

// Threshold (main, :450-462): the sampled weight below which about 3 * n of
// the m entries lie, from kSamples evenly spaced entries (rank by counting).
__global__ __launch_bounds__(kSamples) void k_threshold(int n, long long m, const int* w,
                                                        Ctrl* ctrl) {
  __shared__ int s[kSamples];
  const int t = threadIdx.x;
  const int x = w[(long long)t * m / kSamples];
  s[t] = x;
  __syncthreads();
  int rank = 0;
  for (int k = 0; k < kSamples; k++) rank += (s[k] < x) || (s[k] == x && k < t);
  const int idx = (int)std::min<double>(kSamples - 1, 3.0 * n * kSamples / (double)m);
  if (rank == idx) ctrl->threshold = x;
}

// Slot of `key` in the sorted window s[0..kWin]: the last x with s[x] <= key.
__device__ __forceinline__ int window_slot(const int* s, long long key) {
  int lo = 0, hi = kWin;  // s[lo] <= key holds
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s[mid] <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// initWL (:187-237), edge-parallel.  HEAVY = false: the entries i -> k with
// k > i and weight <= threshold, as (i, k, w, j).  HEAVY = true (after
// k_flatten): those with weight > threshold whose ends lie in different
// trees, as (root i, root k, w, j).  Every weight is checked for sign.
// A step takes kTile entries: each thread four consecutive ones, loaded as
// 16-byte vectors when both arrays are 16-byte aligned (`vec`), and finds
// the vertex of its first entry by one binary search in the LDS window of
// nindex and of the next three by stepping forward.  (One search per entry,
// with the entries of a thread kTile / 4 apart, made the searches' LDS
// reads the step's bottleneck: a pass that only streams the same bytes
// took 0.78 ms where the list build took 1.87.)  The window of the step
// after is loaded into the other half of s_idx, so no barrier separates
// this step's reads of it from the next step's writes.
template <bool HEAVY>
__global__ __launch_bounds__(kThreads) void k_first(int n, long long m, const int* nindex,
                                                    const int* nlist, const int* eweight,
                                                    int* parent, const unsigned long long* gbits,
                                                    Ctrl* ctrl, int4* wl, unsigned cap,
                                                    long long per, bool vec) {
  static_assert(kWin + 1 == kThreads && kTile == 4 * kThreads && kPer == 4, "one window entry and four list entries per thread");
  __shared__ RegAppender A;
  __shared__ int s_idx[2][kWin + 1];
  __shared__ int s_i0;
  const int thr = ctrl->threshold;
  const bool giant = HEAVY && ctrl->giant >= 0;
  const long long jb0 = (long long)blockIdx.x * per;
  const long long jb1 = std::min<long long>(m, jb0 + per);
  if (jb0 >= jb1) return;  // block-uniform
  // vertex of entry jb0: the last i with nindex[i] <= jb0, by a 64-way search
  if (threadIdx.x < 64) {
    const int lane = lane_id();
    long long lo = 0, hi = n;  // nindex[lo] <= jb0 < nindex[hi + 1] (nindex[n] = m > jb0)
    while (hi - lo > 63) {
      const long long span = hi - lo;
      const long long p = lo + span * lane / 64;
      const unsigned long long le = __ballot(nindex[p] <= jb0);
      const int L = 63 - __clzll((long long)le);  // lane 0 always qualifies
      const long long nlo = lo + span * L / 64;
      const long long nhi = L == 63 ? hi : lo + span * (L + 1) / 64 - 1;
      lo = nlo;
      hi = nhi < nlo ? nlo : nhi;
    }
    const long long p = lo + lane;
    const unsigned long long le = __ballot(p <= hi && nindex[p] <= jb0);
    if (lane == 0) s_i0 = (int)(lo + 63 - __clzll((long long)le));
  }
  __syncthreads();
  unsigned* const cnt = &ctrl->count[0];
  unsigned* const flags = &ctrl->flags;
  long long t0 = jb0;
  int iw = s_i0;
  int fill = 0;
  int par = 0;
  while (t0 < jb1) {
    // four consecutive entries per thread from a 16-byte aligned base; the
    // entries below t0 (at most three) are not this step's.  Their loads do
    // not depend on the window and are issued before it.
    const long long ta = t0 & ~3ll;
    const long long tmax = std::min<long long>(ta + kTile, jb1);
    const long long j0 = ta + 4 * (long long)threadIdx.x;
    int pk[kPer], pw[kPer];
    if (vec && j0 + 3 < tmax) {
      // the heavy pass streams the lists past its membership lookups into
      // a 2 MB bitmap; non-temporal loads leave more of L2 to the bitmap
      // (R-MAT-24: 2.38 -> 2.16 ms; the light pass gains nothing)
      typedef int v4i __attribute__((ext_vector_type(4)));
      const v4i* const k4p = reinterpret_cast<const v4i*>(nlist + j0);
      const v4i* const w4p = reinterpret_cast<const v4i*>(eweight + j0);
      const v4i k4 = HEAVY ? __builtin_nontemporal_load(k4p) : *k4p;
      const v4i w4 = HEAVY ? __builtin_nontemporal_load(w4p) : *w4p;
      pk[0] = k4.x; pk[1] = k4.y; pk[2] = k4.z; pk[3] = k4.w;
      pw[0] = w4.x; pw[1] = w4.y; pw[2] = w4.z; pw[3] = w4.w;
    } else {
#pragma unroll
      for (int q = 0; q < kPer; q++) {
        pk[q] = j0 + q < tmax ? nlist[j0 + q] : 0;
        pw[q] = j0 + q < tmax ? eweight[j0 + q] : 0;
      }
    }
    int* const W = s_idx[par];  // nindex[iw .. iw + kWin], one entry per thread
    W[threadIdx.x] = nindex[std::min<long long>((long long)iw + threadIdx.x, n)];
    __syncthreads();
    const long long te = std::max<long long>(t0, std::min<long long>(tmax, W[kWin]));
    if (te > t0) {  // block-uniform
      int4 it[kPer];
      bool ok[kPer];
      bool neg = false;
      int slot = 0, nxt = 0;
      if (j0 + 3 >= t0 && j0 < te) {
        slot = window_slot(W, std::max<long long>(j0, t0));
        nxt = W[slot + 1];  // slot < kWin: j < te <= W[kWin]
      }
#pragma unroll
      for (int q = 0; q < kPer; q++) {
        const long long j = j0 + q;
        ok[q] = false;
        it[q] = make_int4(0, 0, 0, 0);
        if (j >= t0 && j < te) {
          while (j >= nxt) nxt = W[++slot + 1];  // stops below kWin (j < W[kWin])
          const int i = iw + slot;
          const int k = pk[q];
          const int w = pw[q];
          neg |= w < 0;
          if (k > i && (HEAVY ? w > thr : w <= thr)) {
            if (HEAVY) {
              if (!(giant && member(gbits, i) && member(gbits, k))) {
                const int a = ld(&parent[i]), b = ld(&parent[k]);
                ok[q] = a != b;
                it[q] = make_int4(a, b, w, (int)j);
              }
            } else {
              ok[q] = true;
              it[q] = make_int4(i, k, w, (int)j);
            }
          }
        }
      }
      if (neg) __hip_atomic_fetch_or(flags, kFlagNegative, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      append_reg(A, fill, par, it, ok, wl, cnt, cap, flags);
    }
    const long long niw = (long long)iw + (te < W[kWin] ? window_slot(W, te) : kWin);
    t0 = te;
    iw = (int)std::min<long long>(niw, n);
    par ^= 1;
  }
  flush_reg(A, fill, wl, cnt, cap, flags);
}

// Proposes `key` to the minimum of `root` for the active lanes.  The lanes
// whose root is the first active lane's (neighbouring list entries of one
// vertex, and in later rounds the many edges into the largest tree) first
// take their minimum across the wavefront and send one atomic for all of
// them: every atomic on one address executes at the memory side, one after
// the other.  The other lanes propose on their own after it.  (Written with
// the combined proposal first and the others after it, this runs 1.9x
// faster on R-MAT-24 than the if/else form the compiler laid out the other
// way round: 1.91 -> 1.03 ms for the two proposal-only rounds.)  All lanes
// of the wavefront call it.
__device__ __forceinline__ void propose(unsigned long long* minv, int root, unsigned long long key,
                                        bool active) {
  const unsigned long long act = __ballot(active);
  if (!act) return;
  const int lane = lane_id();
  const int first = __ffsll((long long)act) - 1;
  const int r0 = __shfl(root, first);
  const bool same = active && root == r0;
  unsigned long long k = same ? key : kNoEdge;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(k, d);
    k = o < k ? o : k;
  }
  if (lane == first) min64(&minv[r0], k);
  const unsigned long long rest = act & ~__ballot(same);
  if ((rest >> lane) & 1ull) min64(&minv[root], key);
}

// propose(), and then the remaining lanes also combine runs of neighbouring
// lanes with one root: k_first lists a vertex's entries next to each other,
// so in the first round of a phase the lanes of a wavefront hold runs of
// one source vertex, and without this every lane of a run sends its own
// test and atomic to the same address (the first rounds: R-MAT-24 1.03 ->
// 0.90 ms, grid 256^3 1.26 -> 0.96 ms).  The later rounds' lists hold few
// such runs, and k_link became slower with it (3.49 -> 4.40 ms).
__device__ __forceinline__ void propose_runs(unsigned long long* minv, int root, unsigned long long key,
                                        bool active) {
  const unsigned long long act = __ballot(active);
  if (!act) return;
  const int lane = lane_id();
  const int first = __ffsll((long long)act) - 1;
  const int r0 = __shfl(root, first);
  const bool same = active && root == r0;
  unsigned long long k = same ? key : kNoEdge;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(k, d);
    k = o < k ? o : k;
  }
  if (lane == first) min64(&minv[r0], k);
  const unsigned long long rest = act & ~__ballot(same);
  if (!rest) return;
  // runs of neighbouring lanes with one root (list entries of one vertex):
  // a suffix minimum within each run, proposed by the run's first lane
  const bool mine = (rest >> lane) & 1ull;
  k = mine ? key : kNoEdge;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long o = __shfl_down(k, d);
    const int ro = __shfl_down(root, d);
    if (lane + d < 64 && ro == root && o < k) k = o;
  }
  const int rp = __shfl_up(root, 1);
  const bool prev = lane > 0 && ((rest >> (lane - 1)) & 1ull) && rp == root;
  if (mine && !prev) min64(&minv[root], k);
}

// kernel1 (:240-256): keeps the listed edges whose ends are still in
// different trees, as (root, root, w, j), and proposes each to both roots'
// minima.  Persistent blocks over contiguous ranges of the input list.
__global__ __launch_bounds__(kThreads) void k_link(const int4* in, const unsigned* in_cnt,
                                                   int4* out, unsigned* out_cnt, unsigned cap,
                                                   int* parent, unsigned long long* minv,
                                                   unsigned* flags, long long per) {
  __shared__ Appender A;
  const long long total = std::min<long long>(*in_cnt, cap);
  const long long b0 = (long long)blockIdx.x * per;
  const long long b1 = std::min<long long>(total, b0 + per);
  if (threadIdx.x == 0) A.fill = 0;
  if (b0 >= b1) return;  // block-uniform
  __syncthreads();
  for (long long t0 = b0; t0 < b1; t0 += kTile) {
    int4 it[kPer];
    bool ok[kPer];
#pragma unroll
    for (int r = 0; r < kPer; r++) {
      const long long idx = t0 + threadIdx.x + r * kThreads;
      ok[r] = false;
      it[r] = make_int4(0, 0, 0, 0);
      if (idx < b1) {
        int4 e = in[idx];
        e.x = find(e.x, parent);
        e.y = find(e.y, parent);
        if (e.x != e.y) {
          ok[r] = true;
          it[r] = e;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < kPer; r++) {
      const unsigned long long key = key_of(it[r]);
      propose(minv, it[r].x, key, ok[r]);
      propose(minv, it[r].y, key, ok[r]);
    }
    append(A, it, ok, out, out_cnt, cap, flags);
  }
  flush(A, out, out_cnt, cap, flags);
}

// The first round of a phase (kernel1 :240-256 on the list initWL wrote):
// every listed end is already a root -- the light phase starts from single
// vertices, and the heavy phase lists the roots of a flattened forest with
// no union since -- and the two ends of every item differ.  kernel1 would
// find each end's root (two gathers) and write the whole list out again
// unchanged, so this round only proposes each item to both minima, and
// kernel2 reads the list initWL wrote (R-MAT-24 forest 12.7 -> 12.5 ms,
// grid 256^3 10.0 -> 9.2 ms).
__global__ __launch_bounds__(kThreads) void k_propose(const int4* in, const unsigned* in_cnt,
                                                      unsigned cap, unsigned long long* minv) {
  const long long total = std::min<long long>(*in_cnt, cap);
  const long long stride = (long long)gridDim.x * kThreads;
  // block-uniform loop: propose() takes all lanes of the wavefront
  for (long long base = (long long)blockIdx.x * kThreads; base < total; base += stride) {
    const long long i = base + threadIdx.x;
    const bool ok = i < total;
    const int4 e = ok ? in[i] : make_int4(0, 0, 0, 0);
    const unsigned long long key = key_of(e);
    propose_runs(minv, e.x, key, ok);
    propose_runs(minv, e.y, key, ok);
  }
}

// kernel2 (:259-271): an edge that is the minimum of either of its trees
// joins them and is in the forest.
__global__ __launch_bounds__(kThreads) void k_join(const int4* wl, const unsigned* cnt,
                                                   unsigned cap, int* parent,
                                                   const unsigned long long* minv,
                                                   unsigned char* in_mst) {
  const long long total = std::min<long long>(*cnt, cap);
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < total; i += stride) {
    const int4 e = wl[i];
    const unsigned long long key = key_of(e);
    if (key == ld64(&minv[e.x]) || key == ld64(&minv[e.y])) {
      join(e.x, e.y, parent);
      in_mst[e.w] = 1;  // each edge is listed once: one writer per byte
    }
  }
}

// kernel3 (:274-282) resets the minima of the listed edges' trees, two
// scattered stores per edge; the edges of a large tree all name its root,
// and agent-scope stores to one address serialise at the memory side
// (R-MAT-24: 10.4 ms per round; 0.3-0.7 ms with a load before each store).
// Here the whole array is reset with coalesced 16-byte stores (8n bytes,
// ~25 us for 2^24 vertices).  No other kernel runs meanwhile.
__global__ __launch_bounds__(kThreads) void k_reset(int n, unsigned long long* minv) {
  const long long stride = (long long)gridDim.x * kThreads;
  uint4* w = reinterpret_cast<uint4*>(minv);  // minv is 256-byte aligned
  const uint4 none = make_uint4(~0u, ~0u, ~0u, ~0u);
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n / 2; i += stride) w[i] = none;
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) minv[n - 1] = kNoEdge;
}

// Root of the largest tree after the light phase (parents flattened): the
// most frequent root among kSamples evenly spaced vertices, if it occurs
// at least twice; ties go to the smaller root.
__global__ __launch_bounds__(kSamples) void k_pick(int n, const int* parent, Ctrl* ctrl) {
  __shared__ int s[kSamples];
  __shared__ unsigned long long s_best[kSamples / 64];
  const int t = threadIdx.x;
  const int r = parent[(long long)t * n / kSamples];  // no union runs during this kernel
  s[t] = r;
  __syncthreads();
  unsigned c = 0;
  for (int k = 0; k < kSamples; k++) c += s[k] == r;
  unsigned long long key = ((unsigned long long)c << 32) | (unsigned)(INT_MAX - r);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(key, d);
    key = o > key ? o : key;
  }
  if (lane_id() == 0) s_best[t >> 6] = key;
  __syncthreads();
  if (t == 0) {
    unsigned long long best = 0;
    for (int k = 0; k < kSamples / 64; k++) best = s_best[k] > best ? s_best[k] : best;
    ctrl->giant = (best >> 32) >= 2 ? INT_MAX - (int)(unsigned)best : -1;
  }
}

// gbits[v] = (v's root is the largest tree's root), one ballot word per 64
// vertices; with it the heavy phase drops the edges inside that tree after
// two lookups in an array of n/8 bytes instead of gathering both roots.
__global__ __launch_bounds__(kThreads) void k_members(int n, const int* parent, const Ctrl* ctrl,
                                                      unsigned long long* gbits) {
  const int R = ctrl->giant;
  const long long stride = (long long)gridDim.x * kThreads;
  const long long words = ((long long)n + 63) / 64;
  for (long long base = (long long)blockIdx.x * kThreads; base < words * 64; base += stride) {
    const long long v = base + threadIdx.x;
    const unsigned long long b = __ballot(R >= 0 && v < n && parent[v] == R);
    if (lane_id() == 0 && (v >> 6) < words) gbits[v >> 6] = b;
  }
}

// Every vertex pointed at its root (between the phases).
__global__ __launch_bounds__(kThreads) void k_flatten(int n, int* parent) {
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long v = (long long)blockIdx.x * kThreads + threadIdx.x; v < n; v += stride) {
    const int p = ld(&parent[v]);
    if (p != v) {
      const int r = find(p, parent);
      if (r != p) st(&parent[v], r);
    }
  }
}

__global__ void k_clear_count(unsigned* c) { *c = 0u; }

// ---- host side ---------------------------------------------------------------------

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, int code, const char* what) {
  if (e == hipSuccess) return ECLMST_OK;
  return fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

// Host exceptions never cross the C ABI: every entry point that can meet
// one (a host allocation that fails) turns it into a status here.
int host_exception() {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    return fail(ECLMST_ERR_ALLOC, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(ECLMST_ERR_ALLOC, std::string("host resources: ") + e.what());
  } catch (...) {
    return fail(ECLMST_ERR_ALLOC, "host error");
  }
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

unsigned list_cap(long long m) { return (unsigned)std::min<long long>(m / 2 + 64, 0xFFFFFF00ll); }

struct Layout {
  Ctrl* ctrl;
  int* parent;
  unsigned long long* minv;
  unsigned long long* gbits;
  int4* wl[2];
  unsigned cap;
};

size_t layout_bytes(long long n, long long m) {
  size_t b = 256;
  b += align_up(4 * (size_t)n, 256);
  b += align_up(8 * (size_t)n, 256);
  b += align_up(8 * (size_t)((n + 63) / 64), 256);
  b += 2 * align_up(sizeof(int4) * (size_t)list_cap(m), 256);
  return b;
}

Layout carve(void* base, long long n, long long m) {
  char* p = static_cast<char*>(base);
  Layout L;
  L.ctrl = reinterpret_cast<Ctrl*>(p);
  p += 256;
  L.parent = reinterpret_cast<int*>(p);
  p += align_up(4 * (size_t)n, 256);
  L.minv = reinterpret_cast<unsigned long long*>(p);
  p += align_up(8 * (size_t)n, 256);
  L.gbits = reinterpret_cast<unsigned long long*>(p);
  p += align_up(8 * (size_t)((n + 63) / 64), 256);
  L.cap = list_cap(m);
  for (int k = 0; k < 2; k++) {
    L.wl[k] = reinterpret_cast<int4*>(p);
    p += align_up(sizeof(int4) * (size_t)L.cap, 256);
  }
  return L;
}

int cu_count() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    return 256;
  return cus;
}

// Small pinned host buffer for the per-round control reads, allocated on
// first use and held until the process ends.
Ctrl* pinned_ctrl() {
  static Ctrl* h = nullptr;
  static std::once_flag once;
  std::call_once(once, [] {
    void* p = nullptr;
    if (hipHostMalloc(&p, sizeof(Ctrl) * 64, hipHostMallocDefault) == hipSuccess) h = (Ctrl*)p;
  });
  return h;
}

struct Runner {
  int n;
  long long m;
  const int* nindex;
  const int* nlist;
  const int* eweight;
  unsigned char* in_mst;
  Layout L;
  hipStream_t s;
  int blocks;  // persistent grid
  Ctrl* h;     // pinned (per call slot)

  int read_ctrl(Ctrl* out) {
    int rc = hip_fail(hipMemcpyAsync(h, L.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s),
                      ECLMST_ERR_COPY, "copying from device failed");
    if (!rc) rc = hip_fail(hipStreamSynchronize(s), ECLMST_ERR_LAUNCH, "kernel execution failed");
    if (!rc) *out = *h;
    return rc;
  }

  // Boruvka rounds on list `a` (its size already in ctrl->count[a]).  The
  // first round needs no finds and no new list (k_propose), and is queued
  // without reading a size back.  After its unions the forest is flattened
  // once: that round joins the most trees, and the CAS links it leaves form
  // chains that the second round, still on nearly the whole list, would
  // otherwise walk for every edge (R-MAT-24: 13.8 -> 12.85 ms per forest;
  // flattening after every round costs more than it saves).
  int rounds(int a, unsigned cnt) {
    if (cnt > 0) {  // first round: proposals straight from list a (k_propose)
      const int g = (int)std::min<long long>((cnt + kThreads - 1) / kThreads, (long long)blocks * 8);
      k_propose<<<g, kThreads, 0, s>>>(L.wl[a], &L.ctrl->count[a], L.cap, L.minv);
      k_join<<<g, kThreads, 0, s>>>(L.wl[a], &L.ctrl->count[a], L.cap, L.parent, L.minv, in_mst);
      const int gr = (int)std::min<long long>((n / 2 + kThreads - 1) / kThreads, (long long)blocks * 8);
      k_reset<<<std::max(gr, 1), kThreads, 0, s>>>(n, L.minv);
      const int gf = (int)std::min<long long>(((long long)n + kThreads - 1) / kThreads, (long long)blocks * 8);
      k_flatten<<<std::max(gf, 1), kThreads, 0, s>>>(n, L.parent);
    }
    while (cnt > 0) {
      const int b = 1 - a;
      k_clear_count<<<1, 1, 0, s>>>(&L.ctrl->count[b]);
      const long long tiles = ((long long)cnt + kTile - 1) / kTile;
      const int nb = (int)std::max<long long>(1, std::min<long long>(blocks, tiles));
      const long long per = (tiles + nb - 1) / nb * kTile;
      k_link<<<nb, kThreads, 0, s>>>(L.wl[a], &L.ctrl->count[a], L.wl[b], &L.ctrl->count[b], L.cap,
                                     L.parent, L.minv, &L.ctrl->flags, per);
      Ctrl c;
      int rc = read_ctrl(&c);
      if (rc) return rc;
      if (c.flags & kFlagOverflow || c.count[b] > L.cap)
        return fail(ECLMST_ERR_FORMAT, "edge worklist overflow (graph not symmetric?)");
      cnt = c.count[b];
      if (cnt > 0) {
        const int g = (int)std::min<long long>((cnt + kThreads - 1) / kThreads, (long long)blocks * 8);
        k_join<<<g, kThreads, 0, s>>>(L.wl[b], &L.ctrl->count[b], L.cap, L.parent, L.minv, in_mst);
        const int gr = (int)std::min<long long>((n / 2 + kThreads - 1) / kThreads, (long long)blocks * 8);
        k_reset<<<std::max(gr, 1), kThreads, 0, s>>>(n, L.minv);
      }
      a = b;
    }
    return hip_fail(hipGetLastError(), ECLMST_ERR_LAUNCH, "kernel launch failed");
  }

  template <bool HEAVY>
  int phase() {
    k_clear_count<<<1, 1, 0, s>>>(&L.ctrl->count[0]);
    const long long tiles = (m + kTile - 1) / kTile;
    const int nb = (int)std::max<long long>(1, std::min<long long>(blocks, tiles));
    const long long per = (tiles + nb - 1) / nb * kTile;
    const bool vec = ((reinterpret_cast<uintptr_t>(nlist) | reinterpret_cast<uintptr_t>(eweight)) & 15) == 0;
    k_first<HEAVY><<<nb, kThreads, 0, s>>>(n, m, nindex, nlist, eweight, L.parent, L.gbits, L.ctrl, L.wl[0],
                                           L.cap, per, vec);
    Ctrl c;
    int rc = read_ctrl(&c);
    if (rc) return rc;
    if (c.flags & kFlagNegative) return fail(ECLMST_ERR_FORMAT, "negative edge weight");
    if (c.flags & kFlagOverflow || c.count[0] > L.cap)
      return fail(ECLMST_ERR_FORMAT, "edge worklist overflow (graph not symmetric?)");
    return rounds(0, c.count[0]);
  }

  int run(int threshold, float* device_ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (device_ms) {
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
      (void)hipEventRecord(e0, s);
    }
    const int g = (int)std::min<long long>(((long long)std::max<long long>(n, m / 16) + kThreads - 1) / kThreads,
                                           (long long)blocks * 8);
    k_init<<<std::max(g, 1), kThreads, 0, s>>>(n, L.parent, L.minv, m, in_mst, L.ctrl);
    int thr = threshold;
    int rc = ECLMST_OK;
    if (thr < 0) {
      if (m / n >= 4 && m >= kSamples) {
        k_threshold<<<1, kSamples, 0, s>>>(n, m, eweight, L.ctrl);
        Ctrl c;
        rc = read_ctrl(&c);
        thr = c.threshold;
      } else {
        thr = INT_MAX;
      }
    }
    if (!rc) {
      // the kernels read the threshold from the control block
      rc = hip_fail(hipMemcpyAsync(&L.ctrl->threshold, &thr, sizeof(int), hipMemcpyHostToDevice, s),
                    ECLMST_ERR_COPY, "copying to device failed");
      if (!rc) rc = hip_fail(hipStreamSynchronize(s), ECLMST_ERR_LAUNCH, "kernel execution failed");
    }
    if (!rc) rc = phase<false>();
    if (!rc && thr != INT_MAX) {
      const int gf = (int)std::min<long long>((n + kThreads - 1) / kThreads, (long long)blocks * 8);
      k_flatten<<<std::max(gf, 1), kThreads, 0, s>>>(n, L.parent);
      k_pick<<<1, kSamples, 0, s>>>(n, L.parent, L.ctrl);
      const int gm = (int)std::min<long long>(((n + 63) / 64 * 64 + kThreads - 1) / kThreads, (long long)blocks * 8);
      k_members<<<std::max(gm, 1), kThreads, 0, s>>>(n, L.parent, L.ctrl, L.gbits);
      rc = phase<true>();
    }
    if (device_ms) {
      (void)hipEventRecord(e1, s);
      if (!rc) rc = hip_fail(hipEventSynchronize(e1), ECLMST_ERR_LAUNCH, "kernel execution failed");
      float ms = 0.f;
      if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
      *device_ms = ms;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
    if (!rc) rc = hip_fail(hipStreamSynchronize(s), ECLMST_ERR_LAUNCH, "kernel execution failed");
    return rc;
  }
};

int check_args(int32_t nodes, int32_t edges, const void* nindex, const void* nlist,
               const void* eweight, const void* out) {
  if (nodes < 1 || edges < 0) return fail(ECLMST_ERR_ARG, "node or edge count too low");
  if (!nindex || !out || (edges > 0 && (!nlist || !eweight)))
    return fail(ECLMST_ERR_ARG, "null array");
  return ECLMST_OK;
}

std::mutex g_pin_mu;  // the pinned control buffer is shared by the calls of the process

int run_device(int32_t nodes, int32_t edges, const int32_t* d_nindex, const int32_t* d_nlist,
               const int32_t* d_eweight, uint8_t* d_in_mst, void* d_work, size_t work_bytes,
               int32_t threshold, hipStream_t s, float* device_ms) {
  int rc = check_args(nodes, edges, d_nindex, d_nlist, d_eweight, d_in_mst);
  if (rc) return rc;
  if (!d_work || work_bytes < layout_bytes(nodes, edges))
    return fail(ECLMST_ERR_WORKSPACE, "workspace too small");
  if ((reinterpret_cast<uintptr_t>(d_work) & 15) != 0)
    return fail(ECLMST_ERR_ARG, "workspace not 16-byte aligned");
  std::lock_guard<std::mutex> lock(g_pin_mu);
  Ctrl* h = pinned_ctrl();
  if (!h) return fail(ECLMST_ERR_ALLOC, "could not allocate pinned memory");
  Runner r{nodes, edges, d_nindex, d_nlist, d_eweight, d_in_mst, carve(d_work, nodes, edges), s,
           cu_count() * 8, h};
  if (edges == 0) {
    k_init<<<1, kThreads, 0, s>>>(nodes, r.L.parent, r.L.minv, 0, d_in_mst, r.L.ctrl);
    if (device_ms) *device_ms = 0.f;
    return hip_fail(hipStreamSynchronize(s), ECLMST_ERR_LAUNCH, "kernel execution failed");
  }
  return r.run(threshold, device_ms);
}

}  // namespace eclmst

using namespace eclmst;

extern "C" {

const char* eclmst_last_error(void) { return g_err.c_str(); }

const char* eclmst_version(void) { return "eclmst-mi355x 0.1 (gfx950, wave64)"; }

int32_t eclmst_threshold(int32_t nodes, int32_t edges, const int32_t* eweight) {
  if (nodes < 1 || edges < kSamples || edges / nodes < 4 || !eweight) return INT_MAX;
  int s[kSamples];
  for (int t = 0; t < kSamples; t++) s[t] = eweight[(long long)t * edges / kSamples];
  std::sort(s, s + kSamples);
  const int idx = (int)std::min<double>(kSamples - 1, 3.0 * nodes * kSamples / (double)edges);
  return s[idx];
}

size_t eclmst_workspace_bytes(int32_t nodes, int32_t edges) try {
  if (nodes < 1 || edges < 0) return 0;
  return layout_bytes(nodes, edges);
} catch (...) {
  return (host_exception(), size_t(0));
}

int eclmst_run_device(int32_t nodes, int32_t edges, const int32_t* d_nindex, const int32_t* d_nlist,
                      const int32_t* d_eweight, uint8_t* d_in_mst, void* d_work, size_t work_bytes,
                      int32_t threshold, void* stream, float* device_ms) try {
  return run_device(nodes, edges, d_nindex, d_nlist, d_eweight, d_in_mst, d_work, work_bytes,
                    threshold, static_cast<hipStream_t>(stream), device_ms);
} catch (...) {
  return host_exception();
}

static int compute(int32_t nodes, int32_t edges, const int32_t* nindex, const int32_t* nlist,
                   const int32_t* eweight, uint8_t* in_mst, double* seconds) {
  int rc = check_args(nodes, edges, nindex, nlist, eweight, in_mst);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(ECLMST_ERR_DEVICE, "there is no HIP device");
  int* d_ni = nullptr;
  int* d_nl = nullptr;
  int* d_w = nullptr;
  unsigned char* d_out = nullptr;
  void* d_work = nullptr;
  hipStream_t s = nullptr;
  const size_t wb = layout_bytes(nodes, edges);
  auto cleanup = [&]() {
    if (d_ni) (void)hipFree(d_ni);
    if (d_nl) (void)hipFree(d_nl);
    if (d_w) (void)hipFree(d_w);
    if (d_out) (void)hipFree(d_out);
    if (d_work) (void)hipFree(d_work);
    if (s) (void)hipStreamDestroy(s);
  };
  const size_t me = std::max<size_t>((size_t)edges, 1);
  if (hipMalloc((void**)&d_ni, 4 * ((size_t)nodes + 1)) != hipSuccess ||
      hipMalloc((void**)&d_nl, 4 * me) != hipSuccess || hipMalloc((void**)&d_w, 4 * me) != hipSuccess ||
      hipMalloc((void**)&d_out, me) != hipSuccess || hipMalloc(&d_work, wb) != hipSuccess) {
    cleanup();
    return fail(ECLMST_ERR_ALLOC, "could not allocate device memory");
  }
  rc = hip_fail(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), ECLMST_ERR_DEVICE,
                "stream creation failed");
  // the three host arrays in one pass through pinned staging (eclcc_stage.hip)
  const eclstage::CopySpan in[3] = {{d_ni, const_cast<int32_t*>(nindex), 4 * ((size_t)nodes + 1)},
                                    {d_nl, const_cast<int32_t*>(nlist), 4 * (size_t)edges},
                                    {d_w, const_cast<int32_t*>(eweight), 4 * (size_t)edges}};
  if (!rc) rc = hip_fail(eclstage::to_device(in, 3, s), ECLMST_ERR_COPY, "copying to device failed");
  const int32_t thr = eclmst_threshold(nodes, edges, eweight);
  float ms = 0.f;
  if (!rc) rc = run_device(nodes, edges, d_ni, d_nl, d_w, d_out, d_work, wb, thr, s, &ms);
  if (!rc) rc = hip_fail(hipStreamSynchronize(s), ECLMST_ERR_LAUNCH, "kernel execution failed");
  if (!rc && edges > 0)
    rc = hip_fail(eclstage::to_host(in_mst, d_out, (size_t)edges, s), ECLMST_ERR_COPY, "copying from device failed");
  if (seconds) *seconds = 1e-3 * ms;
  cleanup();
  return rc;
}

int eclmst_compute(int32_t nodes, int32_t edges, const int32_t* nindex, const int32_t* nlist,
                   const int32_t* eweight, uint8_t* in_mst) try {
  return compute(nodes, edges, nindex, nlist, eweight, in_mst, nullptr);
} catch (...) {
  return host_exception();
}

int eclmst_compute_timed(int32_t nodes, int32_t edges, const int32_t* nindex,
                         const int32_t* nlist, const int32_t* eweight, uint8_t* in_mst,
                         double* seconds) try {
  return compute(nodes, edges, nindex, nlist, eweight, in_mst, seconds);
} catch (...) {
  return host_exception();
}

}  // extern "C"
