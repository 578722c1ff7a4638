// eclcc_stage.hip -- copies between large pageable host arrays and the
// device for the drop-in calls, which take the caller's host arrays and fill
// the caller's host output as the reference's computeCC and gpuMST do
// (ECL-CC-atomic.cu:295-296, 318; ECL-MST_10-atomic.cu:292-297, 365).  Built
// into libeclcc.so and libeclmst.so, each with a pool of its own (the
// functions are hidden, eclcc_stage.h).
//
// On the MI355X box a copy from pageable memory is staged by the runtime and
// ran at 27-30 GB/s (2.08 GB of R-MAT-24 entries in 70-76 ms), a copy from
// pinned memory at 57.6 GB/s, and a copy of 64 MiB of labels back into
// pageable memory at about 7 GB/s (9.5 ms).  Pinning the caller's arrays in
// place (hipHostRegister) is not available here, so large copies go through
// a pool of pinned staging slots instead: kStageThreads host threads each
// take every kStageThreads-th chunk and move it through one of their own
// slots on their own stream, so the host copies of some chunks overlap the
// DMA of others.  The pool (kStageThreads x kStageSlots x kStageChunk = 128
// MiB of pinned memory, and two streams per thread) is made on the first
// large copy and held until the process ends; one staged copy runs at a
// time.  Small copies, and any copy when the pool cannot be made, use the
// plain pageable path.

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "eclcc_stage.h"

namespace eclstage {
namespace {

constexpr size_t kStageChunk = size_t(8) << 20;
constexpr int kStageThreads = 8;
constexpr int kStageSlots = 2;
// below these, one pageable copy costs less than waking the pool (a
// pageable copy back to the host is four times slower than one to the
// device, so downloads are staged from a smaller size)
constexpr size_t kStageMinBytes = size_t(64) << 20;
constexpr size_t kStageMinDownBytes = size_t(8) << 20;

struct StagePool {
  std::mutex mu;  // one staged copy at a time
  bool tried = false;
  bool ok = false;
  int device = -1;
  char* buf = nullptr;
  // uploads and downloads use streams and events of their own (the slots
  // are shared: every staged copy is complete before the next begins)
  hipStream_t streams[2][kStageThreads] = {};
  hipEvent_t done[2][kStageThreads][kStageSlots] = {};
};

StagePool& pool() {
  static StagePool p;
  return p;
}

// Makes the pool on the current device (under p.mu); false if it cannot.
bool make_pool(StagePool& p) {
  if (p.tried) return p.ok;
  p.tried = true;
  if (hipGetDevice(&p.device) != hipSuccess) return false;
  void* b = nullptr;
  if (hipHostMalloc(&b, kStageChunk * kStageThreads * kStageSlots, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  p.buf = static_cast<char*>(b);
  for (int d = 0; d < 2; d++)
    for (int t = 0; t < kStageThreads; t++) {
      if (hipStreamCreateWithFlags(&p.streams[d][t], hipStreamNonBlocking) != hipSuccess) return false;
      for (int k = 0; k < kStageSlots; k++)
        if (hipEventCreateWithFlags(&p.done[d][t][k], hipEventDisableTiming) != hipSuccess) return false;
    }
  p.ok = true;
  return true;
}

// One chunk of a staged copy: `len` bytes between host and device.
struct Chunk {
  char* dev;
  char* host;
  size_t len;
};

std::vector<Chunk> cut(const CopySpan* spans, int nspans) {
  std::vector<Chunk> out;
  for (int i = 0; i < nspans; i++)
    for (size_t off = 0; off < spans[i].bytes; off += kStageChunk)
      out.push_back(Chunk{static_cast<char*>(spans[i].dev) + off, static_cast<char*>(spans[i].host) + off,
                          spans[i].bytes - off < kStageChunk ? spans[i].bytes - off : kStageChunk});
  return out;
}

void note(std::atomic<int>& err, hipError_t e) {
  if (e == hipSuccess) return;
  int expect = hipSuccess;
  err.compare_exchange_strong(expect, (int)e);
}

// Runs worker(t) on kStageThreads threads (this one included); the first
// error any of them noted.
template <class F>
hipError_t run_workers(std::atomic<int>& err, F worker) {
  std::vector<std::thread> th;
  try {
    th.reserve(kStageThreads - 1);
    for (int t = 1; t < kStageThreads; t++) th.emplace_back(worker, t);
  } catch (...) {
    err.store(hipErrorOutOfMemory);  // the started threads stop at their next chunk
  }
  worker(0);
  for (std::thread& x : th) x.join();
  return (hipError_t)err.load();
}

// Locks and makes the pool for a copy of `bytes` (total) at least `min`
// bytes long; false (unlocked) when the plain path is to be taken.
bool take_pool(StagePool& p, std::unique_lock<std::mutex>& lock, size_t bytes, size_t min) {
  int dev = -1;
  if (bytes < min || hipGetDevice(&dev) != hipSuccess) return false;
  lock.lock();
  if (make_pool(p) && p.device == dev) return true;
  lock.unlock();
  return false;
}

}  // namespace

hipError_t to_device(const CopySpan* spans, int nspans, hipStream_t s) {
  size_t total = 0;
  for (int i = 0; i < nspans; i++) total += spans[i].bytes;
  if (total == 0) return hipSuccess;
  StagePool& p = pool();
  std::unique_lock<std::mutex> lock(p.mu, std::defer_lock);
  if (!take_pool(p, lock, total, kStageMinBytes)) {
    for (int i = 0; i < nspans; i++)
      if (spans[i].bytes) {
        const hipError_t e = hipMemcpyAsync(spans[i].dev, spans[i].host, spans[i].bytes, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
      }
    return hipSuccess;
  }
  // work queued on `s` before the call finishes before the staged copies
  const hipError_t e0 = hipStreamSynchronize(s);
  if (e0 != hipSuccess) return e0;
  const std::vector<Chunk> chunks = cut(spans, nspans);
  std::atomic<int> err{hipSuccess};
  return run_workers(err, [&](int t) {
    (void)hipSetDevice(p.device);
    size_t round = 0;
    for (size_t c = (size_t)t; c < chunks.size() && err.load() == hipSuccess; c += kStageThreads, round++) {
      const int k = (int)(round % kStageSlots);
      char* slot = p.buf + ((size_t)t * kStageSlots + k) * kStageChunk;
      hipError_t e = hipEventSynchronize(p.done[0][t][k]);  // the slot's last DMA
      if (e == hipSuccess) {
        std::memcpy(slot, chunks[c].host, chunks[c].len);
        e = hipMemcpyAsync(chunks[c].dev, slot, chunks[c].len, hipMemcpyHostToDevice, p.streams[0][t]);
      }
      if (e == hipSuccess) e = hipEventRecord(p.done[0][t][k], p.streams[0][t]);
      note(err, e);
    }
    note(err, hipStreamSynchronize(p.streams[0][t]));
  });
}

hipError_t to_host(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipStreamSynchronize(s);
  StagePool& p = pool();
  std::unique_lock<std::mutex> lock(p.mu, std::defer_lock);
  if (!take_pool(p, lock, bytes, kStageMinDownBytes)) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
  }
  // the data is written by work on `s`
  const hipError_t e0 = hipStreamSynchronize(s);
  if (e0 != hipSuccess) return e0;
  const CopySpan one{const_cast<void*>(src), dst, bytes};
  const std::vector<Chunk> chunks = cut(&one, 1);
  std::atomic<int> err{hipSuccess};
  return run_workers(err, [&](int t) {
    (void)hipSetDevice(p.device);
    // this thread's chunks are t, t + T, t + 2T, ...; chunk i of them moves
    // through slot i mod kStageSlots, its DMA issued kStageSlots chunks ahead
    auto issue = [&](size_t i) {
      const size_t c = (size_t)t + i * kStageThreads;
      char* slot = p.buf + ((size_t)t * kStageSlots + i % kStageSlots) * kStageChunk;
      hipError_t e = hipMemcpyAsync(slot, chunks[c].dev, chunks[c].len, hipMemcpyDeviceToHost, p.streams[1][t]);
      if (e == hipSuccess) e = hipEventRecord(p.done[1][t][i % kStageSlots], p.streams[1][t]);
      note(err, e);
    };
    const size_t mine = chunks.size() > (size_t)t ? (chunks.size() - (size_t)t + kStageThreads - 1) / kStageThreads : 0;
    for (size_t i = 0; i < mine && i < (size_t)kStageSlots; i++) issue(i);
    for (size_t i = 0; i < mine && err.load() == hipSuccess; i++) {
      const size_t c = (size_t)t + i * kStageThreads;
      const hipError_t e = hipEventSynchronize(p.done[1][t][i % kStageSlots]);
      note(err, e);
      if (e != hipSuccess) break;
      std::memcpy(chunks[c].host, p.buf + ((size_t)t * kStageSlots + i % kStageSlots) * kStageChunk,
                  chunks[c].len);
      if (i + kStageSlots < mine) issue(i + kStageSlots);
    }
    note(err, hipStreamSynchronize(p.streams[1][t]));
  });
}

}  // namespace eclstage
