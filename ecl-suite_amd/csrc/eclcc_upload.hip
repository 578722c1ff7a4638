// eclcc_upload.hip -- host-to-device copies of large pageable arrays for the
// drop-in call (eclcc_compute), which takes the caller's host CSR as the
// reference's computeCC does (ECL-CC-atomic.cu:295-296).
//
// A copy from pageable memory is staged by the runtime and ran at 27-30
// GB/s on the MI355X box (2.08 GB of R-MAT-24 entries in 70-76 ms), while a
// copy from pinned memory runs at 57.6 GB/s.  Pinning the caller's array in
// place (hipHostRegister) is not available here, so large copies go through
// a pool of pinned staging slots instead: kStageThreads host threads each
// copy every kStageThreads-th chunk of the source into one of their own
// slots and send it to the device on their own stream, so the host copies
// of some chunks overlap the DMA of others.  The pool (kStageThreads x
// kStageSlots x kStageChunk = 128 MiB of pinned memory, and one stream per
// thread) is made on the first large copy and held until the process ends;
// one large copy runs at a time.  Small copies, and any copy when the pool
// cannot be made, use the plain pageable path.

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "eclcc_internal.h"

namespace eclcc {
namespace {

constexpr size_t kStageChunk = size_t(8) << 20;
constexpr int kStageThreads = 8;
constexpr int kStageSlots = 2;
// below this, one pageable copy costs less than waking the pool
constexpr size_t kStageMinBytes = size_t(64) << 20;

struct StagePool {
  std::mutex mu;  // one staged copy at a time
  bool tried = false;
  bool ok = false;
  int device = -1;
  char* buf = nullptr;
  hipStream_t streams[kStageThreads] = {};
  hipEvent_t done[kStageThreads][kStageSlots] = {};
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
  for (int t = 0; t < kStageThreads; t++) {
    if (hipStreamCreateWithFlags(&p.streams[t], hipStreamNonBlocking) != hipSuccess) return false;
    for (int k = 0; k < kStageSlots; k++)
      if (hipEventCreateWithFlags(&p.done[t][k], hipEventDisableTiming) != hipSuccess) return false;
  }
  p.ok = true;
  return true;
}

}  // namespace

int upload_host(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return ECLCC_OK;
  StagePool& p = pool();
  std::unique_lock<std::mutex> lock(p.mu, std::defer_lock);
  int dev = -1;
  bool staged = bytes >= kStageMinBytes && hipGetDevice(&dev) == hipSuccess;
  if (staged) {
    lock.lock();
    staged = make_pool(p) && p.device == dev;
    if (!staged) lock.unlock();
  }
  if (!staged)
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), ECLCC_ERR_COPY,
                     "copying to device failed");
  // work queued on `s` before the call finishes before the staged copies
  int rc = check_hip(hipStreamSynchronize(s), ECLCC_ERR_COPY, "copying to device failed");
  if (rc) return rc;
  const size_t nchunks = (bytes + kStageChunk - 1) / kStageChunk;
  std::atomic<int> err{hipSuccess};
  auto worker = [&](int t) {
    (void)hipSetDevice(p.device);
    size_t round = 0;
    for (size_t c = (size_t)t; c < nchunks && err.load() == hipSuccess; c += kStageThreads, round++) {
      const int k = (int)(round % kStageSlots);
      char* slot = p.buf + ((size_t)t * kStageSlots + k) * kStageChunk;
      hipError_t e = hipEventSynchronize(p.done[t][k]);  // the slot's last DMA
      const size_t off = c * kStageChunk;
      const size_t len = bytes - off < kStageChunk ? bytes - off : kStageChunk;
      if (e == hipSuccess) {
        std::memcpy(slot, static_cast<const char*>(src) + off, len);
        e = hipMemcpyAsync(static_cast<char*>(dst) + off, slot, len, hipMemcpyHostToDevice, p.streams[t]);
      }
      if (e == hipSuccess) e = hipEventRecord(p.done[t][k], p.streams[t]);
      if (e != hipSuccess) {
        int expect = hipSuccess;
        err.compare_exchange_strong(expect, (int)e);
      }
    }
    const hipError_t e = hipStreamSynchronize(p.streams[t]);
    if (e != hipSuccess) {
      int expect = hipSuccess;
      err.compare_exchange_strong(expect, (int)e);
    }
  };
  std::vector<std::thread> th;
  th.reserve(kStageThreads - 1);
  try {
    for (int t = 1; t < kStageThreads; t++) th.emplace_back(worker, t);
  } catch (...) {
    err.store(hipErrorOutOfMemory);  // the started threads stop at their next chunk
  }
  worker(0);
  for (std::thread& x : th) x.join();
  if (err.load() != hipSuccess)
    return check_hip((hipError_t)err.load(), ECLCC_ERR_COPY, "copying to device failed");
  return ECLCC_OK;
}

}  // namespace eclcc
