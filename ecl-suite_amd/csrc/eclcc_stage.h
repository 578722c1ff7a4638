// eclcc_stage.h -- staged copies between pageable host arrays and the device
// (eclcc_stage.hip), used by the drop-in calls of libeclcc.so and
// libeclmst.so.  Hidden: each library holds its own staging pool.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace eclstage {

// `bytes` between device memory `dev` and pageable host memory `host`.
struct CopySpan {
  void* dev;
  void* host;
  size_t bytes;
};

// Copies the spans to the device, ordered after the work already on `s`.
// Large totals go through pinned staging slots on several streams and are
// complete on return; small ones are queued on `s` as plain copies.
__attribute__((visibility("hidden"))) hipError_t to_device(const CopySpan* spans, int nspans, hipStream_t s);

// Copies `bytes` from the device to pageable host memory after the work on
// `s`; complete on return.
__attribute__((visibility("hidden"))) hipError_t to_host(void* host, const void* dev, size_t bytes, hipStream_t s);

}  // namespace eclstage
