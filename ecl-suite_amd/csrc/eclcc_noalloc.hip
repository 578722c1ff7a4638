// eclcc_noalloc.hip -- eclcc_run_device_noalloc (include/eclcc.h): the
// device-resident run for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  It has no counterpart in the reference, which allocates on
// every call (ECL-CC-atomic.cu:290-293).
//
// The workspace of eclcc_run_device (bit words, hub task list, counters)
// lives here in one fixed device array, `scratch`, of
// ECLCC_NOALLOC_SCRATCH_BYTES.  Each GPU the library runs on holds its own
// copy until the process ends.  In host code `scratch` only names the
// symbol: its device address, per device, comes from hipGetSymbolAddress on
// the first call made with that device current, and only that address is
// passed to the kernels.  Nothing in the workspace has to start at zero:
// init rewrites every bit word, the counters and the overflow flag of the
// run before any kernel reads them, so the previous call's contents never
// reach a result.  Calls that overlap on one GPU (different streams, or
// host threads without ordering) share the array and race; callers
// serialise them.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "eclcc_internal.h"

namespace eclcc {
namespace {

static __device__ __align__(256) unsigned char scratch[ECLCC_NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// The device address of this device's `scratch`, looked up once per device.
int scratch_address(void** out) {
  int dev = -1;
  int rc = check_hip(hipGetDevice(&dev), ECLCC_ERR_DEVICE, "no usable HIP device");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(ECLCC_ERR_DEVICE, "device index beyond the library's per-device table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    rc = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), ECLCC_ERR_DEVICE,
                   "could not look up the scratch array");
    if (rc) return rc;
    if (!ptr) return set_error(ECLCC_ERR_DEVICE, "scratch array has no device address");
    g_scratch[dev] = ptr;
  }
  *out = g_scratch[dev];
  return ECLCC_OK;
}

}  // namespace
}  // namespace eclcc

using namespace eclcc;

extern "C" int eclcc_run_device_noalloc(int32_t nodes, int32_t edges, const int32_t* d_nindex,
                                        const int32_t* d_nlist, int32_t* d_nstat,
                                        const eclcc_params* params, void* stream) try {
  if (nodes < 1 || edges < 0) return set_error(ECLCC_ERR_ARG, "node or edge count too low");
  if (!d_nindex || (edges > 0 && !d_nlist)) return set_error(ECLCC_ERR_ARG, "null CSR pointer");
  if (!d_nstat) return set_error(ECLCC_ERR_ARG, "null label pointer");
  Tuning t;
  int rc = resolve_tuning(params, &t);
  if (rc) return rc;
  const long long cap = hub_capacity(edges, t.wave_max, 1);
  const size_t need = workspace_bytes(nodes, cap, 1);
  if (need > ECLCC_NOALLOC_SCRATCH_BYTES)
    return set_error(ECLCC_ERR_BUDGET,
                     "graph needs " + std::to_string(need) + " bytes of workspace; "
                     "eclcc_run_device_noalloc holds " +
                         std::to_string((unsigned long long)ECLCC_NOALLOC_SCRATCH_BYTES) +
                         " (use eclcc_run_device with a workspace of eclcc_workspace_bytes())");
  void* ptr = nullptr;
  if ((rc = scratch_address(&ptr))) return rc;
  Workspace w;
  if (!carve_workspace(ptr, ECLCC_NOALLOC_SCRATCH_BYTES, nodes, cap, 1, &w))
    return set_error(ECLCC_ERR_STATE, "scratch carve failed");
  KParams p{};
  p.nidx = d_nindex;
  p.nlist = d_nlist;
  p.nstat = d_nstat;
  p.n = nodes;
  p.ngraphs = 1;
  p.nslots = nodes;
  fill_params(&p, w, t);
  return check_hip(launch_pipeline(p, false, Shape{t.hub_blocks, t.block_threads},
                                   (hipStream_t)stream, nullptr),
                   ECLCC_ERR_LAUNCH, "kernel launch failed");
} catch (...) {
  return host_exception();
}
