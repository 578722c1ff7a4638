// gather_probe.hip -- bin/gather_probe: the rate of random 4-byte gathers on
// this GPU, the reference point for the label walks of ECL-CC and ECL-MST.
//
// out[i] = table[idx[i]] for 2^26 uniformly random indices into a 64 MiB
// table (the label array of a 2^24-vertex graph), one independent gather
// per lane, in three load flavours: plain, relaxed agent-scope atomic (the
// flavour of the library's label loads) and non-temporal; then the same
// addresses with a relaxed agent-scope CAS (the hooks' flavour) and atomic
// add, each leaving the table as it is (x -> x, + 0), for the rate of
// device-scope atomics on distinct random addresses; then a streaming copy
// of the 256 MiB index array.  Prints one JSON line per kernel: best
// time of `reps` launches (first launch not counted), gathers per second,
// and whether the output equals the expected values.  Run it under
// rocprofv3 --pmc TCC_EA0_RDREQ_* to see the request size of each flavour.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

constexpr int kTable = 1 << 24;
constexpr int kGathers = 1 << 26;
constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void g_plain(const int* t, const int* __restrict__ idx,
                                                    int* __restrict__ out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;  // grid covers kGathers exactly
  out[i] = t[idx[i]];
}

__global__ __launch_bounds__(kThreads) void g_agent(const int* t, const int* __restrict__ idx,
                                                    int* __restrict__ out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  out[i] = __hip_atomic_load(&t[idx[i]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kThreads) void g_nontemporal(const int* t, const int* __restrict__ idx,
                                                          int* __restrict__ out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  out[i] = __builtin_nontemporal_load(&t[idx[i]]);
}

// table[x] == x: the CAS always succeeds and writes the value back, the add
// adds 0, so every launch sees the same table and returns out[i] = idx[i]
__global__ __launch_bounds__(kThreads) void g_cas(int* t, const int* __restrict__ idx,
                                                  int* __restrict__ out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  int expect = idx[i];
  __hip_atomic_compare_exchange_strong(&t[idx[i]], &expect, idx[i], __ATOMIC_RELAXED,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  out[i] = expect;
}

__global__ __launch_bounds__(kThreads) void g_add(int* t, const int* __restrict__ idx,
                                                  int* __restrict__ out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  // the addend is 0 (indices are non-negative), but not known to be, so the
  // add stays a read-modify-write instead of becoming a load
  out[i] = __hip_atomic_fetch_add(&t[idx[i]], idx[i] >> 31, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kThreads) void g_copy(const int4* __restrict__ a, int4* __restrict__ b) {
  const int i = blockIdx.x * kThreads + threadIdx.x;  // grid covers kGathers / 4 exactly
  b[i] = a[i];
}

bool ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "gather_probe: %s: %s\n", what, hipGetErrorString(e));
  return false;
}

}  // namespace

int main(int argc, char** argv) {
  const int reps = argc > 1 ? std::atoi(argv[1]) : 6;
  if (reps < 2 || reps > 1000) {
    std::fprintf(stderr, "usage: %s [reps 2..1000]\n", argv[0]);
    return 2;
  }
  std::vector<int> h_idx(kGathers), h_t(kTable);
  unsigned long long s = 88172645463325252ull;  // xorshift64
  for (int i = 0; i < kGathers; i++) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    h_idx[i] = (int)(s % kTable);
  }
  for (int i = 0; i < kTable; i++) h_t[i] = i;
  int *t = nullptr, *idx = nullptr, *out = nullptr;
  hipEvent_t e0, e1;
  if (!ok(hipMalloc(&t, sizeof(int) * kTable), "hipMalloc") ||
      !ok(hipMalloc(&idx, sizeof(int) * (size_t)kGathers), "hipMalloc") ||
      !ok(hipMalloc(&out, sizeof(int) * (size_t)kGathers), "hipMalloc") ||
      !ok(hipMemcpy(t, h_t.data(), sizeof(int) * kTable, hipMemcpyHostToDevice), "hipMemcpy") ||
      !ok(hipMemcpy(idx, h_idx.data(), sizeof(int) * (size_t)kGathers, hipMemcpyHostToDevice), "hipMemcpy") ||
      !ok(hipEventCreate(&e0), "hipEventCreate") || !ok(hipEventCreate(&e1), "hipEventCreate"))
    return 1;
  const char* names[] = {"plain", "agent_relaxed", "nontemporal", "cas_agent", "add_agent", "copy"};
  std::vector<int> h_out(kGathers);
  for (int k = 0; k < 6; k++) {
    float best = 1e30f;
    for (int r = 0; r < reps; r++) {
      if (!ok(hipEventRecord(e0, nullptr), "hipEventRecord")) return 1;
      switch (k) {
        case 0: g_plain<<<kGathers / kThreads, kThreads>>>(t, idx, out); break;
        case 1: g_agent<<<kGathers / kThreads, kThreads>>>(t, idx, out); break;
        case 2: g_nontemporal<<<kGathers / kThreads, kThreads>>>(t, idx, out); break;
        case 3: g_cas<<<kGathers / kThreads, kThreads>>>(t, idx, out); break;
        case 4: g_add<<<kGathers / kThreads, kThreads>>>(t, idx, out); break;
        default: g_copy<<<kGathers / 4 / kThreads, kThreads>>>((const int4*)idx, (int4*)out); break;
      }
      float ms = 0.f;
      if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(e1, nullptr), "hipEventRecord") ||
          !ok(hipEventSynchronize(e1), "kernel") || !ok(hipEventElapsedTime(&ms, e0, e1), "timing"))
        return 1;
      if (r > 0 && ms < best) best = ms;
    }
    // every kernel writes out[i] = idx[i] (table[x] = x, which the atomics
    // leave as it is; the copy copies idx)
    if (!ok(hipMemcpy(h_out.data(), out, sizeof(int) * (size_t)kGathers, hipMemcpyDeviceToHost), "hipMemcpy"))
      return 1;
    bool same = true;
    for (int i = 0; i < kGathers && same; i++) same = h_out[i] == h_idx[i];
    if (k < 5)
      std::printf("{\"kernel\": \"%s\", \"gathers\": %d, \"table_bytes\": %d, \"ms\": %.4f, "
                  "\"gathers_per_s\": %.4e, \"values_correct\": %s}\n",
                  names[k], kGathers, 4 * kTable, best, kGathers / (best * 1e-3), same ? "true" : "false");
    else
      std::printf("{\"kernel\": \"copy\", \"bytes_read\": %lld, \"bytes_written\": %lld, \"ms\": %.4f, "
                  "\"GBps\": %.1f, \"values_correct\": %s}\n",
                  4ll * kGathers, 4ll * kGathers, best, 8.0 * kGathers / (best * 1e-3) / 1e9,
                  same ? "true" : "false");
    std::fflush(stdout);
    if (!ok(hipMemset(out, 0, sizeof(int) * (size_t)kGathers), "hipMemset")) return 1;
  }
  (void)hipFree(t);
  (void)hipFree(idx);
  (void)hipFree(out);
  return 0;
}
