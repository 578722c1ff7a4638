// eclcc_server.hip -- continuous-batching CC server (include/eclcc.h,
// eclcc_server_*).  The reference has no serving mode; it runs one graph per
// process (ECL-CC-atomic.cu:338-400).  Here one worker thread:
//   * admits waiting requests in arrival order while their device bytes fit
//     in a device arena sized by the HBM budget and their labels fit in a
//     pinned staging arena (admission by free-HBM budget);
//   * packs the admitted requests into one batch (the batched kernels of
//     eclcc_kernels.hip, one launch sequence for the whole batch) on one of
//     several streams, so batches that are admitted while earlier ones still
//     run overlap on the device;
//   * copies labels device -> pinned staging asynchronously, and when a
//     batch's completion event fires copies them to the caller's array,
//     stamps the latency and frees the batch's arena space.
// Each request's labels are the same as eclcc_compute() on that graph alone:
// the batched kernels keep graphs in disjoint label slots.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "eclcc_arena.h"
#include "eclcc_internal.h"

using namespace eclcc;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id;
  int32_t n, m;
  bool resident;           // true: nindex, nlist and out are device pointers
  const int32_t* nindex;
  const int32_t* nlist;
  int32_t* out;
  Clock::time_point t_submit;
  double latency_ms = 0;
  int status = ECLCC_OK;
  std::string error;       // the failure's message when status != ECLCC_OK
  bool done = false;
  // placement inside its batch
  long long slot0 = 0;
  size_t stage_off = 0;  // byte offset of its labels in the pinned arena
};

// Arena bytes of a request besides the batch workspace: its CSR and labels
// when they come from the host, and its graph-table and chunk-map entries.
size_t request_device_bytes(int32_t n, int32_t m, bool resident) {
  const size_t slots = ((size_t)n + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
  const size_t tables = align256(sizeof(GraphDesc)) + align256(4 * (slots / kSlotAlign) + 4);
  if (resident) return tables;
  return align256(4 * ((size_t)n + 1)) + align256(4 * (size_t)m) + align256(4 * (size_t)n) + tables;
}

// Pinned staging bytes of a batch that holds this one request: its labels
// (host requests only), then the batch's graph table and chunk map, which
// go up from pinned memory.  The same sums as the worker's admission loop.
size_t request_stage_bytes(int32_t n, bool resident) {
  const size_t slots = ((size_t)n + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
  const size_t labels = resident ? 0 : align256(4 * (size_t)n);
  return labels + align256(sizeof(GraphDesc)) + align256(4 * (slots / kSlotAlign));
}

struct Batch {
  std::vector<Request*> reqs;
  std::vector<GraphDesc> desc;   // host copies stay alive until the batch retires
  std::vector<int32_t> chunk;
  size_t dev_off = 0, dev_bytes = 0;
  size_t stage_off = 0, stage_bytes = 0;
  size_t table_off = 0;  // graph table and chunk map, pinned, after the labels
  int stream = -1;
  hipEvent_t done_ev = nullptr;
};

}  // namespace

struct eclcc_server {
  eclcc_server_config cfg{};
  Tuning tuning{};
  int device = 0;
  char* d_arena = nullptr;
  char* h_stage = nullptr;  // pinned labels staging
  Arena dev, stage;
  std::vector<hipStream_t> streams;
  std::vector<Batch*> stream_batch;  // in-flight batch per stream (null = free)

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;
  std::unordered_map<int64_t, Request*> reqs;
  int64_t next_id = 0;
  bool stop = false;
  std::thread worker;
  std::string fatal;

  // stats
  int64_t completed = 0, batches = 0;
  size_t peak_bytes = 0;
  int32_t max_batch_seen = 0;
  int64_t inflight = 0;

  ~eclcc_server() {
    for (hipStream_t s : streams) (void)hipStreamDestroy(s);
    if (d_arena) (void)hipFree(d_arena);
    if (h_stage) (void)hipHostFree(h_stage);
    for (auto& kv : reqs) delete kv.second;
  }

  // Launches one batch on stream k.  Returns a status; on failure the
  // requests are completed with that status by the caller.
  int launch(Batch* b, int k) {
    try {
      return launch_tables(b, k);
    } catch (...) {
      return host_exception();  // the batch's tables could not be built
    }
  }

  int launch_tables(Batch* b, int k) {
    hipStream_t s = streams[k];
    const int G = (int)b->reqs.size();
    // layout inside the batch's arena block
    std::vector<GraphDesc>& desc = b->desc;
    desc.resize(G);
    long long vb = 0, hub_cap = 0;
    for (int g = 0; g < G; g++) {
      Request* r = b->reqs[g];
      desc[g].vbase = vb;
      desc[g].n = r->n;
      desc[g].m = r->m;
      r->slot0 = vb;
      vb += ((long long)r->n + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
      hub_cap += hub_capacity(r->m, tuning.wave_max, 1);
    }
    const long long nchunks = vb / kSlotAlign;
    std::vector<int32_t>& chunk = b->chunk;
    chunk.assign((size_t)nchunks, 0);
    for (int g = 0; g < G; g++) {
      const long long c0 = desc[g].vbase / kSlotAlign;
      const long long c1 = (desc[g].vbase + desc[g].n + kSlotAlign - 1) / kSlotAlign;
      for (long long c = c0; c < c1; c++) chunk[(size_t)c] = g;
    }
    char* base = d_arena + b->dev_off;
    size_t off = 0;
    auto carve = [&](size_t bytes) {
      char* p = base + off;
      off += align256(bytes);
      return p;
    };
    GraphDesc* d_desc = (GraphDesc*)carve(sizeof(GraphDesc) * G);
    int32_t* d_chunk = (int32_t*)carve(4 * (size_t)nchunks);
    // host requests: CSR and labels in the arena; resident ones: the caller's arrays
    for (int g = 0; g < G; g++) {
      Request* r = b->reqs[g];
      if (r->resident) {
        desc[g].nidx = r->nindex;
        desc[g].nlist = r->nlist;
        desc[g].nstat = r->out;
      } else {
        desc[g].nidx = (const int*)carve(4 * ((size_t)r->n + 1));
        desc[g].nlist = (const int*)carve(4 * (size_t)r->m);
        desc[g].nstat = (int*)carve(4 * (size_t)r->n);
      }
    }
    const size_t wbytes = workspace_bytes(vb, hub_cap, G);
    void* d_work = carve(wbytes);
    if (off > b->dev_bytes) return set_error(ECLCC_ERR_STATE, "server batch layout overflow");
    Workspace w;
    if (!carve_workspace(d_work, wbytes, vb, hub_cap, G, &w))
      return set_error(ECLCC_ERR_STATE, "server workspace carve failed");
    const hipMemcpyKind h2d = hipMemcpyHostToDevice;
    // the tables go up from pinned memory: an asynchronous copy from
    // pageable memory would be staged by the runtime on the host
    char* h_tab = h_stage + b->table_off;
    std::memcpy(h_tab, desc.data(), sizeof(GraphDesc) * G);
    std::memcpy(h_tab + align256(sizeof(GraphDesc) * G), chunk.data(), 4 * (size_t)nchunks);
    int rc = check_hip(hipMemcpyAsync(d_desc, h_tab, sizeof(GraphDesc) * G, h2d, s),
                       ECLCC_ERR_COPY, "copying to device failed");
    if (!rc)
      rc = check_hip(hipMemcpyAsync(d_chunk, h_tab + align256(sizeof(GraphDesc) * G),
                                    4 * (size_t)nchunks, h2d, s),
                     ECLCC_ERR_COPY, "copying to device failed");
    for (int g = 0; g < G && !rc; g++) {
      Request* r = b->reqs[g];
      if (r->resident) continue;
      rc = check_hip(hipMemcpyAsync((void*)desc[g].nidx, r->nindex, 4 * ((size_t)r->n + 1), h2d, s),
                     ECLCC_ERR_COPY, "copying to device failed");
      if (!rc && r->m > 0)
        rc = check_hip(hipMemcpyAsync((void*)desc[g].nlist, r->nlist, 4 * (size_t)r->m, h2d, s),
                       ECLCC_ERR_COPY, "copying to device failed");
    }
    if (rc) return rc;
    KParams p{};
    p.ngraphs = G;
    p.nslots = vb;
    p.desc = d_desc;
    p.chunk_graph = d_chunk;
    fill_params(&p, w, tuning);
    rc = check_hip(launch_pipeline(p, true, Shape{tuning.hub_blocks, tuning.block_threads}, s, nullptr),
                   ECLCC_ERR_LAUNCH, "kernel launch failed");
    for (int g = 0; g < G && !rc; g++) {
      Request* r = b->reqs[g];
      if (r->resident) continue;
      rc = check_hip(hipMemcpyAsync(h_stage + r->stage_off, desc[g].nstat, 4 * (size_t)r->n,
                                    hipMemcpyDeviceToHost, s),
                     ECLCC_ERR_COPY, "copying from device failed");
    }
    if (!rc) rc = check_hip(hipEventRecord(b->done_ev, s), ECLCC_ERR_LAUNCH, "event record failed");
    return rc;
  }

  // `status` != ECLCC_OK: the worker thread's message (set_error on this
  // thread) is kept with each request, so the client's wait reports it.
  void finish(Batch* b, int status) {
    const auto now = Clock::now();
    const char* msg = status == ECLCC_OK ? "" : eclcc_last_error();
    std::lock_guard<std::mutex> lock(mu);
    for (Request* r : b->reqs) {
      if (status == ECLCC_OK && !r->resident)
        std::memcpy(r->out, h_stage + r->stage_off, 4 * (size_t)r->n);
      r->status = status;
      if (status != ECLCC_OK) {
        try {
          r->error = msg;
        } catch (...) {
          // out of host memory: the client gets the status without the text
        }
      }
      r->latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
      r->done = true;
      completed++;
      inflight--;
    }
    dev.release(b->dev_off, b->dev_bytes);
    stage.release(b->stage_off, b->stage_bytes);
    cv_done.notify_all();
  }

  void run() {
    (void)hipSetDevice(device);
    for (;;) {
      // 1. retire finished batches
      bool busy = false;
      for (size_t k = 0; k < streams.size(); k++) {
        Batch* b = stream_batch[k];
        if (!b) continue;
        const hipError_t q = hipEventQuery(b->done_ev);
        if (q == hipErrorNotReady) {
          busy = true;
          continue;
        }
        int status = (q == hipSuccess) ? ECLCC_OK
                                       : check_hip(q, ECLCC_ERR_LAUNCH, "kernel execution failed");
        finish(b, status);
        (void)hipEventDestroy(b->done_ev);
        delete b;
        stream_batch[k] = nullptr;
      }
      // 2. admit a new batch on a free stream
      int free_k = -1;
      for (size_t k = 0; k < streams.size(); k++)
        if (!stream_batch[k]) { free_k = (int)k; break; }
      Batch* nb = nullptr;
      {
        std::unique_lock<std::mutex> lock(mu);
        if (stop && pending.empty() && !busy) return;
        // while other batches run, a small batch waits for company until its
        // oldest request has waited batch_wait_us: fuller batches carry less
        // fixed cost per graph (an idle GPU takes whatever is waiting at once)
        bool fill = free_k >= 0 && !pending.empty();
        if (fill && busy && cfg.batch_wait_us > 0 && (int)pending.size() < cfg.max_batch) {
          const auto waited = Clock::now() - pending.front()->t_submit;
          fill = waited >= std::chrono::microseconds(cfg.batch_wait_us);
        }
        if (fill) {
          // the batch and its request list are allocated before any request
          // leaves the queue; if the host is out of memory the requests wait
          nb = new (std::nothrow) Batch;
          if (nb) {
            try {
              nb->reqs.reserve((size_t)cfg.max_batch);
            } catch (...) {
              delete nb;
              nb = nullptr;
            }
          }
        }
        if (nb) {
          size_t sbytes = 0, req_bytes = 0, exact_d = 0, tab_bytes = 0;
          long long vb = 0, hub_cap = 0;
          // the requests stay queued until the batch has its memory
          size_t take = 0;
          while (take < pending.size() && (int)nb->reqs.size() < cfg.max_batch &&
                 inflight + (int64_t)nb->reqs.size() < cfg.max_inflight) {
            Request* r = pending[take];
            const long long vb2 = vb + ((long long)r->n + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
            const long long cap2 = hub_cap + hub_capacity(r->m, tuning.wave_max, 1);
            const size_t req2 = req_bytes + request_device_bytes(r->n, r->m, r->resident);
            const size_t need_d = 4096 + req2 + workspace_bytes(vb2, cap2, (long long)nb->reqs.size() + 1);
            const size_t need_s = sbytes + (r->resident ? 0 : align256(4 * (size_t)r->n));
            const size_t tab2 = align256(sizeof(GraphDesc) * (nb->reqs.size() + 1)) +
                                align256(4 * (size_t)(vb2 / kSlotAlign));
            if (dev.used() + need_d > dev.capacity() ||
                stage.used() + need_s + tab2 > stage.capacity())
              break;
            r->stage_off = sbytes;
            sbytes = need_s;
            tab_bytes = tab2;
            vb = vb2;
            hub_cap = cap2;
            req_bytes = req2;
            exact_d = need_d;
            nb->reqs.push_back(r);  // within the reserved capacity
            take++;
          }
          if (nb->reqs.empty()) {
            delete nb;
            nb = nullptr;
          } else {
            const bool ok = dev.alloc(exact_d, &nb->dev_off);
            const bool ok2 = ok && stage.alloc(sbytes + tab_bytes, &nb->stage_off);
            if (!ok2) {
              if (ok) dev.release(nb->dev_off, exact_d);
              // fragmented: the requests stay queued until a batch retires
              delete nb;
              nb = nullptr;
            } else {
              pending.erase(pending.begin(), pending.begin() + (std::ptrdiff_t)take);
              nb->dev_bytes = exact_d;
              nb->stage_bytes = sbytes + tab_bytes;
              nb->table_off = nb->stage_off + sbytes;
              for (Request* r : nb->reqs) r->stage_off += nb->stage_off;
              inflight += (int64_t)nb->reqs.size();
              batches++;
              max_batch_seen = std::max<int32_t>(max_batch_seen, (int32_t)nb->reqs.size());
              peak_bytes = std::max(peak_bytes, dev.used());
            }
          }
        }
        if (!nb && !busy) {
          // nothing on the GPU and nothing admissible: sleep until a submit
          if (stop && pending.empty()) return;
          cv_work.wait_for(lock, std::chrono::milliseconds(5));
          continue;
        }
      }
      if (nb) {
        int rc = check_hip(hipEventCreateWithFlags(&nb->done_ev, hipEventDisableTiming),
                           ECLCC_ERR_DEVICE, "event create failed");
        if (!rc) rc = launch(nb, free_k);
        if (rc) {
          (void)hipStreamSynchronize(streams[free_k]);
          finish(nb, rc);
          if (nb->done_ev) (void)hipEventDestroy(nb->done_ev);
          delete nb;
        } else {
          stream_batch[free_k] = nb;
        }
        continue;
      }
      // GPU busy, nothing new to launch: poll again shortly
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
};

extern "C" {

int eclcc_server_create(const eclcc_server_config* cfg, const eclcc_params* params, eclcc_server** out) try {
  if (!out) return set_error(ECLCC_ERR_ARG, "null output pointer");
  *out = nullptr;
  std::unique_ptr<eclcc_server> s(new (std::nothrow) eclcc_server);
  if (!s) return set_error(ECLCC_ERR_ALLOC, "host allocation failed");
  if (cfg) s->cfg = *cfg;
  int rc = resolve_tuning(params, &s->tuning);
  if (rc) return rc;
  if (hipGetDevice(&s->device) != hipSuccess) return set_error(ECLCC_ERR_DEVICE, "no HIP device");
  if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 4096;
  if (s->cfg.streams <= 0) s->cfg.streams = 4;
  if (s->cfg.max_batch <= 0) s->cfg.max_batch = 1024;
  if (s->cfg.batch_wait_us < 0) return set_error(ECLCC_ERR_ARG, "batch_wait_us must be >= 0");
  size_t freeb = 0, total = 0;
  if ((rc = check_hip(hipMemGetInfo(&freeb, &total), ECLCC_ERR_DEVICE, "hipMemGetInfo failed"))) return rc;
  size_t budget = s->cfg.hbm_budget_bytes ? s->cfg.hbm_budget_bytes : freeb / 10 * 8;
  if (budget > freeb) return set_error(ECLCC_ERR_BUDGET, "HBM budget exceeds free device memory");
  budget &= ~size_t(255);
  s->cfg.hbm_budget_bytes = budget;
  if ((rc = check_hip(hipMalloc((void**)&s->d_arena, budget), ECLCC_ERR_ALLOC, "could not allocate server arena")))
    return rc;
  s->dev.reset(budget);
  // labels are 4 bytes per vertex, a small share of a request's CSR bytes
  const size_t stage_bytes = align256(std::max<size_t>(budget / 8, size_t(1) << 20));
  if ((rc = check_hip(hipHostMalloc((void**)&s->h_stage, stage_bytes, hipHostMallocDefault),
                      ECLCC_ERR_ALLOC, "could not allocate pinned staging")))
    return rc;
  s->stage.reset(stage_bytes);
  s->streams.resize(s->cfg.streams, nullptr);
  s->stream_batch.resize(s->cfg.streams, nullptr);
  for (auto& st : s->streams)
    if ((rc = check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), ECLCC_ERR_DEVICE,
                        "stream create failed")))
      return rc;
  eclcc_server* raw = s.get();
  s->worker = std::thread([raw] { raw->run(); });
  *out = s.release();
  return ECLCC_OK;
} catch (...) {
  return host_exception();
}

}  // extern "C"

namespace {

// Device bytes of a batch that holds this one request (the admission loop's
// need_d for a first request).
size_t request_batch_bytes(const eclcc_server* s, int32_t nodes, int32_t edges, bool resident) {
  return 4096 + request_device_bytes(nodes, edges, resident) +
         workspace_bytes(((long long)nodes + kSlotAlign - 1) / kSlotAlign * kSlotAlign,
                         hub_capacity(edges, s->tuning.wave_max, 1), 1);
}

bool admissible_alone(const eclcc_server* s, int32_t nodes, int32_t edges, bool resident) {
  return request_batch_bytes(s, nodes, edges, resident) <= s->dev.capacity() &&
         request_stage_bytes(nodes, resident) <= s->stage.capacity();
}

int64_t submit_request(eclcc_server* s, int32_t nodes, int32_t edges, const int32_t* nindex,
                       const int32_t* nlist, int32_t* labels_out, bool resident) {
  if (!s) return set_error(ECLCC_ERR_ARG, "null server");
  if (nodes < 1 || edges < 0 || !nindex || (edges > 0 && !nlist) || !labels_out)
    return set_error(ECLCC_ERR_ARG, "bad request arguments");
  // admissible alone, by the worker's own sums (device arena and pinned
  // staging); a request that is not would stay at the head of the queue
  // for good and hold up every request behind it
  if (!admissible_alone(s, nodes, edges, resident))
    return set_error(ECLCC_ERR_BUDGET, "request larger than the HBM budget");
  Request* r = new (std::nothrow) Request;
  if (!r) return set_error(ECLCC_ERR_ALLOC, "host allocation failed");
  r->n = nodes;
  r->m = edges;
  r->resident = resident;
  r->nindex = nindex;
  r->nlist = nlist;
  r->out = labels_out;
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->stop) {
      delete r;
      return set_error(ECLCC_ERR_STATE, "server is stopping");
    }
    r->id = s->next_id++;
    try {  // queued in both or in neither
      s->reqs[r->id] = r;
      try {
        s->pending.push_back(r);
      } catch (...) {
        s->reqs.erase(r->id);
        throw;
      }
    } catch (...) {
      delete r;
      throw;
    }
  }
  s->cv_work.notify_one();
  return r->id;
}

}  // namespace

extern "C" {

int64_t eclcc_server_submit(eclcc_server* s, int32_t nodes, int32_t edges, const int32_t* nindex,
                            const int32_t* nlist, int32_t* labels_out) try {
  return submit_request(s, nodes, edges, nindex, nlist, labels_out, false);
} catch (...) {
  return host_exception();
}

int64_t eclcc_server_submit_device(eclcc_server* s, int32_t nodes, int32_t edges,
                                   const int32_t* d_nindex, const int32_t* d_nlist,
                                   int32_t* d_labels) try {
  return submit_request(s, nodes, edges, d_nindex, d_nlist, d_labels, true);
} catch (...) {
  return host_exception();
}

int eclcc_server_submit_device_many(eclcc_server* s, int32_t count, const int32_t* nodes,
                                    const int32_t* edges, const int32_t* const* d_nindex,
                                    const int32_t* const* d_nlist, int32_t* const* d_labels,
                                    int64_t* ids) try {
  if (!s) return set_error(ECLCC_ERR_ARG, "null server");
  if (count < 0 || (count > 0 && (!nodes || !edges || !d_nindex || !d_nlist || !d_labels || !ids)))
    return set_error(ECLCC_ERR_ARG, "bad request arguments");
  // check every request before queueing any
  for (int32_t i = 0; i < count; i++) {
    if (nodes[i] < 1 || edges[i] < 0 || !d_nindex[i] || (edges[i] > 0 && !d_nlist[i]) || !d_labels[i])
      return set_error(ECLCC_ERR_ARG, "bad request arguments (request " + std::to_string(i) + ")");
    if (!admissible_alone(s, nodes[i], edges[i], true))
      return set_error(ECLCC_ERR_BUDGET,
                       "request " + std::to_string(i) + " larger than the HBM budget");
  }
  std::vector<Request*> rs((size_t)count, nullptr);
  const auto now = Clock::now();
  for (int32_t i = 0; i < count; i++) {
    Request* r = new (std::nothrow) Request;
    if (!r) {
      for (Request* q : rs) delete q;
      return set_error(ECLCC_ERR_ALLOC, "host allocation failed");
    }
    r->n = nodes[i];
    r->m = edges[i];
    r->resident = true;
    r->nindex = d_nindex[i];
    r->nlist = d_nlist[i];
    r->out = d_labels[i];
    r->t_submit = now;
    rs[(size_t)i] = r;
  }
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->stop) {
      for (Request* q : rs) delete q;
      return set_error(ECLCC_ERR_STATE, "server is stopping");
    }
    // all queued or none: a host allocation failure part way takes the
    // ones already queued back out
    int32_t queued = 0;
    try {
      for (; queued < count; queued++) {
        Request* r = rs[(size_t)queued];
        r->id = s->next_id++;
        ids[queued] = r->id;
        s->reqs[r->id] = r;
        try {
          s->pending.push_back(r);
        } catch (...) {
          s->reqs.erase(r->id);
          throw;
        }
      }
    } catch (...) {
      for (int32_t i = queued - 1; i >= 0; i--) {
        s->pending.pop_back();
        s->reqs.erase(rs[(size_t)i]->id);
      }
      for (Request* q : rs) delete q;
      throw;
    }
  }
  s->cv_work.notify_one();
  return ECLCC_OK;
} catch (...) {
  return host_exception();
}

int eclcc_server_wait(eclcc_server* s, int64_t id, int32_t timeout_ms, double* latency_ms) try {
  if (!s) return set_error(ECLCC_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return set_error(ECLCC_ERR_STATE, "unknown request id");
  Request* r = it->second;
  auto pred = [r] { return r->done; };
  if (timeout_ms < 0) {
    s->cv_done.wait(lock, pred);
  } else if (!s->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms), pred)) {
    return set_error(ECLCC_ERR_TIMEOUT, "request not finished");
  }
  if (latency_ms) *latency_ms = r->latency_ms;
  const int status = r->status;
  std::string msg;
  if (status != ECLCC_OK) msg.swap(r->error);
  s->reqs.erase(it);
  delete r;
  if (status != ECLCC_OK)
    return set_error(status, msg.empty() ? std::string("request failed") : "request failed: " + msg);
  return ECLCC_OK;
} catch (...) {
  return host_exception();
}

int eclcc_server_stats(eclcc_server* s, int64_t* completed, int64_t* batches, size_t* peak_bytes,
                       int32_t* max_batch_seen) try {
  if (!s) return set_error(ECLCC_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lock(s->mu);
  if (completed) *completed = s->completed;
  if (batches) *batches = s->batches;
  if (peak_bytes) *peak_bytes = s->peak_bytes;
  if (max_batch_seen) *max_batch_seen = s->max_batch_seen;
  return ECLCC_OK;
} catch (...) {
  return host_exception();
}

void eclcc_server_free(eclcc_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  delete s;
}

}  // extern "C"
