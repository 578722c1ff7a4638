"""Batched and served benchmark modes (``bench.py --mode batched|served``).

batched: the graph mix (n log-uniform on [1e3, 1e6], 50/50 Erdos-Renyi with
average degree uniform on [4, 32] / R-MAT with edge factor 8), packed into one
device allocation; thresholds and grid shapes chosen by the autotuner on a
sample of the mix; one step = one hipGraph replay of the whole batch.  Compared
with one graph per launch on the same resident data.

served: requests drawn from the same mix arrive as a Poisson stream; the
native server admits them by HBM budget, batches whatever is waiting and
overlaps batches on several streams.  Compared with serving the same
requests one at a time (copy in, one graph per launch, copy out).
Request CSR sits in pinned host memory, as a receive path would leave it.
"""
from __future__ import annotations

import ctypes as C
import statistics
import sys
import threading
import time
from typing import List

import numpy as np

from . import _native as N
from . import autotune, workloads
from .batch import Batch
from .cc import default_params

METRIC = "CC edges/s + % HBM peak on RMAT-24; batched/served graphs/s with p50/p99 ms"
HBM_PEAK = 8.0e12


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(graphs) -> int:
    return sum(g.algorithmic_bytes() for g in graphs)


def _labels_equal(a: List[np.ndarray], b: List[np.ndarray]) -> bool:
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def batched(args) -> dict:
    import torch

    torch.cuda.set_device(0)
    count = args.batch_graphs
    t0 = time.perf_counter()
    graphs = workloads.batch_mix(count, progress=lambda d, c: log(f"generated {d}/{c} graphs"))
    n_tot = sum(g.nodes for g in graphs)
    m_tot = sum(g.edges for g in graphs)
    log(f"generated {count} graphs: {n_tot} vertices, {m_tot} CSR entries in "
        f"{time.perf_counter() - t0:.1f}s")

    # autotune on a sample of the mix
    t0 = time.perf_counter()
    sample = graphs[: min(args.tune_graphs, count)]
    with Batch(sample) as sb:
        best, trials = autotune.tune(sb, budget_s=args.tune_budget)
    log(f"autotuned on {len(sample)} graphs in {time.perf_counter() - t0:.1f}s: {best}")
    params = default_params(**best)

    t0 = time.perf_counter()
    b = Batch(graphs, params)
    log(f"uploaded batch ({b.device_bytes() / 2**30:.1f} GiB) in {time.perf_counter() - t0:.1f}s")
    stream = torch.cuda.current_stream()
    for _ in range(max(1, args.warmup)):
        b.run(stream, use_graph=True)
    torch.cuda.synchronize()
    log("timed region (batched)")
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        b.run(stream, use_graph=True)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    kms, _ = b.run_profiled(stream)
    lab_batch = b.all_labels()

    # one graph per launch, synchronising after each (the plain API's usage)
    log("one graph per launch")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for g in range(count):
        b.run_single(g, stream)
        torch.cuda.synchronize()
    plain_sync_s = time.perf_counter() - t0
    lab_plain = b.all_labels()
    # one graph per launch, all enqueued back to back (no host wait)
    start.record(stream)
    for g in range(count):
        b.run_single(g, stream)
    end.record(stream)
    torch.cuda.synchronize()
    plain_async_s = start.elapsed_time(end) * 1e-3
    same = _labels_equal(lab_batch, lab_plain) and _labels_equal(lab_plain, b.all_labels())
    B = algorithmic_bytes(graphs)
    gps = count / (ms * 1e-3)
    out = {
        "metric": METRIC,
        "mode": "batched",
        "value": gps,
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"batch_mix {count} graphs (n log-uniform 1e3..1e6, ER deg U[4,32] / "
                               f"R-MAT ef 8, seed {workloads.BATCH_SEED})",
                   "graphs": count, "nodes": n_tot, "edges": m_tot, "algorithmic_bytes": B,
                   "params": best, "hipgraph": True},
        "edges_per_s": m_tot / (ms * 1e-3),
        "pct_hbm_peak": 100.0 * B / (ms * 1e-3) / HBM_PEAK,
        "plain_sync_graphs_per_s": count / plain_sync_s,
        "plain_async_graphs_per_s": count / plain_async_s,
        "speedup_vs_plain_sync": (count / (ms * 1e-3)) / (count / plain_sync_s),
        "speedup_vs_plain_async": (count / (ms * 1e-3)) / (count / plain_async_s),
        "labels_batched_equal_plain": bool(same),
        "autotune_trials": len(trials),
        "kernels_ms": kms,
    }
    b.close()
    return out


class PinnedPool:
    """Request graphs copied into one pinned host allocation."""

    def __init__(self, graphs):
        import torch

        total = sum(g.nodes + 1 + g.edges for g in graphs)
        self.buf = torch.empty(total, dtype=torch.int32, pin_memory=True)
        arr = self.buf.numpy()
        self.items = []
        off = 0
        for g in graphs:
            ni = arr[off: off + g.nodes + 1]
            ni[:] = g.nindex
            off += g.nodes + 1
            nl = arr[off: off + g.edges]
            nl[:] = g.nlist
            off += g.edges
            self.items.append(type(g)(ni, nl, None, g.name))


def served(args) -> dict:
    import torch

    from .server import Server

    torch.cuda.set_device(0)
    R = args.requests
    t0 = time.perf_counter()
    graphs = workloads.batch_mix(R, base_seed=workloads.BATCH_SEED + 1,
                                 progress=lambda d, c: log(f"generated {d}/{c} requests"))
    pool = PinnedPool(graphs).items
    del graphs
    n_tot = sum(g.nodes for g in pool)
    m_tot = sum(g.edges for g in pool)
    log(f"generated {R} requests ({m_tot} CSR entries) in {time.perf_counter() - t0:.1f}s")
    params = autotune.load("batch_mix") or default_params()

    # plain: one request at a time, resident device buffers, copy in/out
    max_n = max(g.nodes for g in pool)
    max_m = max(g.edges for g in pool)
    dev = torch.device("cuda:0")
    d_ni = torch.empty(max_n + 1, dtype=torch.int32, device=dev)
    d_nl = torch.empty(max(max_m, 1), dtype=torch.int32, device=dev)
    d_ns = torch.empty(max_n, dtype=torch.int32, device=dev)
    wb = N.cc_lib().eclcc_workspace_bytes(max_n, max_m, C.byref(params))
    d_w = torch.empty(int(wb), dtype=torch.uint8, device=dev)
    out_plain = [torch.empty(g.nodes, dtype=torch.int32, pin_memory=True) for g in pool]
    stream = torch.cuda.current_stream()
    L = N.cc_lib()

    def plain_one(i):
        g = pool[i]
        d_ni[: g.nodes + 1].copy_(torch.from_numpy(g.nindex), non_blocking=True)
        if g.edges:
            d_nl[: g.edges].copy_(torch.from_numpy(g.nlist), non_blocking=True)
        N.check_cc(L.eclcc_run_device(g.nodes, g.edges, d_ni.data_ptr(), d_nl.data_ptr(),
                                      d_ns.data_ptr(), d_w.data_ptr(), int(wb), C.byref(params),
                                      stream.cuda_stream))
        out_plain[i].copy_(d_ns[: g.nodes], non_blocking=True)
        torch.cuda.synchronize()

    plain_one(0)
    log("plain one-at-a-time")
    lat_plain = []
    t0 = time.perf_counter()
    for i in range(R):
        ts = time.perf_counter()
        plain_one(i)
        lat_plain.append((time.perf_counter() - ts) * 1e3)
    plain_s = time.perf_counter() - t0
    plain_gps = R / plain_s

    budget = int(args.serve_budget_gib * 2**30)
    srv = Server(hbm_budget_bytes=budget, streams=args.serve_streams, max_batch=args.serve_max_batch,
                 params=params)
    try:
        # capacity: every request submitted at once
        log("server burst (capacity)")
        t0 = time.perf_counter()
        ids = [srv.submit(g) for g in pool]
        res = [srv.wait(r) for r in ids]
        burst_s = time.perf_counter() - t0
        cap_gps = R / burst_s
        same = all(np.array_equal(lab, out_plain[i].numpy()) for i, (lab, _) in enumerate(res))

        # Poisson arrivals at a fraction of the measured capacity
        qps = args.qps if args.qps > 0 else args.load * cap_gps
        rng = np.random.default_rng(args.seed)
        arrivals = np.cumsum(rng.exponential(1.0 / qps, R))
        log(f"server Poisson run at {qps:.0f} requests/s")
        submit_t = [0.0] * R
        rids = [0] * R

        def producer():
            base = time.perf_counter()
            for i in range(R):
                target = base + arrivals[i]
                while True:
                    now = time.perf_counter()
                    if now >= target:
                        break
                    if target - now > 2e-4:
                        time.sleep((target - now) * 0.5)
                submit_t[i] = time.perf_counter()
                rids[i] = srv.submit(pool[i])

        th = threading.Thread(target=producer)
        t0 = time.perf_counter()
        th.start()
        th.join()
        lats = []
        done_t = []
        for i in range(R):
            lab, lat = srv.wait(rids[i])
            same = same and np.array_equal(lab, out_plain[i].numpy())
            lats.append(lat)
            done_t.append(submit_t[i] + lat * 1e-3)
        span = max(done_t) - t0
        st = srv.stats()
    finally:
        srv.close()
    B = sum(g.algorithmic_bytes() for g in pool)
    served_gps = R / span
    return {
        "metric": METRIC,
        "mode": "served",
        "value": served_gps,
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": R,
        "warmup": 1,
        "ms_per_step": 1e3 * span / R,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"served: {R} requests from the batch mix (seed "
                               f"{workloads.BATCH_SEED + 1}), Poisson arrivals at {qps:.0f}/s",
                   "requests": R, "nodes": n_tot, "edges": m_tot, "algorithmic_bytes": B,
                   "qps": qps, "hbm_budget_bytes": budget, "streams": args.serve_streams,
                   "max_batch": args.serve_max_batch, "params": params.as_dict()},
        "p50_ms": float(np.percentile(lats, 50)),
        "p99_ms": float(np.percentile(lats, 99)),
        "mean_ms": float(np.mean(lats)),
        "capacity_graphs_per_s": cap_gps,
        "plain_graphs_per_s": plain_gps,
        "plain_p50_ms": float(np.percentile(lat_plain, 50)),
        "plain_p99_ms": float(np.percentile(lat_plain, 99)),
        "speedup_served_vs_plain": served_gps / plain_gps,
        "speedup_capacity_vs_plain": cap_gps / plain_gps,
        "labels_equal_plain": bool(same),
        "server_stats": st,
    }


def run(mode: str, args) -> dict:
    if mode == "batched":
        return batched(args)
    if mode == "served":
        return served(args)
    raise ValueError(mode)
