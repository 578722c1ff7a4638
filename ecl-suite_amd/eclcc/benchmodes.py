"""Batched and served benchmark modes (``bench.py --mode batched|served``).

batched: the graph mix (n log-uniform on [1e3, 1e6], 50/50 Erdos-Renyi with
average degree uniform on [4, 32] / R-MAT with edge factor 8), packed into one
device allocation; thresholds and grid shapes chosen by the autotuner on a
sample of the mix; one step = one hipGraph replay of the whole batch.  Compared
with one graph per launch on the same resident data.

served: requests on graphs drawn from the same mix, resident in HBM, arrive
as a Poisson stream; the native server admits them by HBM budget, batches
whatever is waiting and overlaps batches on several streams.  Compared with
serving the same requests one at a time (one graph per launch, waiting for
each).  Requests with host arrays (pinned, as a receive path would leave
them) are measured beside it against the per-call allocate/copy/compute/
copy/free pattern.
"""
from __future__ import annotations

import ctypes as C
import gc
import statistics
import sys
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


def BASELINE_PARAMS():
    """Parameters of the one-graph-per-launch baselines: the built-in
    defaults, the fastest of the measured sets for the mix's graphs run one
    at a time."""
    return default_params()


def batch_params(args, graphs):
    """The batch's parameters: the committed autotuner table (tuning/
    mi355x.json, key batch_mix) unless a tuning budget is given, in which
    case the autotuner runs afresh on a sample of the mix."""
    if args.tune_budget > 0:
        t0 = time.perf_counter()
        sample = graphs[: min(args.tune_graphs, len(graphs))]
        with Batch(sample) as sb:
            best, trials = autotune.tune(sb, budget_s=args.tune_budget)
        log(f"autotuned on {len(sample)} graphs in {time.perf_counter() - t0:.1f}s: {best}")
        return default_params(**best), f"autotuned on {len(sample)} graphs ({len(trials)} trials)"
    p = autotune.load("batch_mix")
    if p is None:
        return default_params(), "built-in defaults"
    return p, "tuning/mi355x.json (batch_mix)"


def batched(args, check=None) -> dict:
    """check(graphs, labels) -> bool, if given, is called on the batched
    run's labels outside the timed region."""
    import torch

    torch.cuda.set_device(0)
    count = args.batch_graphs
    t0 = time.perf_counter()
    graphs = workloads.batch_mix(count, progress=lambda d, c: log(f"generated {d}/{c} graphs"))
    n_tot = sum(g.nodes for g in graphs)
    m_tot = sum(g.edges for g in graphs)
    log(f"generated {count} graphs: {n_tot} vertices, {m_tot} CSR entries in "
        f"{time.perf_counter() - t0:.1f}s")

    params, params_source = batch_params(args, graphs)
    best = params.as_dict()

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
    oracle_ok = None
    if check is not None:
        log("labels against the oracle (outside the timed region)")
        oracle_ok = bool(check(graphs, lab_batch))

    # one graph per launch, synchronising after each (the plain API's usage),
    # with the built-in defaults: the fastest parameters for graphs of this
    # mix run one at a time (13.4 K graphs/s on the served mix, against 9.1 K
    # with the batch's tuned path_mode 2, which pays off only in full batches)
    plain_sync_s = plain_async_s = same = None
    if not getattr(args, "no_per_graph", False):
        log("one graph per launch")
        b.set_params(BASELINE_PARAMS())
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
        "plain_sync_graphs_per_s": count / plain_sync_s if plain_sync_s else None,
        "plain_async_graphs_per_s": count / plain_async_s if plain_async_s else None,
        "speedup_vs_plain_sync": gps / (count / plain_sync_s) if plain_sync_s else None,
        "speedup_vs_plain_async": gps / (count / plain_async_s) if plain_async_s else None,
        "labels_batched_equal_plain": None if same is None else bool(same),
        "labels_equal_oracle": oracle_ok,
        "params_source": params_source,
        "baseline_params": BASELINE_PARAMS().as_dict(),
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


def served(args, check=None) -> dict:
    """Requests name graphs resident in HBM (a graph store); the server writes
    each request's labels in place.  Latency = submit -> labels in HBM.  The
    host-array path (CSR copied in, labels copied out) is measured beside it.
    check(graphs, labels) -> bool, if given, is called on the served labels
    after the timed runs."""
    import torch

    from .server import Server

    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    R = args.requests
    t0 = time.perf_counter()
    graphs = workloads.batch_mix(R, base_seed=workloads.BATCH_SEED + 1,
                                 progress=lambda d, c: log(f"generated {d}/{c} requests"))
    n_tot = sum(g.nodes for g in graphs)
    m_tot = sum(g.edges for g in graphs)
    log(f"generated {R} requests ({m_tot} CSR entries) in {time.perf_counter() - t0:.1f}s")
    # the server's batches are smaller than a whole mix: its own tuned entry
    params = autotune.load("served") or autotune.load("batch_mix") or default_params()
    base = BASELINE_PARAMS()

    # the graph store: every request's CSR resident in HBM, labels in place
    ni_all = torch.empty(sum(g.nodes + 1 for g in graphs), dtype=torch.int32, device=dev)
    nl_all = torch.empty(max(m_tot, 1), dtype=torch.int32, device=dev)
    lab_all = torch.empty(n_tot, dtype=torch.int32, device=dev)
    store = []
    oi = oe = ov = 0
    for g in graphs:
        ni = ni_all[oi: oi + g.nodes + 1]
        ni.copy_(torch.from_numpy(np.ascontiguousarray(g.nindex)))
        nl = nl_all[oe: oe + g.edges] if g.edges else nl_all[:1]
        if g.edges:
            nl.copy_(torch.from_numpy(np.ascontiguousarray(g.nlist)))
        store.append((g.nodes, g.edges, ni, nl, lab_all[ov: ov + g.nodes]))
        oi, oe, ov = oi + g.nodes + 1, oe + g.edges, ov + g.nodes
    torch.cuda.synchronize()
    log(f"graph store in HBM: {(ni_all.numel() + nl_all.numel() + lab_all.numel()) * 4 / 2**30:.1f} GiB")

    # plain: one request at a time, one graph per launch, wait for it
    L = N.cc_lib()
    max_n = max(g.nodes for g in graphs)
    max_m = max(g.edges for g in graphs)
    wb = int(L.eclcc_workspace_bytes(max_n, max_m, C.byref(base)))
    d_w = torch.empty(wb, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()

    def plain_one(i):
        n, m, ni, nl, lab = store[i]
        N.check_cc(L.eclcc_run_device(n, m, ni.data_ptr(), nl.data_ptr(), lab.data_ptr(),
                                      d_w.data_ptr(), wb, C.byref(base), stream.cuda_stream))
        stream.synchronize()

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
    ref_labels = [lab.cpu().numpy() for (_, _, _, _, lab) in store]

    budget = int(args.serve_budget_gib * 2**30)
    srv = Server(hbm_budget_bytes=budget, streams=args.serve_streams, max_batch=args.serve_max_batch,
                 params=params, batch_wait_us=args.serve_wait_us)
    try:
        def check_all():
            return all(np.array_equal(lab.cpu().numpy(), ref)
                       for (_, _, _, _, lab), ref in zip(store, ref_labels))

        # capacity: every request submitted at once, in one call through the
        # C ABI with its argument arrays built beforehand (the Python
        # wrapper's per-tensor checks took longer than the GPU work); the
        # span ends at the last completion by the server's own clock
        # (median of 3)
        h = srv._h
        sub_many = N.cc_lib().eclcc_server_submit_device_many
        wait_c = N.cc_lib().eclcc_server_wait
        cap_nodes = np.array([n for (n, _, _, _, _) in store], dtype=np.int32)
        cap_edges = np.array([m for (_, m, _, _, _) in store], dtype=np.int32)
        cap_ni = np.array([ni.data_ptr() for (_, _, ni, _, _) in store], dtype=np.uint64)
        cap_nl = np.array([nl.data_ptr() if m else 0 for (_, m, _, nl, _) in store], dtype=np.uint64)
        cap_lb = np.array([lab.data_ptr() for (_, _, _, _, lab) in store], dtype=np.uint64)
        caps = []
        same = True
        for rep in range(3):
            lab_all.fill_(-1)
            torch.cuda.synchronize()
            ids = np.zeros(R, dtype=np.int64)
            t0 = time.perf_counter()
            N.check_cc(sub_many(h, R, cap_nodes.ctypes.data, cap_edges.ctypes.data, cap_ni.ctypes.data,
                                cap_nl.ctypes.data, cap_lb.ctypes.data, ids.ctypes.data))
            t1 = time.perf_counter()
            lat = C.c_double(0.0)
            worst = 0.0
            for r in ids:
                N.check_cc(wait_c(h, int(r), -1, C.byref(lat)))
                worst = max(worst, lat.value)
            caps.append(R / ((t1 - t0) + worst * 1e-3))
            same = same and check_all()
        log(f"server burst capacity {[round(c) for c in caps]} graphs/s")
        cap_gps = statistics.median(caps)

        # Poisson arrivals at a fraction of the measured capacity, three
        # independent seeded streams; requests are submitted through the C ABI
        # directly so that the producer keeps up.  The run with the median
        # sustained rate is reported, all three beside it.
        qps = args.qps if args.qps > 0 else args.load * cap_gps
        sub = N.cc_lib().eclcc_server_submit_device
        argv = [(n, m, ni.data_ptr(), nl.data_ptr() if m else None, lab.data_ptr())
                for (n, m, ni, nl, lab) in store]

        def poisson_run(seed):
            arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / qps, R))
            lab_all.fill_(-1)
            torch.cuda.synchronize()
            gc.collect()
            submit_t = [0.0] * R
            rids = [0] * R
            b0 = srv.stats()["batches"]
            # the producer spins on the clock (a sleep's wake-up granularity is
            # coarser than the gaps between arrivals) and submits each request
            # when its arrival time comes
            t0 = time.perf_counter()
            for i in range(R):
                target = t0 + arrivals[i]
                while time.perf_counter() < target:
                    pass
                submit_t[i] = time.perf_counter()
                rids[i] = sub(h, *argv[i])
            for r in rids:
                N.check_cc(int(r))
            lats, done_t = [], []
            lat = C.c_double(0.0)
            for i in range(R):
                N.check_cc(N.cc_lib().eclcc_server_wait(h, int(rids[i]), -1, C.byref(lat)))
                lats.append(lat.value)
                done_t.append(submit_t[i] + lat.value * 1e-3)
            span = max(done_t) - t0
            return {"seed": seed, "graphs_per_s": R / span, "span_s": span,
                    "p50_ms": float(np.percentile(lats, 50)), "p99_ms": float(np.percentile(lats, 99)),
                    "mean_ms": float(np.mean(lats)),
                    "achieved_arrival_rate": R / (submit_t[-1] - t0) if R > 1 else qps,
                    "batches": srv.stats()["batches"] - b0, "labels_equal_plain": check_all()}

        runs = []
        for k in range(3):
            log(f"server Poisson run {k + 1}/3 at {qps:.0f} requests/s")
            runs.append(poisson_run(args.seed + k))
            log(f"  {runs[-1]['graphs_per_s']:.0f} graphs/s, p50 {runs[-1]['p50_ms']:.3f} ms, "
                f"p99 {runs[-1]['p99_ms']:.3f} ms")
        med = sorted(runs, key=lambda r: r["graphs_per_s"])[1]
        same = same and all(r["labels_equal_plain"] for r in runs)
        span = med["span_s"]
        achieved_qps = med["achieved_arrival_rate"]
        same = same and check_all()
        oracle_ok = None
        if check is not None:
            log("served labels against the oracle (after the timed runs)")
            oracle_ok = bool(check(graphs, [lab.cpu().numpy() for (_, _, _, _, lab) in store]))
        st = srv.stats()
        st["poisson_batches"] = med["batches"]

        # host-array requests: CSR copied in, labels copied out (PCIe path)
        log("server burst, host-array requests")
        pool = PinnedPool(graphs).items
        t0 = time.perf_counter()
        hids = [srv.submit(g) for g in pool]
        host_ok = True
        for i, r in enumerate(hids):
            lab, _ = srv.wait(r)
            host_ok = host_ok and np.array_equal(lab, ref_labels[i])
        host_cap_gps = R / (time.perf_counter() - t0)
    finally:
        srv.close()

    # the reference's per-call pattern for host arrays: allocate, copy in,
    # compute, copy out, free (eclcc_compute), on a bounded sample
    k = min(R, 128)
    out = np.empty(max_n, dtype=np.int32)
    t0 = time.perf_counter()
    for g in pool[:k]:
        N.check_cc(L.eclcc_compute(g.nodes, g.edges, g.nindex.ctypes.data,
                                   g.nlist.ctypes.data if g.edges else None, out.ctypes.data))
    host_plain_gps = k / (time.perf_counter() - t0)
    B = sum(g.algorithmic_bytes() for g in graphs)
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
        "config": {"workload": f"served: {R} requests on HBM-resident graphs of the batch mix (seed "
                               f"{workloads.BATCH_SEED + 1}), Poisson arrivals at {qps:.0f}/s",
                   "requests": R, "nodes": n_tot, "edges": m_tot, "algorithmic_bytes": B,
                   "qps": qps, "load": qps / cap_gps, "achieved_arrival_rate": achieved_qps,
                   "hbm_budget_bytes": budget,
                   "streams": args.serve_streams, "max_batch": args.serve_max_batch,
                   "batch_wait_us": args.serve_wait_us,
                   "params": params.as_dict(), "baseline_params": base.as_dict()},
        "p50_ms": med["p50_ms"],
        "p99_ms": med["p99_ms"],
        "mean_ms": med["mean_ms"],
        "poisson_runs": runs,
        "capacity_graphs_per_s": cap_gps,
        "capacity_bursts_graphs_per_s": caps,
        "plain_graphs_per_s": plain_gps,
        "plain_p50_ms": float(np.percentile(lat_plain, 50)),
        "plain_p99_ms": float(np.percentile(lat_plain, 99)),
        "speedup_served_vs_plain": served_gps / plain_gps,
        "speedup_capacity_vs_plain": cap_gps / plain_gps,
        "labels_equal_plain": bool(same and host_ok),
        "labels_equal_oracle": oracle_ok,
        "server_stats": st,
        "host_requests": {
            "capacity_graphs_per_s": host_cap_gps,
            "plain_compute_per_call_graphs_per_s": host_plain_gps,
            "plain_sample": f"first {k} requests through eclcc_compute (allocate, copy in, compute, "
                            f"copy out, free per call)",
            "speedup_capacity_vs_plain": host_cap_gps / host_plain_gps,
        },
    }


def run(mode: str, args, check=None) -> dict:
    if mode == "batched":
        return batched(args, check)
    if mode == "served":
        return served(args, check)
    raise ValueError(mode)
