#!/usr/bin/env python3
"""Benchmark of the race-free ECL-CC path on one MI355X.

Default (``python bench.py``): plain ECL-CC on R-MAT scale 24 (edge factor
16, a,b,c = .57,.19,.19, symmetrised, deduplicated, IDs permuted, seed 24),
CSR resident in HBM.  One step = one full CC pass (init ... flatten).  Prints
ONE JSON line with edges/s (edges = CSR entries, the reference's definition,
ECL-CC-atomic.cu:316), the HBM roofline of the pass, the CPU baseline (the
oracle's ECL-CC restatement on the host cores, threaded and serial, same
graph), and the other half of the metric as two objects of the same line:
``batched`` (the 4,096-graph mix in one hipGraph launch: graphs/s against
one graph per launch) and ``served`` (1,024 requests from the mix under
Poisson arrivals: graphs/s, p50/p99 ms against one request at a time).

Other modes: ``--mode batched`` (many graphs, one hipGraph launch, against
one graph per launch), ``--mode served`` (continuous batching under
Poisson arrivals, p50/p99 latency), ``--mode racy`` (the cost of race
freedom) and ``--mode mst`` (the ECL-MST minimum spanning forest of the
same graph with the reference's default weights).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "ecl-suite_amd"))
sys.path.insert(0, str(REPO))

METRIC = "CC edges/s + % HBM peak on RMAT-24; batched/served graphs/s with p50/p99 ms"
HBM_PEAK = 8.0e12  # B/s, MI355X


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(name: str):
    from eclcc import workloads

    if name == "rmat24":
        return workloads.rmat24()
    if name == "grid256":
        return workloads.grid256()
    if name == "er65536":
        return workloads.er65536(1)
    if name.startswith("rmat"):
        return workloads.rmat(int(name[4:]))
    if name.startswith("erstars-"):
        # erstars-<stars>x<leaves>: stars beside a random graph of 4M vertices
        # and 16M pairs (seed 7), their centres outside its largest tree
        stars, leaves = (int(x) for x in name[len("erstars-"):].split("x"))
        return workloads.er_with_stars(4_000_000, 16_000_000, stars, leaves, 7)
    raise SystemExit(f"unknown workload {name}")


def cpu_share() -> tuple:
    """Host cores this process may use: the CPUs in its affinity mask, capped
    by the CPU share the host grants it (OMP_NUM_THREADS, when set), and
    os.cpu_count() for reference."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS", "")
    cores = min(avail, int(share)) if share.isdigit() and int(share) > 0 else avail
    return cores, avail, os.cpu_count() or 1, share or None


def cpu_baseline(g, budget_s: float, gpu_labels):
    """The oracle's ECL-CC restatement (ECL-CC-atomic.cu:69-264) on the host:
    CPU-N threaded on every core this process may use, and CPU-1 serial, on
    the full graph; each the median of as many runs as fit the budget (at
    least one, at most 9)."""
    import numpy as np
    from oracle import oracle

    cores, avail, host, share = cpu_share()

    def timed(fn, budget):
        times, out = [], None
        t_end = time.perf_counter() + budget
        while len(times) < 9 and (not times or time.perf_counter() < t_end):
            t0 = time.perf_counter()
            out = fn()
            times.append(time.perf_counter() - t0)
        return statistics.median(times), len(times), out

    medN, runsN, (labels, used) = timed(lambda: oracle.eclcc_threads(g.nindex, g.nlist, cores),
                                        budget_s / 2)
    med1, runs1, labels1 = timed(lambda: oracle.eclcc_serial(g.nindex, g.nlist), budget_s / 2)
    return {
        "value": g.edges / medN,
        "unit": "edges/s",
        "cores": used,
        "kind": "port",
        "sample": f"full {g.name} (n={g.nodes}, m={g.edges}): the oracle's ECL-CC restatement, "
                  f"CPU-N on {used} threads (median of {runsN}) and CPU-1 serial (median of {runs1})",
        "seconds": medN,
        "host_cpus": host,
        "affinity_cpus": avail,
        "cpu_share_env": share,
        "labels_match_gpu": bool(np.array_equal(labels, gpu_labels)),
        "cpu1": {
            "value": g.edges / med1,
            "unit": "edges/s",
            "cores": 1,
            "seconds": med1,
            "runs": runs1,
            "labels_match_gpu": bool(np.array_equal(labels1, gpu_labels)),
        },
    }


def plain(args) -> dict:
    import torch
    from eclcc import DeviceGraph, params_for_graph, roofline
    from eclcc.cc import KERNELS

    t0 = time.perf_counter()
    g = make_workload(args.workload)
    log(f"generated {g.name}: n={g.nodes} m={g.edges} in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    # the defaults for this graph's class (degree skew), as eclcc_compute picks them
    dg = DeviceGraph(g, "cuda:0", params_for_graph(g))
    stream = torch.cuda.current_stream(dev)
    for _ in range(args.warmup):
        dg.run(stream)
    torch.cuda.synchronize()
    # per-kernel durations: HIP events between the kernels on the launch stream
    per = [[] for _ in KERNELS]
    tots = []
    for _ in range(max(3, min(args.steps, 10))):
        kms, tot = dg.run_profiled(stream)
        for i, x in enumerate(kms):
            per[i].append(x)
        tots.append(tot)
    labels = dg.labels()
    log("timed region")
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(args.steps):
        dg.run(stream)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    if not (labels == dg.labels()).all():
        raise SystemExit("labels changed between runs")
    B = g.algorithmic_bytes()
    kernel_ms = {k: statistics.mean(v) for k, v in zip(KERNELS, per)}
    pass_ms = statistics.mean(tots)
    hubs, tasks = dg.hub_stats()
    # roofline of the kernel that takes the most time, on its own algorithmic bytes
    kb = roofline.kernel_bytes(g.nindex, g.nlist, dg.params.thread_max_deg, dg.params.wave_max_deg,
                               dg.params.giant_filter, dg.params.sample_shift,
                               dg.params.sample_min_deg)
    dom = max((k for k in KERNELS if kb[k]), key=lambda k: kernel_ms[k])
    dom_gbs = kb[dom] / (kernel_ms[dom] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": g.edges / (ms * 1e-3),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": g.name, "nodes": g.nodes, "edges": g.edges,
                   "algorithmic_bytes": B, "params": dg.params.as_dict()},
        "pct_hbm_peak": 100.0 * B / (ms * 1e-3) / HBM_PEAK,
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": dom_gbs,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": dom_gbs * 1e9 / HBM_PEAK,
            "traffic": None,
            "algorithmic_bytes": kb[dom],
            "bytes_rule": roofline.PER_UNIT.get(dom, "bookkeeping"),
            "avg_launch_ms": kernel_ms[dom],
        },
        "pass_roofline": {
            "bound": "hbm",
            "kernels": list(KERNELS),
            "achieved": B / (pass_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": B / (pass_ms * 1e-3) / HBM_PEAK,
            "traffic": None,
            "algorithmic_bytes": B,
        },
        "kernels_ms": kernel_ms,
        "kernel_algorithmic_bytes": kb,
        "traffic_source": None,
        "pass_ms_events": pass_ms,
        "hub_vertices": hubs,
        "hub_tasks": tasks,
    }
    tr = REPO / "profiles" / f"{args.workload}_traffic.json"
    if tr.exists():
        t = json.loads(tr.read_text())
        if (t.get("edges") == g.edges and t.get("nodes") == g.nodes
                and t.get("params", dg.params.as_dict()) == dg.params.as_dict()):
            kt = t.get("kernels", {}).get(dom)
            if kt is not None:
                out["roofline"]["traffic"] = kt["fetch_bytes"] + kt["write_bytes"]
                # the bytes the kernel moves (line granularity included) over the
                # same launch time: how close its gathers run to the HBM peak
                gbs = out["roofline"]["traffic"] / (kernel_ms[dom] * 1e-3) / 1e9
                out["roofline"]["traffic_GBps"] = gbs
                out["roofline"]["traffic_frac"] = gbs * 1e9 / HBM_PEAK
            out["pass_roofline"]["traffic"] = t["bytes_per_pass"]
            gbs = t["bytes_per_pass"] / (pass_ms * 1e-3) / 1e9
            out["pass_roofline"]["traffic_GBps"] = gbs
            out["pass_roofline"]["traffic_frac"] = gbs * 1e9 / HBM_PEAK
            out["traffic_source"] = f"profiles/{tr.name} (rocprofv3: {t.get('source', 'FETCH_SIZE + WRITE_SIZE, copy-calibrated')})"
    # the same graph with every adjacency list read (giant_filter 1)
    if dg.params.giant_filter == 2 and not args.no_compare:
        del dg
        torch.cuda.empty_cache()
        d1 = DeviceGraph(g, "cuda:0", params_for_graph(g, giant_filter=1))
        for _ in range(args.warmup):
            d1.run(stream)
        per1 = [[] for _ in KERNELS]
        for _ in range(max(3, min(args.steps, 10))):
            kms1, _ = d1.run_profiled(stream)
            for i, x in enumerate(kms1):
                per1[i].append(x)
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(args.steps):
            d1.run(stream)
        end.record(stream)
        torch.cuda.synchronize()
        ms1 = start.elapsed_time(end) / args.steps
        same1 = bool((d1.labels() == labels).all())
        out["read_every_edge"] = {
            "params": d1.params.as_dict(),
            "ms_per_step": ms1,
            "edges_per_s": g.edges / (ms1 * 1e-3),
            "pct_hbm_peak": 100.0 * B / (ms1 * 1e-3) / HBM_PEAK,
            "kernels_ms": {k: statistics.mean(v) for k, v in zip(KERNELS, per1)},
            "labels_equal": same1,
        }
        if not same1:
            raise SystemExit("labels differ between giant_filter 2 and 1")
        del d1
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(g, args.cpu_budget, labels)
    # the batched and served halves of the metric, on the graph mix.  The
    # drop-in call on host arrays comes after them: its three rounds of
    # 2.2 GB device allocations and frees, made right before the served
    # phase, raised the server's p50 latency from 0.45 to 0.59 ms (same box;
    # calls on a 0.5 GB graph did not)
    dg = None
    torch.cuda.empty_cache()
    from eclcc import benchmodes

    if not args.no_served:
        log("served: Poisson arrivals into the continuous-batching server")
        out["served"] = summary(benchmodes.served(args, check=oracle_check))
        torch.cuda.empty_cache()
    if not args.no_batched:
        log("batched: the graph mix in one hipGraph launch")
        out["batched"] = summary(benchmodes.batched(args, check=oracle_check))
        torch.cuda.empty_cache()
    if not args.no_host_call:
        log("drop-in call on host arrays (PCIe included)")
        out["host_array_call"] = host_array_call(g, labels)
    return out


def host_array_call(g, gpu_labels, calls: int = 3) -> dict:
    """The drop-in eclcc_compute on the host CSR, as the reference's main
    calls computeCC (ECL-CC-atomic.cu:363): allocate, copy in, run, copy the
    labels out, free.  Wall time per call (median of `calls` after one
    warm-up call), against the PCIe-bound lower limit of copying the CSR in
    and the labels out at the measured pinned DMA rate.  Never the headline:
    that is the HBM-resident pass."""
    import numpy as np
    from eclcc import compute_cc

    lab = np.empty(g.nodes, dtype=np.int32)
    ni = np.ascontiguousarray(g.nindex, dtype=np.int32)
    nl = np.ascontiguousarray(g.nlist, dtype=np.int32)
    walls, kern = [], []
    for k in range(calls + 1):
        t0 = time.perf_counter()
        secs = compute_cc(g.nodes, g.edges, ni, nl, lab)
        if k:
            walls.append(time.perf_counter() - t0)
            kern.append(secs)
    wall = statistics.median(walls)
    h2d = 4 * (g.nodes + 1) + 4 * g.edges
    return {
        "ms_per_call": 1e3 * wall,
        "kernel_ms": 1e3 * statistics.median(kern),
        "edges_per_s": g.edges / wall,
        "h2d_bytes": h2d,
        "d2h_bytes": 4 * g.nodes,
        "copy_GBps": (h2d + 4 * g.nodes) / wall / 1e9,
        "pcie_peak_GBps": 64.0,  # PCIe 5.0 x16, one direction
        "pcie_frac": (h2d + 4 * g.nodes) / wall / 64.0e9,
        "labels_equal": bool(np.array_equal(lab, gpu_labels)),
        "calls": calls,
    }


def oracle_check(graphs, labels) -> bool:
    """Checker (never timed): every graph's labels equal the oracle's
    minimum-ID labels, computed on the host cores this process may use."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    from oracle import oracle

    def one(i):
        g = graphs[i]
        return bool(np.array_equal(labels[i], oracle.cc_labels(g.nindex, g.nlist)))

    with ThreadPoolExecutor(max_workers=cpu_share()[0]) as ex:
        return all(ex.map(one, range(len(graphs))))


def summary(d: dict) -> dict:
    """A batched/served result as an object of the one bench line (the keys
    the line already carries at its top level are dropped)."""
    drop = {"metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "dtype", "data"}
    return {k: v for k, v in d.items() if k not in drop}


def racy(args) -> dict:
    """The cost of race freedom: the same kernels with plain (racy) label
    loads and stores (lib/libeclcc_plain.so, a research build) against the
    product's relaxed atomics, on the same graph and device buffers, for
    both filter modes.  Labels of the plain build are compared, not trusted."""
    import ctypes as C

    import numpy as np
    import torch
    from eclcc import DeviceGraph, default_params
    from eclcc import _native as N

    g = make_workload(args.workload)
    log(f"generated {g.name}")
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    P = N.plain_lib()
    rows = {}
    for filt in (2, 1):
        params = default_params(giant_filter=filt)
        dg = DeviceGraph(g, "cuda:0", params)

        def plain_run():
            N.check_cc(P.eclcc_run_device(dg.nodes, dg.edges, dg.nindex.data_ptr(),
                                          dg.nlist.data_ptr(), dg.nstat.data_ptr(),
                                          dg.work.data_ptr(), dg.work_bytes, C.byref(params),
                                          stream.cuda_stream))

        def timed(fn):
            for _ in range(args.warmup):
                fn()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            a.record(stream)
            for _ in range(args.steps):
                fn()
            b.record(stream)
            torch.cuda.synchronize()
            return a.elapsed_time(b) / args.steps

        t_atomic = timed(lambda: dg.run(stream))
        lab_atomic = dg.labels()
        t_plain = timed(plain_run)
        lab_plain = dg.labels()
        rows[f"giant_filter={filt}"] = {
            "race_free_ms": t_atomic, "plain_ms": t_plain,
            "race_free_over_plain": t_atomic / t_plain,
            "labels_equal": bool(np.array_equal(lab_atomic, lab_plain)),
        }
        del dg
        torch.cuda.empty_cache()
    return {"metric": METRIC, "mode": "racy", "workload": g.name, "n_gpus": 1,
            "steps": args.steps, "results": rows,
            "plain_library": P.eclcc_version().decode()}


def mst_bench(args) -> dict:
    """ECL-MST (src/racefree/egr-input/MST/ECL-MST_10-atomic.cu) on the same
    workload with main's default weights (:428-437) and threshold (:450-462).
    One step = one forest, graph resident in HBM; the time is the device
    window from the first kernel to the last, which includes the per-round
    reads of the worklist size, as the reference's GPUTimer does (:321-363).
    Checked at full size: the forest has n - c edges and the same components
    as the graph (c from the CC library); checked edge for edge against the
    oracle's Kruskal on R-MAT-21, which is also the CPU baseline."""
    import numpy as np
    import torch
    from eclcc import DeviceGraph, Graph, mst, workloads

    t0 = time.perf_counter()
    g = make_workload(args.workload)
    w = mst.mst_weights(g)
    thr = mst.threshold(g, w)
    log(f"generated {g.name} with weights in {time.perf_counter() - t0:.1f}s; threshold {thr}")
    torch.cuda.set_device(0)
    dm = mst.DeviceMST(g, w)
    for _ in range(args.warmup):
        dm.run(thr)
    torch.cuda.synchronize()
    log("timed region (mst)")
    dev_ms, wall_ms = [], []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        dev_ms.append(dm.run(thr))
        wall_ms.append((time.perf_counter() - t1) * 1e3)
    ms = statistics.mean(dev_ms)
    forest = dm.result()
    del dm
    torch.cuda.empty_cache()
    # spanning-forest check at full size, through the CC library
    dg = DeviceGraph(g, "cuda:0")
    dg.run()
    torch.cuda.synchronize()
    lab = dg.labels()
    del dg
    ncomp = int(np.unique(lab).size)
    src = np.repeat(np.arange(g.nodes, dtype=np.int32), np.diff(g.nindex))
    sel = forest == 1
    fg = Graph.from_edges(g.nodes, src[sel], g.nlist[sel])
    df = DeviceGraph(fg, "cuda:0")
    df.run()
    torch.cuda.synchronize()
    spanning = bool(int(sel.sum()) == g.nodes - ncomp and np.array_equal(df.labels(), lab))
    del df
    weight = int(w[sel].astype(np.int64).sum())
    host_call = mst_host_array_call(g, w, forest)
    # edge-for-edge parity and the CPU baseline on a bounded sample
    gs = workloads.rmat(21)
    ws = mst.mst_weights(gs)
    from oracle import oracle

    times = []
    ref = None
    t_end = time.perf_counter() + args.cpu_budget
    while not times or (len(times) < 3 and time.perf_counter() < t_end):
        t1 = time.perf_counter()
        ref = oracle.mst(gs.nindex, gs.nlist, ws)
        times.append(time.perf_counter() - t1)
    got_s = mst.compute_mst(gs, ws)
    B = 4 * (g.nodes + 1) + 8 * g.edges + g.edges
    out = {
        "metric": "MST edges/s on RMAT-24 (next path after CC)",
        "mode": "mst",
        "value": g.edges / (ms * 1e-3),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "wall_ms_per_step": statistics.mean(wall_ms),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{g.name}, default weights (ECL-MST_10-atomic.cu:428-437)",
                   "nodes": g.nodes, "edges": g.edges, "threshold": thr,
                   "algorithmic_bytes": B},
        "pct_hbm_peak": 100.0 * B / (ms * 1e-3) / HBM_PEAK,
        "pass_roofline": {
            "bound": "hbm",
            "achieved": B / (ms * 1e-3) / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": B / (ms * 1e-3) / HBM_PEAK,
            "traffic": None,
            "algorithmic_bytes": B,
            "bytes_rule": "4(n+1) + 4m + 4m + m: nindex, nlist and weights read once, in_mst written",
        },
        "forest_edges": int(sel.sum()),
        "forest_weight": weight,
        "components": ncomp,
        "spanning_forest": spanning,
        "sample_parity": {"workload": gs.name, "edges_equal_oracle": bool(np.array_equal(got_s, ref))},
        "host_array_call": host_call,
        "cpu_baseline": {
            "value": gs.edges / statistics.median(times),
            "unit": "edges/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{gs.name} (n={gs.nodes}, m={gs.edges}), the oracle's Kruskal "
                      f"(ECL-MST_10-atomic.cu:112-154), median of {len(times)} runs",
            "seconds": statistics.median(times),
        },
    }
    tr = REPO / "profiles" / f"mst_{args.workload}_traffic.json"
    if tr.exists():
        t = json.loads(tr.read_text())
        if t.get("nodes") == g.nodes and t.get("edges") == g.edges and t.get("threshold") == thr:
            gbs = t["bytes_per_forest"] / (ms * 1e-3) / 1e9
            out["pass_roofline"]["traffic"] = t["bytes_per_forest"]
            out["pass_roofline"]["traffic_GBps"] = gbs
            out["pass_roofline"]["traffic_frac"] = gbs * 1e9 / HBM_PEAK
            out["traffic_source"] = f"profiles/{tr.name} (rocprofv3: {t.get('source', 'FETCH_SIZE + WRITE_SIZE, copy-calibrated')})"
    return out


def mst_host_array_call(g, w, forest, calls: int = 3) -> dict:
    """The drop-in eclmst_compute on host arrays, as the reference's main
    calls gpuMST (ECL-MST_10-atomic.cu:450-462): allocate, copy nindex,
    nlist and weights in, run, copy in_mst out, free.  Wall time per call
    (median of `calls` after one warm-up call), against the PCIe link; the
    forest must equal the device-resident run's."""
    import ctypes as C

    import numpy as np
    from eclcc import _native as N

    ni = np.ascontiguousarray(g.nindex, dtype=np.int32)
    nl = np.ascontiguousarray(g.nlist, dtype=np.int32)
    ww = np.ascontiguousarray(w, dtype=np.int32)
    out = np.empty(max(g.edges, 1), dtype=np.uint8)
    L = N.mst_lib()
    walls, dev = [], []
    for k in range(calls + 1):
        sec = C.c_double(0.0)
        t0 = time.perf_counter()
        N.check_mst(L.eclmst_compute_timed(g.nodes, g.edges, ni.ctypes.data, nl.ctypes.data, ww.ctypes.data,
                                           out.ctypes.data, C.byref(sec)))
        if k:
            walls.append(time.perf_counter() - t0)
            dev.append(sec.value)
    wall = statistics.median(walls)
    h2d = 4 * (g.nodes + 1) + 8 * g.edges
    return {
        "ms_per_call": 1e3 * wall,
        "device_ms": 1e3 * statistics.median(dev),
        "edges_per_s": g.edges / wall,
        "h2d_bytes": h2d,
        "d2h_bytes": g.edges,
        "copy_GBps": (h2d + g.edges) / wall / 1e9,
        "pcie_peak_GBps": 64.0,  # PCIe 5.0 x16, one direction
        "pcie_frac": (h2d + g.edges) / wall / 64.0e9,
        "forest_equal": bool(np.array_equal(out[: g.edges], forest)),
        "calls": calls,
    }


def copy_probe(args) -> dict:
    """A plain device copy of known size: the reference point that scales the
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch

    n = 1 << 28  # 1 GiB of int32
    a = torch.ones(n, dtype=torch.int32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        b.copy_(a)
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    return {"mode": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "ms_per_copy": ms,
            "GBps": 8 * n / (ms * 1e-3) / 1e9}


def gather_probe(args) -> dict:
    """Random 4-byte gathers of known size, the reference point for the label
    walks: ecl-suite_amd/bin/gather_probe gathers 2^26 uniformly random
    entries of a 64 MiB table (the label array of a 2^24-vertex graph), one
    independent gather per lane, with plain, relaxed agent-scope (the
    library's label loads) and non-temporal loads, then a relaxed
    agent-scope CAS (the hooks' flavour) and atomic add on the same
    addresses, then copies 256 MiB.
    Under rocprofv3 --pmc TCC_EA0_RDREQ_* it gives each flavour's request
    sizes; here its JSON lines are gathered into one."""
    import subprocess

    exe = REPO / "ecl-suite_amd" / "bin" / "gather_probe"
    if not exe.exists():
        raise SystemExit(f"{exe} is not built (run __graft_entry__.build())")
    r = subprocess.run([str(exe), str(max(2, min(args.steps, 1000)))], capture_output=True, text=True,
                       timeout=300)
    if r.returncode != 0:
        raise SystemExit(f"gather_probe failed ({r.returncode}): {r.stderr.strip()}")
    rows = [json.loads(x) for x in r.stdout.splitlines() if x.startswith("{")]
    return {"mode": "gather", "kernels": {row.pop("kernel"): row for row in rows},
            "values_correct": all(row["values_correct"] for row in rows)}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", default="plain",
                    choices=["plain", "batched", "served", "copy", "gather", "racy", "mst"])
    ap.add_argument("--workload", default="rmat24")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the read-every-edge (giant_filter 1) comparison pass")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-host-call", action="store_true",
                    help="plain mode: leave out the drop-in call on host arrays")
    ap.add_argument("--no-batched", action="store_true",
                    help="plain mode: leave out the batched object of the line")
    ap.add_argument("--no-served", action="store_true",
                    help="plain mode: leave out the served object of the line")
    # batched / served
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--tune-graphs", type=int, default=256)
    ap.add_argument("--no-per-graph", action="store_true",
                    help="batched mode: leave out the one-graph-per-launch runs (for counter runs, "
                         "where the profiler's cost per dispatch would dominate)")
    ap.add_argument("--tune-budget", type=float, default=0.0,
                    help="seconds of autotuning on a sample of the mix; 0 = the committed table")
    ap.add_argument("--requests", type=int, default=1024)
    ap.add_argument("--qps", type=float, default=0.0, help="offered load; 0 = --load x capacity")
    ap.add_argument("--load", type=float, default=0.75)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--serve-budget-gib", type=float, default=32.0)
    ap.add_argument("--serve-streams", type=int, default=3)
    ap.add_argument("--serve-max-batch", type=int, default=128)
    ap.add_argument("--serve-wait-us", type=int, default=100,
                    help="a batch smaller than --serve-max-batch waits this long for company "
                         "while another batch is on the GPU")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(graphs are independent units; see the sharding design). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.mode == "plain":
        out = plain(args)
    elif args.mode == "copy":
        out = copy_probe(args)
    elif args.mode == "gather":
        out = gather_probe(args)
    elif args.mode == "racy":
        out = racy(args)
    elif args.mode == "mst":
        out = mst_bench(args)
    else:
        from eclcc import benchmodes

        out = benchmodes.run(args.mode, args, check=oracle_check)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
