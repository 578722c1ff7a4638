"""Autotuner for the degree thresholds and launch shapes of the CC kernels.

Coordinate descent over a small discrete space, timing the real batched run
(hipGraph replay) on resident data with HIP events on the launch stream.
Labels do not depend on any of these parameters (every setting computes the
same minimum-ID labels); only the time changes.  Results are persisted as a
JSON table keyed by workload class.
"""
from __future__ import annotations

import json
import statistics
import time
from pathlib import Path
from typing import Dict, List, Optional, Tuple

from . import _native as N
from .batch import Batch
from .cc import default_params

SPACE: Dict[str, List[int]] = {
    # second-entry hooks taken by 1 in 2^shift vertices: random graphs need
    # fewer of them (1024-graph ER / R-MAT mix, 128 graphs: 0.92 -> 0.81 ms
    # with 2), power-law and grid graphs more
    "sample_shift": [0, 1, 2, 3],
    "wave_max_deg": [512, 2048, 8192],
    "thread_max_deg": [8, 16, 32],
    "hub_chunk": [2048, 4096, 8192],
    "block_threads": [256, 512],
    "giant_filter": [2, 1, 0],
    "path_mode": [1, 2, 0],
}

TABLE = N.PKG_ROOT / "tuning" / "mi355x.json"


def _params(d: Dict[str, int]) -> N.EclccParams:
    return default_params(**d)


def time_batch(batch: Batch, params: N.EclccParams, reps: int = 3) -> float:
    """Median milliseconds of one batched run (hipGraph replay)."""
    import torch

    batch.set_params(params)
    stream = torch.cuda.current_stream()
    batch.run(stream, use_graph=True)  # capture + warm
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        batch.run(stream, use_graph=True)
        b.record(stream)
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b))
    return statistics.median(times)


def tune(batch: Batch, space: Optional[Dict[str, List[int]]] = None, reps: int = 3,
         passes: int = 2, budget_s: float = 120.0) -> Tuple[Dict[str, int], List[dict]]:
    """Coordinate descent: each parameter in turn takes the value with the
    lowest time while the others stay at their current best."""
    space = space or SPACE
    base = default_params().as_dict()
    best = {k: base[k] for k in space}
    cache: Dict[Tuple, float] = {}
    trials: List[dict] = []
    t_end = time.perf_counter() + budget_s

    def measure(cfg: Dict[str, int]) -> float:
        key = tuple(sorted(cfg.items()))
        if key not in cache:
            cache[key] = time_batch(batch, _params(cfg), reps)
            trials.append({"params": dict(cfg), "ms": cache[key]})
        return cache[key]

    best_ms = measure(best)
    for _ in range(passes):
        changed = False
        for name, values in space.items():
            for v in values:
                if time.perf_counter() > t_end:
                    break
                cfg = dict(best, **{name: v})
                ms = measure(cfg)
                if ms < best_ms * 0.995:
                    best, best_ms, changed = cfg, ms, True
        if not changed:
            break
    batch.set_params(_params(best))
    return best, trials


def save(key: str, params: Dict[str, int], trials: List[dict], path: Path = TABLE,
         note: str = "") -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    table = json.loads(path.read_text()) if path.exists() else {}
    table[key] = {"params": params, "trials": trials, "note": note}
    path.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")


def load(key: str, path: Path = TABLE) -> Optional[N.EclccParams]:
    if not path.exists():
        return None
    entry = json.loads(path.read_text()).get(key)
    return None if entry is None else _params(entry["params"])


# One graph: each parameter varied alone around the per-class choice
# (eclcc_params_for_graph), and the pairs that trade sampled hooks against
# the work left to compute1/compute2.
PLAIN_SPACE: List[Dict[str, int]] = (
    [dict()]
    + [dict(sample_shift=s) for s in (0, 1, 2, 3)]
    + [dict(sample_min_deg=d) for d in (0, 8, 16, 32)]
    + [dict(path_mode=m) for m in (0, 1, 2)]
    + [dict(thread_max_deg=t) for t in (4, 8, 32, 64)]
    + [dict(wave_max_deg=w) for w in (256, 512, 1024, 4096, 8192)]
    + [dict(thread_max_deg=t, sample_shift=s) for t, s in ((8, 1), (32, 1), (64, 1), (8, 2))]
    + [dict(giant_filter=f) for f in (1, 0)]
)


def sweep_plain(g, space: Optional[List[Dict[str, int]]] = None, reps: int = 20):
    """Times one graph resident in HBM under each parameter set of `space`
    (overrides of eclcc_params_for_graph's choice): yields one dict per set
    with the pass time (HIP events over `reps` back-to-back passes), the
    per-kernel times (event pairs, mean of 5) and whether the labels equal
    the first set's."""
    import numpy as np
    import torch

    from .cc import KERNELS, DeviceGraph, params_for_graph

    stream = torch.cuda.current_stream()
    dg = DeviceGraph(g, "cuda:0")
    ref = None
    for kw in space or PLAIN_SPACE:
        dg.set_params(params_for_graph(g, **kw))
        for _ in range(3):
            dg.run(stream)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            dg.run(stream)
        b.record(stream)
        torch.cuda.synchronize()
        kms = [statistics.mean(x) for x in zip(*[dg.run_profiled(stream)[0] for _ in range(5)])]
        lab = dg.labels()
        if ref is None:
            ref = lab
        yield {"graph": g.name, "overrides": kw, "params": dg.params.as_dict(),
               "ms": a.elapsed_time(b) / reps,
               "kernels_ms": {k: v for k, v in zip(KERNELS, kms)},
               "labels_equal": bool(np.array_equal(lab, ref))}


def main(argv=None) -> None:
    """Tunes on a sample of the batch mix and stores the result in the table
    under "batch_mix" (read by the served mode):
    python -m eclcc.autotune [--graphs 256] [--budget 120] [--out path].
    With --plain WORKLOAD (rmat24, grid256, er65536) it sweeps the plain
    path's parameters on that graph instead and prints one JSON line per
    parameter set."""
    import argparse

    from . import workloads

    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=256)
    ap.add_argument("--budget", type=float, default=120.0)
    ap.add_argument("--out", default=str(TABLE))
    ap.add_argument("--plain", default=None, choices=["rmat24", "grid256", "er65536"])
    a = ap.parse_args(argv)
    import torch

    torch.cuda.init()  # before the library touches the device (time_batch times on torch's stream)
    if a.plain:
        g = {"rmat24": workloads.rmat24, "grid256": workloads.grid256,
             "er65536": lambda: workloads.er65536(1)}[a.plain]()
        for row in sweep_plain(g):
            print(json.dumps(row), flush=True)
        return
    graphs = workloads.batch_mix(a.graphs)
    with Batch(graphs) as b:
        best, trials = tune(b, budget_s=a.budget)
    save("batch_mix", best, trials, Path(a.out),
         note=f"coordinate descent on the first {a.graphs} graphs of the batch mix "
              f"(workloads.batch_mix), hipGraph replay, median of 3")
    print(json.dumps({"batch_mix": best}))


if __name__ == "__main__":
    main()
