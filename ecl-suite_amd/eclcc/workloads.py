"""The seeded synthetic inputs the CC path is measured and tested on.

    er65536     Erdos-Renyi G(n=65,536, M=262,144) -- average degree exactly 8
    rmat24      R-MAT scale 24, edge factor 16, (a,b,c,d)=(.57,.19,.19,.05),
                symmetrised, deduplicated, self-loops dropped, IDs permuted
    grid256     3-D 7-point grid 256^3, each edge kept with p = 0.6, IDs permuted
    er_with_stars  a random graph beside stars whose centres lie outside its
                largest tree (numpy's generator)
    batch_mix   graphs with n log-uniform on [1e3, 1e6]; 50/50 Erdos-Renyi
                (average degree uniform on [4, 32]) / R-MAT (edge factor 8)

All draws come from libeclgraph's splitmix64 counter stream, so a workload is
a pure function of its seed.
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import List, Optional

from .graph import Graph, rand64, set_generator_threads

RMAT_ABC = (0.57, 0.19, 0.19)
BATCH_SEED = 4096


def er65536(seed: int = 1) -> Graph:
    g = Graph.er(65536, 262144, seed)
    g.name = f"er65536-deg8-seed{seed}"
    return g


def rmat(scale: int, edge_factor: int = 16, seed: int = 24) -> Graph:
    a, b, c = RMAT_ABC
    g = Graph.rmat(scale, edge_factor << scale, seed, a=a, b=b, c=c, permute=True)
    g.name = f"rmat{scale}-ef{edge_factor}-seed{seed}"
    return g


def rmat24(seed: int = 24) -> Graph:
    return rmat(24, 16, seed)


def grid3d(side: int, keep_p: float = 0.6, seed: int = 3) -> Graph:
    g = Graph.grid3d(side, side, side, keep_p, seed, permute=True)
    g.name = f"grid{side}^3-p{keep_p}-seed{seed}"
    return g


def grid256(seed: int = 3) -> Graph:
    return grid3d(256, 0.6, seed)


def er_with_stars(n_er: int, pairs: int, stars: int, leaves: int, seed: int) -> Graph:
    """A random graph on n_er vertices (`pairs` random pairs) beside `stars`
    stars of `leaves` leaves each, all IDs permuted (numpy's seeded
    generator).  The random part holds most vertices, so its tree is the
    largest, and every star centre is a high-degree vertex outside it: the
    case where the default filter leaves long lists to the hub kernel."""
    import numpy as np

    rng = np.random.default_rng(seed)
    src = [rng.integers(0, n_er, pairs)]
    dst = [rng.integers(0, n_er, pairs)]
    base = n_er
    for _ in range(stars):
        src.append(np.full(leaves, base, dtype=np.int64))
        dst.append(np.arange(base + 1, base + 1 + leaves, dtype=np.int64))
        base += 1 + leaves
    perm = rng.permutation(base)
    g = Graph.from_edges(base, perm[np.concatenate(src)], perm[np.concatenate(dst)])
    g.name = f"er{n_er}+{stars}stars{leaves}-seed{seed}"
    return g


@dataclass
class MixSpec:
    index: int
    kind: str          # "er" or "rmat"
    n: int
    seed: int
    pairs: int = 0     # ER: distinct undirected pairs
    scale: int = 0     # R-MAT
    samples: int = 0   # R-MAT directed samples (edge factor 8 x n)


def _unit(seed: int, ctr: int) -> float:
    return (rand64(seed, ctr) >> 11) * (1.0 / (1 << 53))


def mix_spec(index: int, base_seed: int = BATCH_SEED, n_lo: int = 1000, n_hi: int = 1_000_000) -> MixSpec:
    """Graph `index` of the batched mix: n log-uniform on [n_lo, n_hi],
    generator ER or R-MAT with probability 1/2 each."""
    k = 4 * index
    u = _unit(base_seed, k)
    n = int(math.exp(math.log(n_lo) + u * (math.log(n_hi) - math.log(n_lo))))
    n = max(n_lo, min(n_hi, n))
    kind = "er" if (rand64(base_seed, k + 1) >> 63) == 0 else "rmat"
    seed = rand64(base_seed, k + 3)
    if kind == "er":
        d = 4.0 + 28.0 * _unit(base_seed, k + 2)
        pairs = int(round(n * d / 2.0))
        pairs = min(pairs, n * (n - 1) // 4)
        return MixSpec(index, kind, n, seed, pairs=pairs)
    scale = max(1, math.ceil(math.log2(n)))
    return MixSpec(index, kind, n, seed, scale=scale, samples=8 * n)


def build_mix_graph(spec: MixSpec) -> Graph:
    if spec.kind == "er":
        g = Graph.er(spec.n, spec.pairs, spec.seed)
    else:
        a, b, c = RMAT_ABC
        g = Graph.rmat(spec.scale, spec.samples, spec.seed, n=spec.n, a=a, b=b, c=c, permute=True)
    g.name = f"mix{spec.index}-{spec.kind}-n{spec.n}"
    return g


def batch_mix(count: int, base_seed: int = BATCH_SEED, n_lo: int = 1000, n_hi: int = 1_000_000,
              workers: Optional[int] = None, progress=None) -> List[Graph]:
    """`count` graphs of the batched mix, generated in parallel (each graph
    single-threaded; ctypes releases the GIL).  progress(done, count) is
    called from time to time."""
    import os
    import time

    specs = [mix_spec(i, base_seed, n_lo, n_hi) for i in range(count)]
    workers = workers or min(16, os.cpu_count() or 1)

    def one(spec: MixSpec) -> Graph:
        set_generator_threads(1)
        return build_mix_graph(spec)

    # largest first for load balance, results returned in index order
    order = sorted(range(count), key=lambda i: -specs[i].n)
    out: List[Optional[Graph]] = [None] * count
    last = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for done, (i, g) in enumerate(zip(order, ex.map(one, [specs[i] for i in order])), 1):
            out[i] = g
            if progress is not None and (time.perf_counter() - last > 10 or done == count):
                progress(done, count)
                last = time.perf_counter()
    return out  # type: ignore[return-value]
