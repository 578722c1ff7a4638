"""Algorithmic bytes of one CC pass and of each of its kernels.

Whole pass (the compulsory traffic: read the CSR once, write each label
once):  B(n, m) = 4(n+1) + 4m + 4n.

Per kernel, the bytes its own job cannot avoid moving:

    init      4(n+1) + 4n      read the index, write every label
    sample    16 per sampled vertex: read the entry it hooks and the labels
              of both ends, write its own label
    hubscan, pick              bookkeeping (no per-vertex or per-edge term)
    mark      4n + n/8         read every label, write one bit per vertex
    compute1  4 per adjacency entry + 8 per vertex of the low-degree bin
    compute2  4 per adjacency entry + 8 per vertex of the medium-degree bin
    compute3  4 per adjacency entry + 8 per vertex of the high-degree bin
    flatten   8n               read and write every label

A vertex is active when it has a smaller neighbour (only those hook,
ECL-CC-atomic.cu:119-132).  The sample kernel hooks one entry per vertex:
the second entry of an active vertex if that is smaller (with sample_shift
s > 0 only for the active vertices whose multiplicative hash
(v * 0x9E3779B1 mod 2^32) has its top s bits clear, and with
sample_min_deg d > 0 only for those with at least d entries), the first entry of an
inactive vertex with a neighbour.  Labels are counted as read by the
kernels that need them (sample, mark, flatten), not only as written.  The bins split the active
vertices by degree:
low <= thread_max_deg < medium <= wave_max_deg < high.  That is the work of
giant_filter 0 and 1.  With giant_filter 2 the bins hold only the vertices
outside the sampled largest tree, which the device decides during the run:
the compute kernels' figures are then not known on the host (None).
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np

PER_UNIT = {
    "init": "4(n+1) + 4n bytes per graph",
    "sample": "16 bytes per sampled vertex (hooked entry and both labels read, its label written)",
    "mark": "4 + 1/8 bytes per vertex (label read, membership bit written)",
    "compute1": "4 bytes per entry + 8 bytes per vertex of the low-degree bin",
    "compute2": "4 bytes per entry + 8 bytes per vertex of the medium-degree bin",
    "compute3": "4 bytes per entry + 8 bytes per vertex of the high-degree bin",
    "flatten": "8 bytes per vertex",
}


def pass_bytes(n: int, m: int) -> int:
    return 4 * (n + 1) + 4 * m + 4 * n


def bins(nindex: np.ndarray, nlist: np.ndarray, thread_max: int, wave_max: int,
         sample_shift: int = 0, sample_min_deg: int = 0) -> Dict[str, tuple]:
    """(vertices, entries) of the active vertices in each degree bin."""
    n = nindex.shape[0] - 1
    deg = np.diff(nindex).astype(np.int64)
    has = deg > 0
    smallest = np.full(n, np.iinfo(np.int32).max, dtype=np.int64)
    if nlist.shape[0]:
        starts = nindex[:-1][has]
        smallest[has] = np.minimum.reduceat(nlist, starts)
    ids = np.arange(n)
    active = smallest < ids
    first = np.full(n, np.iinfo(np.int32).max, dtype=np.int64)
    second = np.full(n, np.iinfo(np.int32).max, dtype=np.int64)
    first[has] = nlist[nindex[:-1][has]]
    two = deg >= 2
    second[two] = nlist[nindex[:-1][two] + 1]
    pick = np.ones(n, dtype=bool)
    if sample_shift:
        h = (ids.astype(np.uint64) * np.uint64(0x9E3779B1)) & np.uint64(0xFFFFFFFF)
        pick = (h >> np.uint64(32 - sample_shift)) == 0
    pick &= deg >= sample_min_deg
    sampled = np.where(active, (second < ids) & pick, has & (first != ids))
    low = active & (deg <= thread_max)
    mid = active & (deg > thread_max) & (deg <= wave_max)
    hub = active & (deg > wave_max)
    return {
        "active": (int(active.sum()), int(deg[active].sum())),
        "sampled": (int(sampled.sum()), int(deg[sampled].sum())),
        "low": (int(low.sum()), int(deg[low].sum())),
        "mid": (int(mid.sum()), int(deg[mid].sum())),
        "hub": (int(hub.sum()), int(deg[hub].sum())),
    }


def kernel_bytes(nindex: np.ndarray, nlist: np.ndarray, thread_max: int, wave_max: int,
                 giant_filter: int = 1, sample_shift: int = 0,
                 sample_min_deg: int = 0) -> Dict[str, Optional[int]]:
    n = nindex.shape[0] - 1
    b = bins(nindex, nlist, thread_max, wave_max, sample_shift, sample_min_deg)
    per_bin: Dict[str, Optional[int]] = {k: 4 * e + 8 * v for k, (v, e) in b.items()}
    if giant_filter == 2:
        per_bin.update(low=None, mid=None, hub=None)
    return {
        "init": 4 * (n + 1) + 4 * n,
        "sample": 16 * b["sampled"][0],
        "hubscan": 0,
        "pick": 0,
        "mark": 4 * n + (n + 7) // 8,
        "compute1": per_bin["low"],
        "compute2": per_bin["mid"],
        "compute3": per_bin["hub"],
        "flatten": 8 * n,
    }
