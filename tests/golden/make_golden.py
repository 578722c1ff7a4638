"""Writes the golden CC fixtures under tests/golden/*.npz.

The reference ships no golden vectors or input graphs for this path (its
inputs are downloaded, download_inputs.sh:11-27, and its only check is the
self-verifier, ECL-CC-atomic.cu:365-395).  These fixtures pin the oracle
against an independent implementation: labels come from
scipy.sparse.csgraph.connected_components (scipy 1.15), canonicalised to the
ECL-CC output form label[v] = minimum vertex ID of v's component.

Each fixture holds int32 arrays nindex, nlist and labels.  Run from the repo
root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent / "ecl-suite_amd"))
sys.path.insert(0, str(HERE.parent))

from eclcc import Graph, workloads  # noqa: E402
import graphs as G  # noqa: E402


def scipy_min_labels(nindex, nlist):
    n = nindex.shape[0] - 1
    a = sp.csr_matrix((np.ones(nlist.shape[0], dtype=np.int8), nlist, nindex), shape=(n, n))
    _, comp = connected_components(a, directed=False)
    mins = np.full(comp.max() + 1, n, dtype=np.int64)
    np.minimum.at(mins, comp, np.arange(n))
    return mins[comp].astype(np.int32)


def unsorted_with_dups(seed=4):
    """Adjacency in random order, duplicate entries and self-loops (the
    reference tolerates both: a duplicate re-hook is a no-op and v > v is
    false)."""
    rng = np.random.default_rng(seed)
    n = 300
    src = rng.integers(0, n, 500)
    dst = rng.integers(0, n, 500)
    g = Graph.from_edges(n, np.concatenate([src, dst, [5, 5, 7]]),
                         np.concatenate([dst, src, [5, 9, 7]]), flags=0, name="unsorted-dups")
    nl = g.nlist.copy()
    for v in range(n):
        rng.shuffle(nl[g.nindex[v]:g.nindex[v + 1]])
    g.nlist = nl
    return g


def cases():
    yield "single_vertex", G.empty(1)
    yield "isolated_100", G.empty(100)
    yield "path_257", G.path(257)
    yield "star_1200_hub_last", G.star(1200, 1199)
    yield "clique_40", G.clique(40)
    yield "cliques_30x9_shuffled", G.disjoint_cliques(30, 9, shuffle_seed=2)
    yield "two_stars_bridged", G.two_stars_bridged(1500, 3, 2999, 4000)
    yield "er_2000_m5000", Graph.er(2000, 5000, 11)
    yield "er_3000_sparse", Graph.er(3000, 1400, 12)
    yield "rmat10_ef8", workloads.rmat(10, 8, 13)
    yield "rmat12_ef16", workloads.rmat(12, 16, 14)
    yield "grid12_p05", workloads.grid3d(12, 0.5, 15)
    yield "grid10_p025", workloads.grid3d(10, 0.25, 16)
    yield "unsorted_dups_selfloops", unsorted_with_dups()


def main():
    for name, g in cases():
        lab = scipy_min_labels(g.nindex, g.nlist)
        np.savez_compressed(HERE / f"{name}.npz", nindex=g.nindex.astype(np.int32),
                            nlist=g.nlist.astype(np.int32), labels=lab)
        print(f"{name}: n={g.nodes} m={g.edges} components={len(np.unique(lab))}")


if __name__ == "__main__":
    main()
