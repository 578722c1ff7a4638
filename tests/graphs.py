"""Small hand-built graphs with known components, shared by the CPU oracle
tests and the GPU parity tests."""
from __future__ import annotations

import numpy as np

from eclcc import Graph


def from_undirected(n, pairs, name=""):
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    return Graph.from_edges(n, pairs[:, 0], pairs[:, 1], name=name)


def empty(n):
    return Graph(np.zeros(n + 1, dtype=np.int32), np.zeros(0, dtype=np.int32), name=f"empty{n}")


def path(n):
    return from_undirected(n, [(i, i + 1) for i in range(n - 1)], f"path{n}")


def star(n, centre=0):
    return from_undirected(n, [(centre, i) for i in range(n) if i != centre], f"star{n}-c{centre}")


def clique(n):
    return from_undirected(n, [(i, j) for i in range(n) for j in range(i + 1, n)], f"clique{n}")


def disjoint_cliques(k, size, shuffle_seed=None):
    pairs = []
    for c in range(k):
        b = c * size
        pairs += [(b + i, b + j) for i in range(size) for j in range(i + 1, size)]
    n = k * size
    pairs = np.asarray(pairs, dtype=np.int64)
    if shuffle_seed is not None:
        perm = np.random.default_rng(shuffle_seed).permutation(n)
        pairs = perm[pairs]
    return from_undirected(n, pairs, f"cliques{k}x{size}")


def expected_disjoint_cliques(k, size, shuffle_seed=None):
    n = k * size
    comp = np.arange(n) // size
    if shuffle_seed is not None:
        perm = np.random.default_rng(shuffle_seed).permutation(n)
        newcomp = np.empty(n, dtype=np.int64)
        newcomp[perm] = comp
        comp = newcomp
    out = np.empty(n, dtype=np.int32)
    for c in range(k):
        members = np.nonzero(comp == c)[0]
        out[members] = members.min()
    return out


def two_stars_bridged(n_leaves, hub_a, hub_b, n):
    """Two stars whose hubs are joined through a long path: hubs of high
    degree exercise the block-level kernel, the path exercises find."""
    pairs = []
    leaves = [v for v in range(n) if v not in (hub_a, hub_b)]
    rng = np.random.default_rng(7)
    rng.shuffle(leaves)
    a_leaves, b_leaves = leaves[:n_leaves], leaves[n_leaves:2 * n_leaves]
    rest = leaves[2 * n_leaves:]
    pairs += [(hub_a, x) for x in a_leaves]
    pairs += [(hub_b, x) for x in b_leaves]
    chain = [hub_a] + rest[:50] + [hub_b]
    pairs += [(chain[i], chain[i + 1]) for i in range(len(chain) - 1)]
    return from_undirected(n, pairs, "two-stars")
