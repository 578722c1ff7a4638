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


def er_with_stars(n_er, pairs, stars, leaves, seed):
    """Stars whose centres lie outside the largest tree (the package's
    generator, also a bench workload)."""
    from eclcc import workloads

    return workloads.er_with_stars(n_er, pairs, stars, leaves, seed)


INT32_MAX = 2**31 - 1


def max_entries_multigraph(chain=40, split=20, tail_loops=101, x=30, edges=INT32_MAX):
    """A symmetric multigraph with `edges` CSR entries (default 2^31 - 1, the
    most an int32 edge count allows), built so that lists end right at that
    limit.  Vertices 0..chain-1 form two paths, 0..split-1 and split..chain-1;
    each path edge is repeated c times (the first edge of each path once, so
    that the two path roots have short lists, and the last edge takes the
    rest of the count).  Vertex z = chain has `tail_loops` self-loops and then
    one edge to x (x's list ends with z): z has no smaller neighbour among
    its first entries, so init scans its list up to the final entry.
    Vertex z + 1 is isolated, with its empty list at position `edges`.
    Labels: 0 on the first path, split on the second and on z, z + 1 on z + 1.
    Returns (graph, expected labels)."""
    assert 2 < split < chain - 2 and split <= x < chain
    z = chain
    base = 2 + 2 + 2 + tail_loops  # edges 0-1 and split-(split+1) once, z-x, z's loops
    rest = edges - base
    assert rest > 0 and rest % 2 == 0, "edges - tail_loops must be odd"
    many = [i for i in range(chain - 1) if i not in (0, split - 1, split, chain - 2)]
    c = rest // 2 // (len(many) + 1)
    c_last = rest // 2 - c * len(many)
    assert c >= 1 and c_last >= 1

    def mult(i):  # multiplicity of path edge (i, i + 1)
        if i == split - 1:
            return 0
        if i in (0, split):
            return 1
        return c_last if i == chain - 2 else c

    vert, neigh, cnt = [], [], []
    for v in range(chain):
        for u, k in ((v - 1, mult(v - 1) if v > 0 else 0), (v + 1, mult(v) if v < chain - 1 else 0)):
            if k:
                vert.append(v), neigh.append(u), cnt.append(k)
        if v == x:
            vert.append(v), neigh.append(z), cnt.append(1)
    vert += [z, z]
    neigh += [z, x]
    cnt += [tail_loops, 1]
    cnt = np.asarray(cnt, dtype=np.int64)
    n = chain + 2
    deg = np.bincount(np.asarray(vert), weights=cnt, minlength=n).astype(np.int64)
    nindex = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(deg, out=nindex[1:])
    assert nindex[-1] == edges
    nlist = np.repeat(np.asarray(neigh, dtype=np.int32), cnt)
    g = Graph(nindex.astype(np.int32), nlist, name=f"multigraph-m{edges}")
    expected = np.empty(n, dtype=np.int32)
    expected[:split] = 0
    expected[split:chain] = split
    expected[z] = split
    expected[z + 1] = z + 1
    return g, expected


def random_structure(seed, n_hi=3000):
    """A seeded small graph of an irregular shape, in the CSR form the
    reference accepts but its generators never write: lists in random order,
    duplicate entries and self-loops.  The shape is one of random pairs,
    stars (a hub at a random ID, some leaves cross-linked), a permuted path,
    or shuffled cliques; isolated vertices occur in all of them."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, n_hi))
    kind = seed % 4
    if kind == 0:
        k = int(rng.integers(0, 4 * n + 1))
        a, b = rng.integers(0, n, k), rng.integers(0, n, k)
    elif kind == 1:
        hub = int(rng.integers(0, n))
        leaves = rng.choice(n, size=int(rng.integers(0, n)), replace=False)
        extra = int(rng.integers(0, n // 4 + 1))
        a = np.concatenate([np.full(leaves.size, hub), rng.choice(n, extra)])
        b = np.concatenate([leaves, rng.choice(n, extra)])
    elif kind == 2:
        perm = rng.permutation(n)
        cut = rng.random(max(n - 1, 0)) < 0.9  # keep most path edges
        a, b = perm[:-1][cut], perm[1:][cut]
    else:
        size = int(rng.integers(2, 9))
        perm = rng.permutation(n)
        pa, pb = [], []
        for c0 in range(0, n - n % size, size):
            grp = perm[c0:c0 + size]
            i, j = np.triu_indices(size, 1)
            pa.append(grp[i])
            pb.append(grp[j])
        a = np.concatenate(pa) if pa else np.zeros(0, np.int64)
        b = np.concatenate(pb) if pb else np.zeros(0, np.int64)
    dup = rng.random(a.size) < 0.1  # some edges twice
    a = np.concatenate([a, a[dup]])
    b = np.concatenate([b, b[dup]])
    src = np.concatenate([a, b]).astype(np.int64)
    dst = np.concatenate([b, a]).astype(np.int64)
    order = rng.permutation(src.size)  # random order within each list
    src, dst = src[order], dst[order]
    by = np.argsort(src, kind="stable")
    nlist = dst[by].astype(np.int32)
    nindex = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=nindex[1:])
    return Graph(nindex.astype(np.int32), nlist, name=f"random-structure-{seed}-k{kind}-n{n}")
