"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the
checker of the GPU path: golden vectors from scipy, closed forms, an
independent BFS, and planted mistakes the restated verifier must catch."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import graphs as G
from eclcc import Graph, workloads
from oracle import oracle

GOLDEN = sorted((Path(__file__).parent / "golden").glob("*.npz"))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return z["nindex"], z["nlist"], z["labels"]


def test_fixtures_present():
    assert len(GOLDEN) >= 10


@pytest.mark.parametrize("path", GOLDEN, ids=[p.stem for p in GOLDEN])
def test_oracle_matches_golden(path):
    ni, nl, lab = load(path)
    assert np.array_equal(oracle.cc_labels(ni, nl), lab)
    assert np.array_equal(oracle.cc_bfs(ni, nl), lab)
    assert np.array_equal(oracle.eclcc_serial(ni, nl), lab)
    for t in (1, 3, 8):
        got, used = oracle.eclcc_threads(ni, nl, t)
        assert used == t and np.array_equal(got, lab)
    rc, nc = oracle.verify(ni, nl, lab)
    assert rc == 0 and nc == len(np.unique(lab))
    assert oracle.count_components(lab) == nc


def test_closed_forms():
    n = 500
    assert np.array_equal(oracle.cc_labels(*csr(G.empty(n))), np.arange(n))
    assert not oracle.cc_labels(*csr(G.path(n))).any()
    assert not oracle.cc_labels(*csr(G.star(n, n - 1))).any()
    assert not oracle.cc_labels(*csr(G.clique(30))).any()
    g = G.disjoint_cliques(17, 6, shuffle_seed=9)
    assert np.array_equal(oracle.cc_labels(*csr(g)), G.expected_disjoint_cliques(17, 6, 9))
    full = workloads.grid3d(9, 1.0, 1)
    assert not oracle.eclcc_serial(*csr(full)).any()
    none = workloads.grid3d(9, 0.0, 1)
    assert np.array_equal(oracle.eclcc_serial(*csr(none)), np.arange(9 ** 3))


def csr(g):
    return g.nindex, g.nlist


def test_init_closed_form_by_hand():
    # vertex: list (in stored order) -> first smaller neighbour, else itself
    lists = {0: [3, 1], 1: [4, 0, 2], 2: [1, 3], 3: [4, 0, 2], 4: [3, 1]}
    expect = [0, 0, 1, 0, 3]
    ni = np.zeros(6, dtype=np.int32)
    nl = []
    for v in range(5):
        nl += lists[v]
        ni[v + 1] = len(nl)
    assert oracle.init_labels(ni, np.array(nl, dtype=np.int32)).tolist() == expect
    # a vertex whose smaller neighbour sits far down an unsorted list
    ni = np.array([0, 1, 8, 9, 10, 11, 12, 13, 14], dtype=np.int32)
    nl = np.array([1, 7, 6, 5, 4, 3, 2, 0, 1, 1, 1, 1, 1, 1], dtype=np.int32)
    assert oracle.init_labels(ni, nl).tolist() == [0, 0, 1, 1, 1, 1, 1, 1]


def test_bfs_and_unionfind_agree_on_random_graphs():
    for g in (workloads.rmat(13, 16, 3), Graph.er(5000, 4000, 4), workloads.grid3d(15, 0.3, 5)):
        a = oracle.cc_labels(*csr(g))
        assert np.array_equal(a, oracle.cc_bfs(*csr(g)))
        assert np.array_equal(a, oracle.eclcc_serial(*csr(g)))
        assert np.array_equal(a, oracle.eclcc_threads(*csr(g), 6)[0])
        assert oracle.verify(*csr(g), a)[0] == 0


def test_brute_force_tiny():
    # every graph on 5 vertices (all 1024 edge subsets): labels by closure
    pairs = [(i, j) for i in range(5) for j in range(i + 1, 5)]
    for mask in range(1 << len(pairs)):
        keep = [p for k, p in enumerate(pairs) if (mask >> k) & 1]
        g = G.from_undirected(5, keep if keep else np.zeros((0, 2)))
        reach = np.eye(5, dtype=bool)
        for a, b in keep:
            reach[a, b] = reach[b, a] = True
        for k in range(5):  # Warshall closure
            reach |= reach[:, [k]] & reach[[k], :]
        expect = np.array([np.nonzero(reach[v])[0].min() for v in range(5)], dtype=np.int32)
        ni, nl = csr(g)
        assert np.array_equal(oracle.cc_labels(ni, nl), expect), keep
        assert np.array_equal(oracle.cc_bfs(ni, nl), expect), keep
        assert np.array_equal(oracle.eclcc_serial(ni, nl), expect), keep
        assert oracle.verify(ni, nl, expect) == (0, len(set(expect.tolist())))


def plant(path):
    ni, nl, lab = load(path)
    return ni, nl, lab.copy()


def test_verifier_catches_planted_mistakes():
    ni, nl, lab = plant(GOLDEN[[p.stem for p in GOLDEN].index("er_2000_m5000")])
    assert oracle.verify(ni, nl, lab)[0] == 0
    comps, sizes = np.unique(lab, return_counts=True)
    big = comps[np.argmax(sizes)]
    members = np.nonzero(lab == big)[0]
    # split: one member of a non-singleton component gets its own ID
    bad = lab.copy()
    bad[members[-1]] = members[-1]
    assert oracle.verify(ni, nl, bad)[0] == -1
    # merge: two components share one label
    small = [c for c in comps if c != big][0]
    bad = lab.copy()
    bad[lab == small] = big
    assert oracle.verify(ni, nl, bad)[0] == -5
    # right partition, wrong representative (maximum instead of minimum)
    bad = lab.copy()
    bad[members] = members.max()
    assert oracle.verify(ni, nl, bad)[0] == -7
    bad = lab.copy()
    bad[3] = -1
    assert oracle.verify(ni, nl, bad)[0] == -2
    bad = lab.copy()
    bad[3] = len(lab)
    assert oracle.verify(ni, nl, bad)[0] == -6


def test_max_entries_multigraph_builder_small():
    # the builder of the 2^31 - 1 entry GPU case, at a size the oracle
    # checks in milliseconds: symmetric, lists where stated, labels as stated
    g, exp = G.max_entries_multigraph(chain=12, split=5, tail_loops=11, x=7, edges=10001)
    assert g.edges == 10001 and int(g.nindex[-2]) == 10001
    src = np.repeat(np.arange(g.nodes, dtype=np.int64), np.diff(g.nindex))
    fwd = np.sort(src * g.nodes + g.nlist)
    rev = np.sort(g.nlist.astype(np.int64) * g.nodes + src)
    assert np.array_equal(fwd, rev)
    assert np.array_equal(oracle.cc_labels(g.nindex, g.nlist), exp)
    z = 12
    assert list(g.nlist[g.nindex[z]:g.nindex[z + 1]]) == [z] * 11 + [7]


def test_random_structure_builder_is_symmetric_and_oracle_equals_scipy():
    # the irregular graphs of the batched GPU test: symmetric CSR (lists
    # unsorted, with duplicates and self-loops), and the oracle's labels
    # equal scipy's components canonicalised to minimum IDs
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    for seed in range(16):
        g = G.random_structure(seed)
        n = g.nodes
        src = np.repeat(np.arange(n, dtype=np.int64), np.diff(g.nindex))
        assert np.array_equal(np.sort(src * n + g.nlist), np.sort(g.nlist.astype(np.int64) * n + src))
        a = sp.csr_matrix((np.ones(g.edges), (src, g.nlist)), shape=(n, n))
        c, lab = cg.connected_components(a, directed=False)
        mins = np.full(c, n)
        np.minimum.at(mins, lab, np.arange(n))
        assert np.array_equal(oracle.cc_labels(g.nindex, g.nlist), mins[lab]), g.name
