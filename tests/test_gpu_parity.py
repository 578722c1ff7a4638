"""GPU parity: labels from libeclcc.so on an MI355X equal the oracle's,
bit for bit (integer work: tolerance zero)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import announce
import graphs as G
from eclcc import DeviceGraph, compute, compute_timed, default_params, workloads
from eclcc.cc import KERNELS
from oracle import oracle

pytestmark = pytest.mark.gpu


def check(g, labels):
    exp = oracle.cc_labels(g.nindex, g.nlist)
    assert labels.dtype == np.int32 and labels.shape == (g.nodes,)
    bad = np.nonzero(labels != exp)[0]
    assert bad.size == 0, f"{g.name}: {bad.size} labels differ, first at {bad[:5]}"


SMALL = [
    lambda: G.empty(1),
    lambda: G.empty(1000),
    lambda: G.path(2),
    lambda: G.path(5000),
    lambda: G.star(3000, 0),
    lambda: G.star(3000, 2999),
    lambda: G.star(20000, 12345),
    lambda: G.clique(90),
    lambda: G.disjoint_cliques(40, 12, shuffle_seed=3),
    lambda: G.two_stars_bridged(9000, 17, 30000, 40000),
    lambda: workloads.er65536(1),
    lambda: workloads.grid3d(32, 0.6, 1),
    lambda: workloads.grid3d(24, 1.0, 2),
    lambda: workloads.grid3d(24, 0.0, 2),
    lambda: workloads.rmat(14, 16, 5),
    lambda: workloads.rmat(16, 16, 6),
]


@pytest.mark.parametrize("make", SMALL, ids=[f"g{i}" for i in range(len(SMALL))])
def test_plain_matches_oracle(cuda, make):
    g = make()
    announce(f"plain eclcc_compute {g.name} n={g.nodes} m={g.edges}")
    labels, secs = compute_timed(g)
    assert secs > 0
    check(g, labels)


def test_disjoint_cliques_closed_form(cuda):
    g = G.disjoint_cliques(25, 7, shuffle_seed=11)
    announce(f"closed form {g.name}")
    assert np.array_equal(compute(g), G.expected_disjoint_cliques(25, 7, shuffle_seed=11))


PARAMS = [
    dict(thread_max_deg=0, wave_max_deg=1),          # nearly all work in compute3
    dict(thread_max_deg=1, wave_max_deg=64),
    dict(thread_max_deg=4, wave_max_deg=4096, hub_chunk=64),
    dict(thread_max_deg=64, wave_max_deg=64, block_threads=512, hub_chunk=8192),
    dict(thread_max_deg=16, wave_max_deg=704, block_threads=1024, hub_blocks=3),
    dict(thread_max_deg=1 << 20, wave_max_deg=1 << 20),  # all at lane granularity
    dict(giant_filter=0),
    dict(giant_filter=0, thread_max_deg=2, wave_max_deg=40, hub_chunk=128),
    dict(path_mode=0),
    dict(path_mode=0, giant_filter=0, thread_max_deg=8, wave_max_deg=100),
    dict(giant_filter=2),
    dict(giant_filter=2, thread_max_deg=0, wave_max_deg=1),
    dict(giant_filter=2, thread_max_deg=2, wave_max_deg=40, hub_chunk=128, path_mode=0),
    dict(giant_filter=2, thread_max_deg=1 << 20, wave_max_deg=1 << 20),
    dict(giant_filter=2, thread_max_deg=16, wave_max_deg=704, block_threads=1024, hub_blocks=3),
    dict(sample_shift=1),
    dict(sample_shift=4),
    dict(sample_shift=2, giant_filter=1),
    dict(sample_shift=3, giant_filter=0, path_mode=0),
    dict(path_mode=2),
    dict(path_mode=2, giant_filter=0, thread_max_deg=2, wave_max_deg=40, hub_chunk=128),
    dict(path_mode=2, giant_filter=1, sample_shift=2),
    dict(sample_min_deg=16, path_mode=2),
    dict(sample_min_deg=3, sample_shift=1, giant_filter=1),
    dict(sample_min_deg=1 << 30),
]


@pytest.mark.parametrize("kw", PARAMS, ids=[str(i) for i in range(len(PARAMS))])
def test_thresholds_do_not_change_labels(cuda, kw):
    g = workloads.rmat(15, 16, 9)
    announce(f"params {kw} on {g.name}")
    labels, _ = compute_timed(g, default_params(**kw))
    check(g, labels)


def test_init_kernel_closed_form(cuda):
    # unsorted lists: init takes the FIRST smaller neighbour in list order
    rng = np.random.default_rng(5)
    g = workloads.rmat(14, 16, 2)
    nl = g.nlist.copy()
    for v in range(g.nodes):
        b, e = g.nindex[v], g.nindex[v + 1]
        rng.shuffle(nl[b:e])
    g.nlist = nl
    announce(f"init only {g.name}")
    dg = DeviceGraph(g, "cuda:0")
    dg.run_init_only()
    assert np.array_equal(dg.labels(), oracle.init_labels(g.nindex, g.nlist))
    dg.run()
    check(g, dg.labels())


def test_init_finds_the_smaller_neighbour_at_every_list_position(cuda):
    # vertex 1000 + i has 40 entries, all larger than it except vertex i,
    # placed at position i: covers the first two entries, the per-lane
    # steps and the cooperative scan of longer lists
    from eclcc.graph import Graph
    L = 40
    src, dst = [], []
    for i in range(L):
        hub = 1000 + i
        src.append(hub); dst.append(i)
        for j in range(L - 1):
            src.append(hub); dst.append(2000 + 64 * i + j)
    g = Graph.from_edges(2000 + 64 * L, np.array(src), np.array(dst))
    nl = g.nlist.copy()
    for i in range(L):
        hub = 1000 + i
        b, e = g.nindex[hub], g.nindex[hub + 1]
        lst = [x for x in nl[b:e] if x != i]
        lst.insert(i, i)
        nl[b:e] = lst
    g.nlist = nl
    want = oracle.init_labels(g.nindex, g.nlist)
    assert all(want[1000 + i] == i for i in range(L))
    announce("init only, smaller neighbour at positions 0..39")
    dg = DeviceGraph(g, "cuda:0")
    dg.run_init_only()
    assert np.array_equal(dg.labels(), want)
    dg.run()
    check(g, dg.labels())


def test_filter2_without_a_largest_tree_on_unsorted_lists(cuda):
    # 2^16 vertices in components of four: the 256 evenly spaced samples of
    # pick all fall into different components, so no largest tree is found
    # and the work lists are the possibly active vertices.  Vertex 4c+1 has
    # the list [4c+2, 4c+3, 4c]: its only smaller neighbour sits past the two
    # entries init looks at under filter 2, so it starts as its own root and
    # must hook 4c from compute
    from eclcc.graph import Graph

    n = 1 << 16
    c = np.arange(n // 4) * 4
    src = np.concatenate([c + 1, c + 1, c + 1])
    dst = np.concatenate([c + 2, c + 3, c])
    g = Graph.from_edges(n, src, dst)
    nl = g.nlist.copy()
    for v in c + 1:
        b, e = g.nindex[v], g.nindex[v + 1]
        assert e - b == 3
        nl[b:e] = [v + 1, v + 2, v - 1]
    g.nlist = nl
    announce("filter 2, no largest tree, smaller neighbour at list position 2")
    for kw in (dict(), dict(sample_shift=2), dict(giant_filter=1), dict(giant_filter=0)):
        labels, _ = compute_timed(g, default_params(**kw))
        check(g, labels)
        assert np.array_equal(labels, np.repeat(c, 4)), kw


def test_device_runs_are_repeatable(cuda):
    g = workloads.rmat(16, 16, 12)
    announce(f"repeat {g.name}")
    dg = DeviceGraph(g, "cuda:0")
    first = None
    for _ in range(5):
        dg.run()
        lab = dg.labels()
        if first is None:
            first = lab
            check(g, lab)
        assert np.array_equal(lab, first)
    kms, tot = dg.run_profiled()
    assert len(kms) == len(KERNELS) and tot > 0
    assert np.array_equal(dg.labels(), first)


def test_full_size_grid256(cuda):
    g = workloads.grid256()
    announce(f"full size {g.name} n={g.nodes} m={g.edges}")
    dg = DeviceGraph(g, "cuda:0")
    dg.run()
    lab = dg.labels()
    check(g, lab)
    other = 1 if dg.params.giant_filter == 2 else 2
    announce(f"full size {g.name} giant_filter={other}")
    dg2 = DeviceGraph(g, "cuda:0", default_params(giant_filter=other))
    dg2.run()
    assert np.array_equal(dg2.labels(), lab)
    for kw in (dict(path_mode=2), dict(path_mode=2, sample_shift=1)):
        announce(f"full size {g.name} {kw}")
        dg2.set_params(default_params(**kw))
        dg2.run()
        assert np.array_equal(dg2.labels(), lab), kw
    del dg2
    # size-independent properties
    assert np.all(lab <= np.arange(g.nodes)) and np.all(lab[lab] == lab)


def test_full_size_rmat24(cuda):
    g = workloads.rmat24()
    first = None
    for filt, mode in ((1, 1), (2, 1), (2, 2)):
        announce(f"full size {g.name} n={g.nodes} m={g.edges} giant_filter={filt} path_mode={mode}")
        dg = DeviceGraph(g, "cuda:0", default_params(giant_filter=filt, path_mode=mode))
        dg.run()
        lab = dg.labels()
        hubs, tasks = dg.hub_stats()
        if first is None:
            check(g, lab)
            first = lab
            # every high-degree vertex with a smaller neighbour has tasks
            assert hubs > 0 and tasks >= hubs
        else:
            assert np.array_equal(lab, first)
            # only those outside the largest tree
            assert tasks >= hubs >= 0
        del dg


def test_filter_and_path_modes_agree_full_rmat20(cuda):
    g = workloads.rmat(20, 16, 31)
    announce(f"filter on/off, path modes {g.name}")
    ref = compute(g, default_params(giant_filter=1))
    check(g, ref)
    for kw in (dict(giant_filter=0), dict(path_mode=0), dict(path_mode=0, giant_filter=0),
               dict(giant_filter=2), dict(giant_filter=2, path_mode=0), dict(path_mode=2),
               dict(path_mode=2, giant_filter=0), dict(path_mode=2, sample_shift=1)):
        assert np.array_equal(compute(g, default_params(**kw)), ref), kw


@pytest.mark.parametrize("mode,filt", [(0, 1), (1, 1), (0, 2), (1, 2), (1, 0), (2, 2), (2, 1), (2, 0)])
def test_path_and_filter_modes_on_grid_and_er(cuda, mode, filt):
    for g in (workloads.grid3d(48, 0.6, 4), workloads.er65536(2), workloads.grid3d(40, 0.25, 5)):
        announce(f"path_mode={mode} giant_filter={filt} {g.name}")
        labels, _ = compute_timed(g, default_params(path_mode=mode, giant_filter=filt))
        check(g, labels)


def _star_plus(n, center, leaves, extra_src=(), extra_dst=()):
    src = np.concatenate([np.full(len(leaves), center), np.asarray(extra_src, dtype=np.int64)])
    dst = np.concatenate([np.asarray(leaves), np.asarray(extra_dst, dtype=np.int64)])
    return G.from_undirected(n, np.stack([src, dst], axis=1))


@pytest.mark.parametrize("filt", [1, 2])
def test_high_degree_vertices_outside_the_largest_tree(cuda, filt):
    # samples of the root pick sit at multiples of n/256 = 1000; the stars
    # avoid them, so (a) no graph-wide root is found at all, or (b) a root is
    # found in a path through the sampled vertices but the star lies outside
    n = 256000
    leaves = [v for v in range(1, 3004) if v % 1000]
    a = _star_plus(n, 255999, leaves)
    path = [t * 1000 for t in range(256)]
    b = _star_plus(n, 255999, leaves, path[:-1], path[1:])
    for g, name in ((a, "no sampled root"), (b, "star outside the sampled tree")):
        announce(f"giant_filter={filt}: {name}")
        for kw in (dict(), dict(thread_max_deg=0, wave_max_deg=1), dict(wave_max_deg=100)):
            check(g, compute(g, default_params(giant_filter=filt, **kw)))


@pytest.mark.parametrize("kw", [dict(), dict(hub_chunk=64), dict(hub_chunk=128, wave_max_deg=16, path_mode=0),
                                dict(path_mode=2, sample_shift=1), dict(giant_filter=1)],
                         ids=["default", "chunk64", "wave16-mode0", "mode2-shift1", "filter1"])
def test_large_stars_outside_the_largest_tree(cuda, kw):
    # three stars of 50,000 leaves beside a random graph of 200,000 vertices:
    # with filter 2 mark lists the three centres (outside the largest tree)
    # with their edge tasks, and compute3's blocks hook their lists in
    # tasks of hub_chunk entries
    from eclcc import params_for_graph

    g = G.er_with_stars(200_000, 800_000, 3, 50_000, seed=5)
    p = params_for_graph(g, **kw)
    announce(f"{g.name} n={g.nodes} m={g.edges} params {kw}")
    dg = DeviceGraph(g, "cuda:0", p)
    for rep in range(2):
        dg.nstat.fill_(-3)
        dg.run()
        check(g, dg.labels())
    hubs, tasks = dg.hub_stats()
    if p.giant_filter == 2:
        chunk = p.hub_chunk
        assert hubs >= 3 and tasks >= 3 * ((50_000 + chunk - 1) // chunk), (hubs, tasks)


def test_cli_runs_like_the_reference_minibench(cuda, tmp_path):
    # the process face: `eclcc <file.egr>` prints the reference's lines, and
    # its analyzer reads the time with `time: +(\d+\.\d+) s`
    # (scripts/analyze_results.py:62)
    import re
    import subprocess

    from eclcc import _native as N

    g = G.disjoint_cliques(30, 9, shuffle_seed=4)
    g2 = workloads.rmat(13, 8, 2)
    for graph, flags in ((g, []), (g2, ["--check"])):
        path = tmp_path / "g.egr"
        graph.write(str(path))
        announce(f"eclcc CLI {flags} on n={graph.nodes}")
        r = subprocess.run([str(N.BIN_DIR / "eclcc"), *flags, str(path)], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        out = r.stdout
        assert "verification passed" in out
        assert re.search(r"time: +(\d+\.\d+) s", out)
        want = oracle.count_components(oracle.cc_labels(graph.nindex, graph.nlist))
        assert f"number of connected components: {want}\n" in out


def test_lists_ending_at_the_largest_int32_edge_count(cuda):
    # m = 2^31 - 1 CSR entries (8 GiB of nlist), the most an int32 edge
    # count allows: the last non-empty list ends at entry 2^31 - 2 and is
    # scanned by init to its end, then hooked at lane (compute1), wavefront
    # (compute2) and block (compute3) granularity; the last vertex's empty
    # list starts at 2^31 - 1.  List positions one step past the end must
    # not wrap.  Labels by construction (two paths of repeated edges, a
    # vertex joined to the second path, an isolated vertex).
    import torch

    g, expected = G.max_entries_multigraph()
    assert g.edges == 2**31 - 1 and int(g.nindex[-2]) == 2**31 - 1
    announce(f"plain eclcc_compute host arrays {g.name} n={g.nodes}")
    assert np.array_equal(compute(g), expected)
    dg = DeviceGraph(g, "cuda:0")
    try:
        for kw in (dict(), dict(giant_filter=0), dict(giant_filter=1, path_mode=0),
                   dict(giant_filter=0, thread_max_deg=128, wave_max_deg=256),
                   dict(giant_filter=2, thread_max_deg=128, wave_max_deg=256),
                   dict(giant_filter=0, thread_max_deg=16, wave_max_deg=64),
                   dict(giant_filter=2, thread_max_deg=16, wave_max_deg=64, hub_chunk=64)):
            announce(f"resident {g.name} params {kw}")
            dg.set_params(default_params(**kw))
            dg.nstat.fill_(-7)
            dg.run()
            got = dg.labels()
            bad = np.nonzero(got != expected)[0]
            assert bad.size == 0, f"{kw}: {bad.size} labels differ, first at {bad[:5]}: {got[bad[:5]]}"
    finally:
        del dg
        torch.cuda.empty_cache()


@pytest.mark.parametrize("kw", [dict(), dict(giant_filter=0, thread_max_deg=1, wave_max_deg=8, hub_chunk=64),
                                dict(giant_filter=1, path_mode=0), dict(sample_shift=2)],
                         ids=["default", "f0-small-bins", "f1-halving", "shift2"])
def test_plain_on_irregular_graphs(cuda, kw):
    # unsorted lists, duplicates, self-loops, hubs at any ID, permuted paths,
    # shuffled cliques, 1 to 3000 vertices
    announce(f"plain eclcc_compute on irregular graphs, params {kw}")
    for seed in range(200, 232):
        g = G.random_structure(seed)
        check(g, compute(g, default_params(**kw)))
