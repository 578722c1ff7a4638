"""libeclgraph.so: .egr format compatibility with the reference reader and
writer, and the properties of the seeded generators."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

from eclcc import EclgraphError, Graph, workloads
from eclcc.graph import DEDUPE, DROP_SELF, SYMMETRIZE, rand64, set_generator_threads
from oracle import oracle


def assert_clean_symmetric(g, sorted_lists=True):
    n = g.nodes
    assert g.nindex[0] == 0 and g.nindex[-1] == g.edges
    assert np.all(np.diff(g.nindex) >= 0)
    assert g.nlist.min(initial=0) >= 0 and g.nlist.max(initial=0) < n
    src = np.repeat(np.arange(n), np.diff(g.nindex))
    assert not np.any(src == g.nlist), "self-loop"
    fwd = src.astype(np.int64) * n + g.nlist
    bwd = g.nlist.astype(np.int64) * n + src
    assert np.array_equal(np.sort(fwd), np.sort(bwd)), "not symmetric"
    if sorted_lists:
        same = src[1:] == src[:-1]
        assert np.all(g.nlist[1:][same] > g.nlist[:-1][same]), "unsorted or duplicate"


def test_er_exact_pairs_and_clean():
    g = Graph.er(5000, 20000, 3)
    assert g.nodes == 5000 and g.edges == 40000
    assert_clean_symmetric(g)


def test_er65536_is_config_size():
    g = workloads.er65536(1)
    assert (g.nodes, g.edges) == (65536, 524288)
    assert_clean_symmetric(g)
    assert g.algorithmic_bytes() == 4 * 65537 + 4 * 524288 + 4 * 65536


def test_grid_edge_counts():
    s = 11
    full = Graph.grid3d(s, s, s, 1.0, 1, permute=True)
    assert full.edges == 2 * 3 * s * s * (s - 1)
    assert_clean_symmetric(full)
    assert Graph.grid3d(s, s, s, 0.0, 1).edges == 0
    part = workloads.grid3d(20, 0.6, 2)
    expect = 2 * 3 * 20 * 20 * 19 * 0.6
    assert abs(part.edges - expect) < 0.05 * expect
    assert_clean_symmetric(part)
    # unpermuted grid: neighbours of (0,0,0) are 1, s, s*s
    plain = Graph.grid3d(s, s, s, 1.0, 1, permute=False)
    assert plain.nlist[plain.nindex[0]:plain.nindex[1]].tolist() == [1, s, s * s]


def test_rmat_clean_and_deterministic():
    a = workloads.rmat(13, 16, 7)
    assert_clean_symmetric(a)
    set_generator_threads(1)
    try:
        b = workloads.rmat(13, 16, 7)
    finally:
        set_generator_threads(0)
    assert np.array_equal(a.nindex, b.nindex) and np.array_equal(a.nlist, b.nlist)
    c = workloads.rmat(13, 16, 8)
    assert not (a.edges == c.edges and np.array_equal(a.nlist, c.nlist))
    d = a.degrees()
    assert d.max() > 20 * d.mean(), "R-MAT degrees should be skewed"


def test_rmat_non_power_of_two():
    g = Graph.rmat(12, 8 * 3000, 5, n=3000)
    assert g.nodes == 3000
    assert_clean_symmetric(g)


def test_batch_mix_spec():
    from eclcc.workloads import mix_spec

    specs = [mix_spec(i) for i in range(2000)]
    ns = np.array([s.n for s in specs])
    assert ns.min() >= 1000 and ns.max() <= 1_000_000
    # log-uniform: median near sqrt(1e3*1e6) ~ 31623
    assert 20000 < np.median(ns) < 50000
    kinds = [s.kind for s in specs]
    assert 0.45 < kinds.count("er") / len(kinds) < 0.55
    assert mix_spec(17) == mix_spec(17)


def test_rand64_is_splitmix64():
    # splitmix64 from state 0: first output is 0xE220A8397B1DCDAF
    assert rand64(0, 0) == 0xE220A8397B1DCDAF


def test_egr_roundtrip(tmp_path):
    g = workloads.rmat(10, 8, 2)
    p = tmp_path / "g.egr"
    g.write(p)
    h = Graph.read(p)
    assert np.array_equal(g.nindex, h.nindex) and np.array_equal(g.nlist, h.nlist)
    assert h.eweight is None
    raw = np.fromfile(p, dtype="<i4")
    assert raw[0] == g.nodes and raw[1] == g.edges and raw.size == 2 + g.nodes + 1 + g.edges
    g.eweight = np.arange(g.edges, dtype=np.int32)
    g.write(p)
    assert np.array_equal(Graph.read(p).eweight, g.eweight)


def test_reader_errors(tmp_path):
    with pytest.raises(EclgraphError):
        Graph.read(tmp_path / "missing.egr")
    p = tmp_path / "bad.egr"
    np.array([0, 0, 0], dtype="<i4").tofile(p)
    with pytest.raises(EclgraphError):
        Graph.read(p)  # nodes < 1
    np.array([4, 10, 0, 1, 2], dtype="<i4").tofile(p)
    with pytest.raises(EclgraphError):
        Graph.read(p)  # truncated index


@pytest.mark.skipif(not oracle.REF_ROUNDTRIP.exists(), reason="reference ECLgraph.h driver not built")
def test_reference_reader_writer_interop(tmp_path):
    """Files written here are read by the reference's readECLgraph, and what
    the reference's writeECLgraph writes back is byte-identical."""
    for g in (workloads.grid3d(8, 0.5, 1), workloads.rmat(11, 16, 3)):
        ours = tmp_path / "ours.egr"
        theirs = tmp_path / "theirs.egr"
        g.write(ours)
        out = subprocess.run([str(oracle.REF_ROUNDTRIP), str(ours), str(theirs)], check=True,
                             capture_output=True, text=True).stdout.split()
        assert int(out[0]) == g.nodes and int(out[1]) == g.edges and out[2] == "0"
        assert ours.read_bytes() == theirs.read_bytes()
        back = Graph.read(theirs)
        assert np.array_equal(back.nlist, g.nlist)


def test_mtx_reader(tmp_path):
    p = tmp_path / "m.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n% comment\n"
                 "5 5 4\n2 1\n3 2\n5 4\n4 4\n")
    g = Graph.read_mtx(p, SYMMETRIZE | DROP_SELF | DEDUPE)
    assert g.nodes == 5
    assert_clean_symmetric(g)
    assert oracle.cc_labels(g.nindex, g.nlist).tolist() == [0, 0, 0, 3, 3]
    q = tmp_path / "r.mtx"
    q.write_text("%%MatrixMarket matrix coordinate real general\n4 4 2\n1 4 2.5\n3 2 1e3\n")
    h = Graph.read_mtx(q, SYMMETRIZE | DROP_SELF | DEDUPE)
    assert h.nodes == 4
    assert_clean_symmetric(h)
    # the weights travel with their entries: (0,3) and (3,0) weigh 250
    assert h.eweight.tolist() == [250, 100000, 100000, 250]


def test_from_edges_flags():
    g = Graph.from_edges(4, [0, 1, 1, 2, 3], [1, 0, 1, 2, 2], flags=0)
    assert g.edges == 5  # kept as given
    h = Graph.from_edges(4, [0, 1, 1, 2, 3], [1, 0, 1, 2, 2])
    assert_clean_symmetric(h)
    assert h.edges == 4


def test_check_accepts_generated_and_flags_broken_csr():
    from eclcc.graph import GraphFormatError

    for g in (workloads.er65536(), workloads.rmat(10, 8, seed=3), workloads.grid3d(8, 0.5, 2)):
        g.check()
    # the reverse of entry 2->5 removed: reported at that entry's position
    g = Graph.from_edges(8, [0, 1, 2, 3], [1, 2, 5, 4])
    src = np.repeat(np.arange(g.nodes), np.diff(g.nindex))
    drop = int(np.flatnonzero((src == 5) & (g.nlist == 2))[0])
    keep = np.ones(g.edges, dtype=bool)
    keep[drop] = False
    ni = g.nindex.copy()
    ni[6:] -= 1
    asym = Graph(ni, g.nlist[keep])
    with pytest.raises(GraphFormatError) as e:
        asym.check()
    pos = int(np.flatnonzero((np.repeat(np.arange(8), np.diff(ni)) == 2) & (asym.nlist == 5))[0])
    assert e.value.where == pos and "symmetric" in str(e.value)
    asym.check(symmetry=False)  # index and range are fine
    bad = Graph(g.nindex.copy(), g.nlist.copy())
    bad.nlist[3] = 8
    with pytest.raises(GraphFormatError) as e:
        bad.check(symmetry=False)
    assert e.value.where == 3
    bad = Graph(g.nindex.copy(), g.nlist.copy())
    bad.nindex[4] = bad.nindex[5] + 1
    with pytest.raises(GraphFormatError) as e:
        bad.check()
    assert e.value.where == 5 and "nindex decreases" in str(e.value)


def test_cli_check_flag(tmp_path):
    from eclcc import _native as N

    exe = N.BIN_DIR / "eclcc"
    g = Graph.from_edges(4, [0, 2], [1, 3])
    ok = tmp_path / "ok.egr"
    g.write(str(ok))
    asym = tmp_path / "asym.egr"
    Graph(np.array([0, 1, 1, 1, 1], dtype=np.int32), np.array([1], dtype=np.int32)).write(str(asym))
    r = subprocess.run([str(exe), "--bogus", str(ok)], capture_output=True, text=True)
    assert r.returncode == 255 and "USAGE:" in r.stderr
    r = subprocess.run([str(exe), "--check", str(asym)], capture_output=True, text=True)
    assert r.returncode == 255 and "not symmetric" in r.stderr


def _raw_egr(path, nindex, nlist):
    """An .egr file with exactly these arrays (no checks on the way)."""
    ni = np.asarray(nindex, dtype=np.int32)
    nl = np.asarray(nlist, dtype=np.int32)
    with open(path, "wb") as f:
        np.array([ni.size - 1, nl.size], dtype=np.int32).tofile(f)
        ni.tofile(f)
        nl.tofile(f)


@pytest.mark.parametrize("exe_name", ["eclcc", "eclmst"])
def test_clis_reject_broken_index_and_ids_before_the_gpu(tmp_path, exe_name):
    # a decreasing nindex and an entry outside [0, n) would index the label
    # array and the lists out of bounds on the GPU; both CLIs stop on the
    # host with the reference's ERROR line and exit status -1, without --check
    from eclcc import _native as N

    exe = N.BIN_DIR / exe_name
    cases = {
        "decreasing": ([0, 2, 1, 3, 4], [1, 2, 0, 3], "nindex decreases"),
        "out_of_range": ([0, 1, 2, 3, 4], [1, 0, 7, 2], "out of range"),
        "negative": ([0, 1, 2, 3, 4], [1, 0, -3, 2], "out of range"),
    }
    for name, (ni, nl, msg) in cases.items():
        path = tmp_path / f"{name}.egr"
        _raw_egr(path, ni, nl)
        r = subprocess.run([str(exe), str(path)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 255, (name, r.stdout, r.stderr)
        assert "ERROR:" in r.stderr and msg in r.stderr, (name, r.stderr)
        assert "compute time" not in r.stdout and "Dev" not in r.stdout


def test_kernel_bytes_partition_the_active_entries():
    from eclcc import roofline

    g = workloads.rmat(12, 16, seed=3)
    b = roofline.bins(g.nindex, g.nlist, 16, 2048)
    # hand count with a plain loop
    lo = mid = hub = 0
    ent = [0, 0, 0]
    for v in range(g.nodes):
        lst = g.nlist[g.nindex[v]:g.nindex[v + 1]]
        if len(lst) == 0 or lst.min() >= v:
            continue
        d = len(lst)
        k = 0 if d <= 16 else (1 if d <= 2048 else 2)
        ent[k] += d
        lo, mid, hub = lo + (k == 0), mid + (k == 1), hub + (k == 2)
    assert b["low"] == (lo, ent[0]) and b["mid"] == (mid, ent[1]) and b["hub"] == (hub, ent[2])
    assert b["active"] == (lo + mid + hub, sum(ent))
    kb = roofline.kernel_bytes(g.nindex, g.nlist, 16, 2048)
    assert kb["compute2"] == 4 * ent[1] + 8 * mid
    assert roofline.pass_bytes(g.nodes, g.edges) == g.algorithmic_bytes()
    # a star with its hub last: one active high-degree vertex
    s = Graph.from_edges(3000, np.arange(2999), np.full(2999, 2999))
    b = roofline.bins(s.nindex, s.nlist, 16, 2048)
    assert b["hub"] == (1, 2999) and b["low"] == (0, 0)


def test_sampled_vertex_count_matches_the_rule():
    from eclcc import roofline

    g = workloads.rmat(11, 8, seed=5)
    for shift, min_deg in ((0, 0), (2, 0), (0, 16), (1, 5)):
        want = 0
        for v in range(g.nodes):
            lst = g.nlist[g.nindex[v]:g.nindex[v + 1]]
            if len(lst) == 0:
                continue
            if lst.min() < v:  # active: its second entry, if smaller (1 in 2^shift of
                # them, and only with at least min_deg entries)
                picked = shift == 0 or ((v * 0x9E3779B1) & 0xFFFFFFFF) >> (32 - shift) == 0
                want += int(len(lst) >= 2 and lst[1] < v and picked and len(lst) >= min_deg)
            else:              # inactive: its first entry
                want += int(lst[0] != v)
        assert roofline.bins(g.nindex, g.nlist, 16, 2048, shift, min_deg)["sampled"][0] == want
        assert roofline.kernel_bytes(g.nindex, g.nlist, 16, 2048, 2, shift, min_deg)["sample"] == 16 * want
    assert roofline.kernel_bytes(g.nindex, g.nlist, 16, 2048)["mark"] == 4 * g.nodes + g.nodes // 8


def test_host_allocation_failure_is_a_status_not_an_abort():
    # under a 1.5 GiB address-space limit the ER generator's 2 GiB pair
    # table cannot be allocated: the library returns ECLGRAPH_ERR_ALLOC
    # (no C++ exception crosses the C ABI, so the process does not abort)
    import sys

    from eclcc import _native as N

    code = r"""
import ctypes, resource, sys
resource.setrlimit(resource.RLIMIT_AS, (1536 << 20, 1536 << 20))
L = ctypes.CDLL(sys.argv[1])
L.eclgraph_gen_er.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64,
                              ctypes.POINTER(ctypes.c_void_p)]
L.eclgraph_last_error.restype = ctypes.c_char_p
out = ctypes.c_void_p()
rc = L.eclgraph_gen_er(1 << 20, 100_000_000, 1, ctypes.byref(out))
print(rc, bool(out.value), L.eclgraph_last_error().decode())
"""
    r = subprocess.run([sys.executable, "-c", code, str(N.LIB_DIR / "libeclgraph.so")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rc, made, msg = r.stdout.strip().split(" ", 2)
    assert int(rc) == -3 and made == "False" and "allocation" in msg, r.stdout
