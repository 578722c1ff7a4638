"""ECL-MST minimum spanning forest: the oracle pinned on the CPU, the
library's exports and host logic, and (marked gpu) the HIP path against the
oracle.  Reference: src/racefree/egr-input/MST/ECL-MST_10-atomic.cu."""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from conftest import announce
from eclcc import Graph, workloads
from eclcc import _native as N
from eclcc import mst
from oracle import oracle

INT32_MAX = 2**31 - 1


def from_pairs(n, pairs):
    a = np.array([p[0] for p in pairs], dtype=np.int32)
    b = np.array([p[1] for p in pairs], dtype=np.int32)
    return Graph.from_edges(n, a, b)


def key_weights(g, w):
    """{entry j: key} of the one-direction entries (nlist[j] > i)."""
    out = {}
    for i in range(g.nodes):
        for j in range(g.nindex[i], g.nindex[i + 1]):
            if g.nlist[j] > i:
                out[j] = (int(w[j]) << 32) | j
    return out


# ---- the oracle, pinned -----------------------------------------------------

def test_default_weights_hand_values():
    # 1 + ((i * k) % n), negated if negative, i * k in 32-bit two's complement
    g = from_pairs(70000, [(2, 3), (46341, 65536), (65536, 69999)])
    w = oracle.mst_weights(g.nindex, g.nlist)
    want = {(2, 3): 1 + 6, (3, 2): 1 + 6,
            # 46341 * 65536 = 3037003776 wraps to -1257963520; % 70000 = -63520
            (46341, 65536): 63519, (65536, 46341): 63519,
            # 65536 * 69999 = 4587454464 wraps to 292487168; % 70000 = 27168
            (65536, 69999): 27169, (69999, 65536): 27169}
    for i in range(g.nodes):
        for j in range(g.nindex[i], g.nindex[i + 1]):
            assert w[j] == want[(i, int(g.nlist[j]))]
    # the library computes the same weights
    assert np.array_equal(mst.mst_weights(g), w)


def test_oracle_weight_equals_scipy():
    import scipy.sparse as sp
    from scipy.sparse.csgraph import minimum_spanning_tree, connected_components

    rng = np.random.default_rng(3)
    for t in range(12):
        g = workloads.rmat(9, 4, t) if t % 2 else Graph.er(400, 1200, t)
        # scipy drops zero weights: keep them >= 1; many ties
        w = np.empty(g.edges, dtype=np.int32)
        for i in range(g.nodes):
            for j in range(g.nindex[i], g.nindex[i + 1]):
                k = g.nlist[j]
                w[j] = 1 + (min(i, k) * 7 + max(i, k) * 13 + t) % 5
        t_in = oracle.mst(g.nindex, g.nlist, w)
        A = sp.csr_matrix((w.astype(np.float64), g.nlist, g.nindex), shape=(g.nodes, g.nodes))
        assert int(minimum_spanning_tree(A).sum()) == int(w[t_in == 1].astype(np.int64).sum())
        ncomp = connected_components(A, directed=False)[0]
        assert int(t_in.sum()) == g.nodes - ncomp


def brute_force_forest(g, w):
    """Entries of the spanning forest with the least sum of keys, by trying
    every subset of the edges of a tiny graph."""
    keys = key_weights(g, w)
    ends = {}
    for i in range(g.nodes):
        for j in range(g.nindex[i], g.nindex[i + 1]):
            if g.nlist[j] > i:
                ends[j] = (i, int(g.nlist[j]))
    # components of the whole graph
    lab = oracle.cc_labels(g.nindex, g.nlist)
    need = g.nodes - len(set(lab.tolist()))
    best, best_sum = None, None
    for sub in itertools.combinations(sorted(ends), need):
        parent = list(range(g.nodes))

        def f(x):
            while parent[x] != x:
                x = parent[x]
            return x
        ok = True
        for j in sub:
            a, b = f(ends[j][0]), f(ends[j][1])
            if a == b:
                ok = False
                break
            parent[max(a, b)] = min(a, b)
        if ok:
            s = sum(keys[j] for j in sub)
            if best_sum is None or s < best_sum:
                best, best_sum = set(sub), s
    return best


def test_oracle_equals_brute_force_on_five_vertices():
    all_pairs = list(itertools.combinations(range(5), 2))
    rng = np.random.default_rng(11)
    for mask in range(1, 1 << len(all_pairs), 3):
        pairs = [p for b, p in enumerate(all_pairs) if mask >> b & 1]
        g = from_pairs(5, pairs)
        w = rng.integers(0, 3, g.edges).astype(np.int32)  # ties everywhere
        t_in = oracle.mst(g.nindex, g.nlist, w)
        assert set(np.nonzero(t_in)[0].tolist()) == brute_force_forest(g, w), pairs


def test_oracle_closed_forms():
    # a tree: every edge, from its smaller end only
    g = workloads.rmat(8, 1, 1)
    lab = oracle.cc_labels(g.nindex, g.nlist)
    tree_pairs = [(v, int(lab[v])) for v in range(g.nodes) if lab[v] != v]  # star per component
    t = from_pairs(g.nodes, tree_pairs)
    w = np.arange(t.edges, dtype=np.int32) % 7
    t_in = oracle.mst(t.nindex, t.nlist, w)
    for i in range(t.nodes):
        for j in range(t.nindex[i], t.nindex[i + 1]):
            assert t_in[j] == (t.nlist[j] > i)
    # a cycle: every edge but the one with the largest key
    n = 9
    c = from_pairs(n, [(v, (v + 1) % n) for v in range(n)])
    w = np.full(c.edges, 4, dtype=np.int32)
    t_in = oracle.mst(c.nindex, c.nlist, w)
    keys = key_weights(c, w)
    assert set(np.nonzero(t_in)[0].tolist()) == set(keys) - {max(keys, key=keys.get)}
    # no edges
    e = Graph(nindex=np.zeros(4, dtype=np.int32), nlist=np.zeros(0, dtype=np.int32))
    assert oracle.mst(e.nindex, e.nlist, np.zeros(0, dtype=np.int32)).size == 0


def test_threshold_host_logic():
    g = workloads.rmat(12, 8, 2)
    w = mst.mst_weights(g)
    thr = mst.threshold(g, w)
    s = np.sort(w[(np.arange(256, dtype=np.int64) * g.edges) // 256])
    assert thr == s[min(255, int(3.0 * g.nodes * 256 / g.edges))]
    sparse = workloads.grid3d(8, 0.6, 1)  # average degree < 4: one phase
    assert mst.threshold(sparse, mst.mst_weights(sparse)) == INT32_MAX


def test_mst_argument_errors_are_statuses():
    L = N.mst_lib()
    assert L.eclmst_compute(0, 0, None, None, None, None) == -1
    assert L.eclmst_workspace_bytes(0, 0) == 0
    assert L.eclmst_run_device(5, 0, None, None, None, None, None, 0, -1, None, None) == -1


# ---- the HIP path against the oracle ------------------------------------------

def _cases():
    yield "er", Graph.er(3000, 12000, 5), None
    yield "rmat", workloads.rmat(12, 8, 7), None
    yield "grid", workloads.grid3d(16, 0.6, 2), None
    yield "rmat-ties", workloads.rmat(11, 16, 9), "ties"
    yield "isolated", workloads.rmat(12, 1, 3), None
    yield "er-dense", Graph.er(1000, 40000, 6), "ties"


@pytest.mark.gpu
@pytest.mark.parametrize("name,g,wk", list(_cases()), ids=[c[0] for c in _cases()])
def test_gpu_mst_matches_oracle(cuda, name, g, wk):
    w = mst.mst_weights(g)
    if wk == "ties":
        w = (w % 4).astype(np.int32)
    announce(f"mst {name} n={g.nodes} m={g.edges}")
    want = oracle.mst(g.nindex, g.nlist, w)
    got = mst.compute_mst(g, w)
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_gpu_mst_device_thresholds_agree(cuda):
    g = workloads.rmat(14, 8, 4)
    w = mst.mst_weights(g)
    want = oracle.mst(g.nindex, g.nlist, w)
    dm = mst.DeviceMST(g, w)
    for thr in (-1, INT32_MAX, 0, int(np.median(w)), mst.threshold(g, w)):
        announce(f"mst device threshold {thr}")
        dm.run(thr)
        assert np.array_equal(dm.result(), want), thr


@pytest.mark.gpu
def test_gpu_mst_on_arrays_not_16_byte_aligned(cuda):
    # the list build loads entries and weights as 16-byte vectors only when
    # both arrays are 16-byte aligned; caller arrays at other 4-byte offsets
    # take the per-entry loads and give the same forest
    import ctypes as C

    import torch

    g = workloads.rmat(13, 8, 6)
    w = mst.mst_weights(g)
    want = oracle.mst(g.nindex, g.nlist, w)
    dev = torch.device("cuda:0")
    L = N.mst_lib()
    wb = int(L.eclmst_workspace_bytes(g.nodes, g.edges))
    work = torch.empty(wb, dtype=torch.uint8, device=dev)
    ni = torch.from_numpy(np.ascontiguousarray(g.nindex)).to(dev)
    for off_list, off_w in ((1, 0), (0, 3), (2, 1)):
        announce(f"mst with nlist at +{4 * off_list} B, weights at +{4 * off_w} B")
        nl = torch.zeros(g.edges + 4, dtype=torch.int32, device=dev)
        ww = torch.zeros(g.edges + 4, dtype=torch.int32, device=dev)
        nl[off_list:off_list + g.edges] = torch.from_numpy(g.nlist).to(dev)
        ww[off_w:off_w + g.edges] = torch.from_numpy(w).to(dev)
        out = torch.zeros(g.edges, dtype=torch.uint8, device=dev)
        ms = C.c_float(0.0)
        N.check_mst(L.eclmst_run_device(g.nodes, g.edges, ni.data_ptr(), nl[off_list:].data_ptr(),
                                        ww[off_w:].data_ptr(), out.data_ptr(), work.data_ptr(), wb, -1,
                                        torch.cuda.current_stream().cuda_stream, C.byref(ms)))
        assert np.array_equal(out.cpu().numpy(), want), (off_list, off_w)


@pytest.mark.gpu
def test_gpu_mst_medium_rmat(cuda):
    g = workloads.rmat(18, 16, 5)
    w = mst.mst_weights(g)
    announce(f"mst rmat18 n={g.nodes} m={g.edges}")
    want = oracle.mst(g.nindex, g.nlist, w)
    got, sec = mst.compute_mst(g, w, timed=True)
    assert np.array_equal(got, want)
    assert sec > 0


@pytest.mark.gpu
def test_gpu_mst_rejects_bad_inputs(cuda):
    g = workloads.rmat(10, 8, 1)
    w = mst.mst_weights(g)
    w[5] = -3
    with pytest.raises(N.EclmstError) as e:
        mst.compute_mst(g, w)
    assert e.value.status == -6
    # one direction only (not symmetric): more list entries than the lists hold
    n = 2000
    nindex = np.arange(0, 3 * n + 1, 3, dtype=np.int32)
    nlist = np.array([min(i + d, n - 1) if i + d < n else i for i in range(n) for d in (1, 2, 3)],
                     dtype=np.int32)
    gd = Graph(nindex=nindex, nlist=nlist)
    with pytest.raises(N.EclmstError) as e:
        mst.compute_mst(gd, np.ones(nlist.shape[0], dtype=np.int32))
    assert e.value.status == -6


def test_mst_cli_usage_and_file_errors(tmp_path):
    import subprocess

    exe = N.BIN_DIR / "eclmst"
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 255 and "USAGE:" in r.stderr
    r = subprocess.run([str(exe), str(tmp_path / "none.egr")], capture_output=True, text=True)
    assert r.returncode == 255 and "ERROR:" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("weighted", [False, True])
def test_gpu_mst_cli_matches_the_harness_protocol(cuda, tmp_path, weighted):
    # the reference's main prints "Dev+/-: %12.9f s" and "all good"; its
    # analyzer turns Dev[-+] into "runtime" and reads `time: +(\d+\.\d+) s`
    # (scripts/analyze_results.py:62,83-84)
    import re
    import subprocess

    g = workloads.rmat(12, 8, 6) if weighted else Graph.er(2000, 3000, 4)
    if weighted:
        g.eweight = (mst.mst_weights(g) % 50).astype(np.int32)
    path = tmp_path / "g.egr"
    g.write(str(path))
    announce(f"eclmst CLI on {g.nodes} nodes, weighted file: {weighted}")
    r = subprocess.run([str(N.BIN_DIR / "eclmst"), str(path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "all good" in r.stdout and "differ" not in r.stdout
    assert ("Dev+:" if g.edges // g.nodes >= 4 else "Dev-:") in r.stdout
    line = next(l for l in r.stdout.splitlines() if l.startswith("Dev"))
    assert re.search(r"time: +(\d+\.\d+) s", re.sub(r"Dev[-+]", "runtime", line))


def test_mst_shared_accesses_go_through_atomic_helpers():
    """parent[] and minv[] are touched through the relaxed agent-scope
    helpers everywhere two threads of one kernel may meet on them; the plain
    accesses left are in kernels that only write the array at distinct
    addresses (k_init, k_reset) or only read it (k_pick, k_members)."""
    import re

    src = (N.PKG_ROOT / "csrc" / "eclmst.hip").read_text()
    body = re.sub(r"//[^\n]*", "", src)
    helpers = r"(?:ld|st|cas|ld64|st64|min64)\(&$"
    plain = []
    for m in re.finditer(r"\b(parent|minv)\[", body):
        if re.search(helpers, body[max(0, m.start() - 10):m.start()]):
            continue
        k = body.rfind("__global__", 0, m.start())
        kernel = re.search(r"void (\w+)\(", body[k:k + 300]).group(1)
        plain.append((kernel, m.group(1)))
    assert sorted(set(plain)) == [("k_init", "minv"), ("k_init", "parent"), ("k_members", "parent"),
                                  ("k_pick", "parent"), ("k_reset", "minv")], plain
    for kernel in ("k_init", "k_reset", "k_pick", "k_members"):
        k = body.find(f"void {kernel}(")
        kbody = body[k:body.find("\n}\n", k)]
        assert not re.search(r"\b(ld|st|cas|ld64|st64|min64)\(&(parent|minv)", kbody), kernel
    prod = src.split("// ---- atomics", 1)[1].split("// Root of x", 1)[0]
    for f in ("__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)",
              "__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)",
              "__hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)"):
        assert f in prod, f


def _forest_of_multigraph(g):
    """The forest of a multigraph whose simple graph has no cycle and whose
    lists are sorted where they hold an entry to a larger neighbour (the
    builder in graphs.max_entries_multigraph): every distinct edge i < k is
    in it, at the first copy in i's list (all copies weigh the same, and
    ties go to the smaller entry index).  Returns the entry indices."""
    pos = []
    for i in range(g.nodes):
        b, e = int(g.nindex[i]), int(g.nindex[i + 1])
        seg = g.nlist[b:e]
        if e == b or int(seg[-1]) <= i:
            continue
        for k in np.unique(seg[[0, np.searchsorted(seg, i + 1), e - b - 1]]):
            if k > i:
                pos.append(b + int(np.searchsorted(seg, k)))
    return np.asarray(sorted(set(pos)), dtype=np.int64)


def test_forest_of_multigraph_helper_equals_the_oracle():
    import graphs as G

    g, _ = G.max_entries_multigraph(chain=12, split=5, tail_loops=11, x=7, edges=10001)
    w = mst.mst_weights(g)
    ref = oracle.mst(g.nindex, g.nlist, w)
    assert np.array_equal(np.flatnonzero(ref), _forest_of_multigraph(g))


@pytest.mark.gpu
def test_gpu_mst_at_the_largest_int32_edge_count(cuda):
    # 2^31 - 1 entries (8 GiB of nlist and 8 GiB of weights) with the
    # reference's default weights and threshold: the forest is every edge
    # of the two paths and the extra vertex's edge, each at its first copy
    import graphs as G

    g, _ = G.max_entries_multigraph()
    w = mst.mst_weights(g)
    want = _forest_of_multigraph(g)
    announce(f"eclmst_compute {g.name} n={g.nodes} threshold {mst.threshold(g, w)}")
    got = mst.compute_mst(g, w)
    assert int(got.sum(dtype=np.int64)) == want.size == g.nodes - 3  # three components
    assert bool(got[want].all())


@pytest.mark.gpu
def test_gpu_mst_on_irregular_graphs(cuda):
    # unsorted lists, duplicate entries, self-loops, stars, paths, cliques;
    # the default weights and random weights from a small range (many ties,
    # broken by the entry index): the forest equals the oracle's entry for
    # entry
    import graphs as G

    rng = np.random.default_rng(11)
    for seed in range(24):
        g = G.random_structure(seed)
        for name, w in (("default", mst.mst_weights(g)), ("ties", None)):
            # a duplicate entry must weigh the same in both directions
            src = np.repeat(np.arange(g.nodes), np.diff(g.nindex))
            lo = np.minimum(src, g.nlist).astype(np.int64)
            hi = np.maximum(src, g.nlist).astype(np.int64)
            if name == "ties":
                key = lo * g.nodes + hi
                _, inv = np.unique(key, return_inverse=True)
                w = rng.integers(0, 4, inv.max() + 1 if inv.size else 1).astype(np.int32)[inv]
            announce(f"eclmst_compute {g.name} weights {name}")
            want = oracle.mst(g.nindex, g.nlist, w)
            got = mst.compute_mst(g, w)
            assert np.array_equal(got[:g.edges], want[:g.edges]), (g.name, name)
