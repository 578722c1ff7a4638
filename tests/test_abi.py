"""The C ABI boundary without a GPU: both libraries load and export every
function include/*.h declares; argument errors come back as statuses; the
CLI keeps the reference's usage/error behaviour; the kernels' label accesses
all go through the relaxed-atomic helpers."""
from __future__ import annotations

import ctypes as C
import re
import subprocess

import numpy as np
import pytest

import eclcc
from eclcc import _native as N


@pytest.mark.parametrize("header,lib", [("eclcc.h", "libeclcc.so"), ("eclgraph.h", "libeclgraph.so"),
                                        ("eclmst.h", "libeclmst.so")])
def test_library_exports_every_declared_function(header, lib):
    names = N.header_functions(header)
    assert len(names) >= 7
    dll = C.CDLL(str(N.LIB_DIR / lib))
    missing = [f for f in names if not hasattr(dll, f)]
    assert not missing, f"{lib} lacks {missing}"


def test_params_defaults_and_workspace():
    p = eclcc.default_params()
    assert p.thread_max_deg >= 0 and p.wave_max_deg > p.thread_max_deg
    assert 64 <= p.hub_chunk <= 8192 and p.block_threads in (256, 512, 1024)
    L = N.cc_lib()
    small = L.eclcc_workspace_bytes(1000, 10000, C.byref(p))
    big = L.eclcc_workspace_bytes(100000, 10000000, C.byref(p))
    assert 0 < small < big
    assert L.eclcc_workspace_bytes(0, 10, C.byref(p)) == 0
    bad = eclcc.default_params(block_threads=300)
    assert L.eclcc_workspace_bytes(1000, 10000, C.byref(bad)) == 0
    assert p.giant_filter == 2 and p.path_mode == 1
    for f in (0, 1, 2):
        assert L.eclcc_workspace_bytes(1000, 10000, C.byref(eclcc.default_params(giant_filter=f))) > 0
    assert p.sample_shift == 0 and p.sample_min_deg == 0
    assert L.eclcc_workspace_bytes(1000, 10000, C.byref(eclcc.default_params(sample_min_deg=1 << 30))) > 0
    for sh in range(5):
        assert L.eclcc_workspace_bytes(1000, 10000, C.byref(eclcc.default_params(sample_shift=sh))) > 0
    for bad in (eclcc.default_params(giant_filter=3), eclcc.default_params(giant_filter=-1),
                eclcc.default_params(path_mode=3), eclcc.default_params(sample_shift=5),
                eclcc.default_params(sample_shift=-1), eclcc.default_params(sample_min_deg=-1)):
        assert L.eclcc_workspace_bytes(1000, 10000, C.byref(bad)) == 0


def test_argument_errors_are_statuses():
    L = N.cc_lib()
    rc = L.eclcc_run_device(0, 0, None, None, None, None, 0, None, None)
    assert rc == -1 and b"too low" in L.eclcc_last_error()
    rc = L.eclcc_compute(5, -1, None, None, None)
    assert rc == -1
    with pytest.raises(eclcc.EclccError):
        eclcc.compute_cc(0, 0, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    with pytest.raises(ValueError):
        eclcc.compute_cc(3, 0, np.zeros(3, np.int32), np.zeros(0, np.int32), np.zeros(3, np.int32))
    assert L.eclcc_status_string(-6) == b"request exceeds the HBM budget"
    assert b"gfx950" in L.eclcc_version()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(N, "LIB_DIR", tmp_path)
    monkeypatch.setattr(N, "_cc", None)
    with pytest.raises(N.NativeLibraryMissing):
        N.cc_lib()


def test_cli_usage_and_file_errors(tmp_path):
    exe = N.BIN_DIR / "eclcc"
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 255 and "USAGE:" in r.stderr
    r = subprocess.run([str(exe), str(tmp_path / "none.egr")], capture_output=True, text=True)
    assert r.returncode == 255 and "ERROR: could not open file" in r.stderr


def test_label_accesses_go_through_atomic_helpers():
    src = (N.PKG_ROOT / "csrc" / "eclcc_kernels.hip").read_text()
    body = re.sub(r"//[^\n]*", "", src)
    uses = [m for m in re.finditer(r"\b(ns|nstat)\[", body)]
    assert uses
    plain = []
    for m in uses:
        before = body[max(0, m.start() - 40):m.start()]
        # ns[...], G.nstat[...], or (k < H ? G0 : G1).nstat[...] in the walks
        # that take two chunks per wavefront
        if re.search(r"lab_(ld|st|cas)\(&(?:G\.|\(k < H \? G0 : G1\)\.)?$", before):
            continue
        line = body[body.rfind("\n", 0, m.start()) + 1:body.find("\n", m.end())]
        plain.append(line.strip())
    # the product's helpers are relaxed agent-scope atomics (the plain ones
    # exist only under ECLCC_PLAIN_LABELS, for the research copy)
    prod = body.split("#ifndef ECLCC_PLAIN_LABELS", 1)[1].split("#else", 1)[0]
    assert "__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)" in prod
    assert "__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)" in prod
    assert "plain labels" not in N.cc_lib().eclcc_version().decode()
    # the one plain store is init's, where no other thread reads labels
    assert plain == ["if (live) G.nstat[v] = m;"], plain


def test_params_for_graph_picks_by_degree_skew():
    # graphs of 2^20 vertices or more: even degrees (grid, Erdos-Renyi) take
    # half of the sampled hooks, and from 2^23 vertices the halving walks
    # too; a heavy tail (R-MAT) takes the halving walks and spends its
    # second-entry hooks on lists of 16 entries or more; smaller graphs and
    # the rest of the parameters keep the defaults
    from eclcc import Graph, params_for_graph, workloads

    d = eclcc.default_params().as_dict()
    for g, shift, mode, mind in ((workloads.grid3d(104, 0.6, 1), 1, 1, 0), (Graph.er(1 << 20, 4 << 20, 3), 1, 1, 0),
                                 (workloads.rmat(20, 16, 2), 0, 2, 16), (workloads.rmat(16, 16, 2), 0, 1, 0),
                                 (workloads.er65536(1), 0, 1, 0), (workloads.grid3d(32, 0.6, 1), 0, 1, 0)):
        p = params_for_graph(g).as_dict()
        assert p == dict(d, sample_shift=shift, path_mode=mode, sample_min_deg=mind), g.name
    # a large star is skewed; a large empty graph is not; small ones keep the defaults
    import graphs as G

    # (the hub at a sampled position: the choice reads 4096 evenly spaced degrees)
    star = Graph.from_edges(1 << 20, np.zeros((1 << 20) - 1, dtype=np.int32), np.arange(1, 1 << 20))
    assert params_for_graph(star).as_dict() == dict(d, path_mode=2, sample_min_deg=16)
    assert params_for_graph(G.empty(1 << 20)).as_dict() == dict(d, sample_shift=1)
    assert params_for_graph(G.empty((1 << 23) - 1)).as_dict() == dict(d, sample_shift=1)
    assert params_for_graph(G.empty(1 << 23)).as_dict() == dict(d, sample_shift=1, path_mode=2)
    assert params_for_graph(G.empty(1)).as_dict() == d
    assert params_for_graph(workloads.grid3d(16, 0.6, 1), giant_filter=1).giant_filter == 1
    L = N.cc_lib()
    p = N.EclccParams()
    assert L.eclcc_params_for_graph(0, 0, None, C.byref(p)) == -1
    assert L.eclcc_params_for_graph(1, 0, None, None) == -1
