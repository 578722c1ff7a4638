"""The drop-in boundary exercised from C: tests/c_caller/eclcc_caller.c
includes include/eclcc.h, links libeclcc.so, and calls
eclcc_compute(g.nodes, g.edges, g.nindex, g.nlist, nodestatus) where the
reference's main calls computeCC (src/racefree/egr-input/CC/ECL-CC-atomic.cu
:363).  Its labels must equal the committed golden labels and the oracle's."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import announce
from eclcc import Graph, workloads
from oracle import oracle

HERE = Path(__file__).resolve().parent
CALLER = HERE / "c_caller" / "eclcc_caller"
GOLDEN = sorted((HERE / "golden").glob("*.npz"))


def test_c_caller_is_built_against_the_library():
    assert CALLER.exists(), "tests/c_caller/eclcc_caller not built (__graft_entry__.build())"
    r = subprocess.run(["ldd", str(CALLER)], capture_output=True, text=True, timeout=60)
    assert "libeclcc.so" in r.stdout and "not found" not in r.stdout, r.stdout
    # the header compiles as C with every warning on (the Makefile uses -pedantic)
    assert "eclcc_compute" in (HERE.parent / "include" / "eclcc.h").read_text()
    r = subprocess.run([str(CALLER)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "USAGE:" in r.stderr


def run_caller(g, tmp_path, name):
    egr = tmp_path / f"{name}.egr"
    lab = tmp_path / f"{name}.labels"
    g.write(str(egr))
    r = subprocess.run([str(CALLER), str(egr), str(lab)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return np.fromfile(lab, dtype="<i4")


@pytest.mark.gpu
def test_c_caller_labels_equal_the_golden_labels(cuda, tmp_path):
    for f in GOLDEN:
        d = np.load(f)
        g = Graph(d["nindex"].astype(np.int32), d["nlist"].astype(np.int32))
        announce(f"C caller on {f.stem} n={g.nodes} m={g.edges}")
        got = run_caller(g, tmp_path, f.stem)
        assert np.array_equal(got, d["labels"]), f.stem
        assert np.array_equal(got, oracle.cc_labels(g.nindex, g.nlist)), f.stem


@pytest.mark.gpu
def test_c_caller_on_a_full_size_graph(cuda, tmp_path):
    g = workloads.grid256()
    announce(f"C caller on {g.name} n={g.nodes} m={g.edges}")
    got = run_caller(g, tmp_path, "grid256")
    assert np.array_equal(got, oracle.cc_labels(g.nindex, g.nlist))
