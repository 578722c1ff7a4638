"""The harness-compatible runner (eclcc.harness) and its restatement of the
reference analyser's parse and median rules
(scripts/analyze_results.py:59-124, scripts/multi_run.py:65-81)."""
from __future__ import annotations

import io
import os
import stat
import statistics
import subprocess
import sys

import numpy as np
import pytest

from conftest import announce
from eclcc import harness
from eclcc import _native as N


def results_file(times_by_input, num_inputs=None, mst=False):
    """A results file as multi_run.py and the CLI print it: the progress
    lines, then per run the program's banner, its input line and its time
    line (ECL-CC-atomic.cu:340-341, :314; MST's `Dev+:` line, :364)."""
    names = list(times_by_input)
    out = [f"num_codes: 1\nnum_inputs: {len(names) if num_inputs is None else num_inputs}\n\n",
           "code_path: src/racefree/egr-input/CC\ninput_path: inputs-undirected/\n\n"]
    for k, name in enumerate(names, 1):
        out.append(f"running ECL-CC-atomic.cu, {k} out of {len(names)} inputs\n\n")
        for i, t in enumerate(times_by_input[name], 1):
            out.append(f"run {i} out of {len(times_by_input[name])}\n\n")
            if mst:
                out.append(f"input: inputs-undirected/{name}.egr\nDev+: {t:12.9f} s\nHost:  0.123456789 s\nall good\n")
            else:
                out.append(f"ECL-CC v1.1 (race-free)\ninput graph: 10 nodes and 20 edges "
                           f"(inputs-undirected/{name}.egr)\ncompute time: {t:.9f} s\n"
                           f"throughput: 1.000 Mnodes/s\nverification passed\n\n")
    return "".join(out).splitlines(keepends=True)


def test_nine_runs_of_two_inputs_give_the_median():
    rng = np.random.default_rng(3)
    times = {"rmat16.sym": [float(x) for x in np.round(rng.uniform(1e-4, 2e-3, 9), 9)],
             "USA-road-d.NY": [float(x) for x in np.round(rng.uniform(1e-3, 9e-3, 9), 9)]}
    parsed = harness.parse_results(results_file(times))
    assert parsed == times
    med = harness.median_times(parsed)
    assert med == {k: statistics.median(v) for k, v in times.items()}
    # the median of 9 is the 5th smallest
    assert med["rmat16.sym"] == sorted(times["rmat16.sym"])[4]
    sp = harness.speedups({"rmat16.sym": 2.0, "USA-road-d.NY": 3.0, "gone": 1.0},
                          {"rmat16.sym": 1.0, "USA-road-d.NY": 1.5})
    assert sp == {"rmat16.sym": 2.0, "USA-road-d.NY": 2.0}


def test_mst_device_lines_count_as_times():
    times = {"a": [0.5, 0.25, 0.75], "b": [1.0, 2.0, 3.0]}
    parsed = harness.parse_results(results_file(times, mst=True))
    assert parsed == times  # the Host: line is not a time


def test_parse_rules_reject_incomplete_files():
    ok = {"a": [0.1, 0.2], "b": [0.3, 0.4]}
    with pytest.raises(ValueError, match="num_inputs not found"):
        harness.parse_results([ln for ln in results_file(ok) if "num_inputs" not in ln])
    with pytest.raises(ValueError, match="Expected 3 inputs, but found 2"):
        harness.parse_results(results_file(ok, num_inputs=3))
    with pytest.raises(ValueError, match="Expected 2 runtimes for input 'b', but found 1"):
        harness.parse_results(results_file({"a": [0.1, 0.2], "b": [0.3]}))


def test_runner_prints_the_harness_protocol(tmp_path):
    # a stand-in program in place of bin/eclcc, so this runs without a GPU
    exe = tmp_path / "prog"
    exe.write_text("#!/bin/sh\necho \"input graph: 4 nodes and 6 edges ($1)\"\necho \"compute time: 0.000250000 s\"\n")
    exe.chmod(exe.stat().st_mode | stat.S_IEXEC)
    d = tmp_path / "inputs-undirected"
    d.mkdir()
    for name in ("g1.egr", "g2.egr"):
        (d / name).write_bytes(b"")
    buf = io.StringIO()
    assert harness.run(str(d), str(exe), 9, buf) == 0
    text = buf.getvalue()
    assert "num_inputs: 2" in text and text.count("run 9 out of 9") == 2
    parsed = harness.parse_results(text.splitlines(keepends=True))
    assert sorted(parsed) == ["g1", "g2"] and all(v == [0.00025] * 9 for v in parsed.values())


@pytest.mark.gpu
def test_runner_drives_the_cc_cli(cuda, tmp_path):
    from eclcc import workloads

    d = tmp_path / "inputs-undirected"
    d.mkdir()
    workloads.rmat(14, 16, 5).write(str(d / "rmat14.egr"))
    workloads.grid3d(24, 0.6, 2).write(str(d / "grid24.egr"))
    announce("harness runner: bin/eclcc, 2 inputs x 3 runs")
    out = tmp_path / "racefree_CC.txt"
    with open(out, "w") as f:
        assert harness.run(str(d), str(N.BIN_DIR / "eclcc"), 3, f) == 0
    text = out.read_text()
    assert text.count("verification passed") == 6
    med = harness.median_times(harness.parse_results(text.splitlines(keepends=True)))
    assert sorted(med) == ["grid24", "rmat14"] and all(0 < t < 1 for t in med.values())
    r = subprocess.run([sys.executable, "-m", "eclcc.harness", "median", str(out)], capture_output=True,
                       text=True, cwd=str(N.PKG_ROOT), timeout=60)
    assert r.returncode == 0 and len(r.stdout.splitlines()) == 2
