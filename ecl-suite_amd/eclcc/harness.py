"""Harness-compatible runner: the reference's experiment protocol for the
MI355X executables (bin/eclcc, bin/eclmst).

The reference's harness compiles each .cu file, runs it `num_runs` times per
input and prints progress lines around the program's own output
(scripts/multi_run.py:65-81, :109-110); all_tests.sh:23-36 redirects that to
results/<kind>_<code>.txt, and scripts/analyze_results.py:59-110 parses it:
`num_inputs: N`, the input name from an `inputs-undirected/<name>.egr` or
`inputs-directed/<name>.egr` path, and every `time: <float> s` line (MST's
`Dev+:` / `Dev-:` lines count as times too) -- then checks that there are N
inputs with the same number of times each and takes the median per input.

Here the executable is already built, so the compile step is gone; the lines
the analyser reads are the same, so `analyze_results.py` reads this runner's
output unchanged, and `parse_results` / `median_times` below restate its
parse and median rules for our own tables.

    python -m eclcc.harness run [--runs 9] [--exe bin/eclcc] <input_dir>  > results/racefree_CC.txt
    python -m eclcc.harness median results/racefree_CC.txt
    python -m eclcc.harness speedup results/baseline_CC.txt results/racefree_CC.txt
"""
from __future__ import annotations

import argparse
import os
import re
import statistics
import subprocess
import sys
from typing import Dict, List, Optional, TextIO

from . import _native as N

INPUT_RE = re.compile(r"(?:inputs-undirected|inputs-directed)/(\S+)\.egr")
TIME_RE = re.compile(r"time: +(\d+\.\d+) s")
NUM_INPUTS_RE = re.compile(r"num_inputs:\s*(\d+)")
MST_DEV_RE = re.compile(r"Dev[-+]")


def _emit(out: TextIO, text: str) -> None:
    out.write(text)
    out.flush()


def run(input_dir: str, exe: Optional[str] = None, num_runs: int = 9,
        out: TextIO = sys.stdout) -> int:
    """Runs `exe` (default bin/eclcc) num_runs times on every file of
    input_dir, in directory order, writing the harness's progress lines and
    the program's own output to `out` (multi_run.py:71-81, :104-110).
    Returns the number of runs whose exit status was not 0 (the reference
    does not check; the analyser then finds a missing time)."""
    exe = str(exe or (N.BIN_DIR / "eclcc"))
    inputs = [f for f in os.listdir(input_dir) if os.path.isfile(os.path.join(input_dir, f))]
    _emit(out, f"num_codes: 1\nnum_inputs: {len(inputs)}\n\n")
    _emit(out, f"code_path: {os.path.dirname(exe)}\ninput_path: {input_dir}\n\n")
    failed = 0
    code = os.path.basename(exe)
    for k, name in enumerate(inputs, 1):
        _emit(out, f"running {code}, {k} out of {len(inputs)} inputs\n\n")
        for i in range(num_runs):
            _emit(out, f"run {i + 1} out of {num_runs}\n\n")
            r = subprocess.run([exe, os.path.join(input_dir, name)], stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True)
            _emit(out, r.stdout)
            failed += r.returncode != 0
    return failed


def parse_results(lines) -> Dict[str, List[float]]:
    """The times of each input in a results file (analyze_results.py:59-103):
    raises ValueError when `num_inputs` is missing, when the number of inputs
    found differs from it, or when the inputs do not all have the same
    number of times."""
    results: Dict[str, List[float]] = {}
    current = None
    num_inputs = None
    for line in lines:
        m = NUM_INPUTS_RE.search(line)
        if m:
            num_inputs = int(m.group(1))
        m = INPUT_RE.search(line)
        if m:
            current = m.group(1)
            results.setdefault(current, [])
        m = TIME_RE.search(MST_DEV_RE.sub("runtime", line))
        if m and current:
            results[current].append(float(m.group(1)))
    if not num_inputs:
        raise ValueError("num_inputs not found")
    if len(results) != num_inputs:
        raise ValueError(f"Expected {num_inputs} inputs, but found {len(results)}.")
    # the first input's count is the one every other input must have
    expect = None
    for name, times in results.items():
        if expect is None:
            expect = len(times)
        elif len(times) != expect:
            raise ValueError(f"Expected {expect} runtimes for input '{name}', but found {len(times)}.")
    return results


def median_times(results: Dict[str, List[float]]) -> Dict[str, float]:
    """Median time per input (analyze_results.py:105-110)."""
    return {k: statistics.median(v) for k, v in results.items() if v}


def speedups(old: Dict[str, float], new: Dict[str, float]) -> Dict[str, float]:
    """old / new per input present in both (analyze_results.py:112-124)."""
    return {k: old[k] / new[k] for k in old if k in new and new[k] != 0}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m eclcc.harness")
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("input_dir")
    r.add_argument("--exe", default=None)
    r.add_argument("--runs", type=int, default=9)
    m = sub.add_parser("median")
    m.add_argument("results")
    s = sub.add_parser("speedup")
    s.add_argument("old")
    s.add_argument("new")
    a = ap.parse_args(argv)
    if a.cmd == "run":
        return 1 if run(a.input_dir, a.exe, a.runs) else 0
    if a.cmd == "median":
        with open(a.results) as f:
            for k, v in sorted(median_times(parse_results(f)).items()):
                print(f"{k}\t{v:.9f}")
        return 0
    with open(a.old) as f:
        old = median_times(parse_results(f))
    with open(a.new) as f:
        new = median_times(parse_results(f))
    for k, v in sorted(speedups(old, new).items()):
        print(f"{k}\t{v:.4f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
