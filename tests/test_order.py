"""Block order of the batched walk kernels (order_item / block_order in
eclcc_internal.h), checked on the CPU: built with hipcc as host code
together with a checker that walks every block of many grid sizes and
window settings.  A missed or repeated work item would cost mark its cut of
the work lists without changing any label, so the GPU parity tests cannot
see it."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="needs hipcc")
def test_block_order_is_a_permutation_within_windows(tmp_path):
    exe = tmp_path / "block_order_check"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-x", "hip", "--cuda-host-only",
                    "-I", str(ROOT / "ecl-suite_amd" / "csrc"), str(ROOT / "tests" / "block_order_check.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all permutations" in r.stdout
