"""The server's arena allocator (ecl-suite_amd/csrc/eclcc_arena.h) on the
CPU: built with g++ together with a randomised checker and run."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_arena_blocks_never_overlap_and_always_coalesce(tmp_path):
    exe = tmp_path / "arena_fuzz"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "ecl-suite_amd" / "csrc"),
                    "-o", str(exe), str(ROOT / "tests" / "arena_fuzz.cpp")], check=True, timeout=120)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "arena ok", r.stdout + r.stderr
