"""Writes the expected outputs of the MatrixMarket converter fixtures.

For every tests/golden/mtx/*.mtx this runs the reference's own converter
(scripts/mm2ecl_real.cpp, compiled where it lies by `make -C oracle ref` into
oracle/_ref/mm2ecl_real) and stores the .egr file it writes beside the input
as <name>.egr.  For every tests/golden/mtx/bad/*.mtx it records the error
line the reference prints (it exits with status -1 and writes nothing) in
bad/expected_errors.json.  Run from the repository root:

    make -C oracle ref && python tests/golden/make_mtx_golden.py
"""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent / "mtx"
REF = Path(__file__).resolve().parents[2] / "oracle" / "_ref" / "mm2ecl_real"


def main() -> None:
    if not REF.exists():
        sys.exit(f"{REF} is not built: run `make -C oracle ref` where the reference tree is present")
    for mtx in sorted(HERE.glob("*.mtx")):
        egr = mtx.with_suffix(".egr")
        r = subprocess.run([str(REF), str(mtx), str(egr)], capture_output=True, text=True)
        if r.returncode != 0:
            sys.exit(f"{mtx.name}: the reference converter failed: {r.stderr.strip()}")
        print(f"{mtx.name} -> {egr.name} ({egr.stat().st_size} bytes)")
    errors = {}
    for mtx in sorted((HERE / "bad").glob("*.mtx")):
        out = Path("/tmp") / (mtx.stem + ".egr")
        r = subprocess.run([str(REF), str(mtx), str(out)], capture_output=True, text=True)
        if r.returncode == 0:
            sys.exit(f"{mtx.name}: expected the reference converter to fail")
        errors[mtx.name] = r.stderr.strip().splitlines()[0]
        print(f"{mtx.name}: {errors[mtx.name]}")
    (HERE / "bad" / "expected_errors.json").write_text(json.dumps(errors, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
