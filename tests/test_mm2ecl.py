"""The MatrixMarket converter against the reference's own
(scripts/mm2ecl_real.cpp:50-184): byte-identical .egr files on the committed
fixtures (expected bytes written by the reference's converter, compiled from
its source into oracle/_ref, by tests/golden/make_mtx_golden.py), the same
refusals with the same messages, and -- where oracle/_ref/mm2ecl_real is
built -- byte-identical output on randomised files, run live."""
from __future__ import annotations

import json
import random
import subprocess
from pathlib import Path

import numpy as np
import pytest

from eclcc import EclgraphError, Graph
from eclcc import _native as N
from eclcc.graph import DEDUPE, DROP_SELF, SYMMETRIZE
from oracle import oracle

MTX = Path(__file__).resolve().parent / "golden" / "mtx"
REF = oracle.HERE / "_ref" / "mm2ecl_real"
GOOD = sorted(MTX.glob("*.mtx"))
BAD = sorted((MTX / "bad").glob("*.mtx"))


def convert(mtx: Path, out: Path) -> str:
    g, kind = Graph.mm2ecl(str(mtx))
    g.write(str(out))
    return kind


@pytest.mark.parametrize("mtx", GOOD, ids=[p.stem for p in GOOD])
def test_converter_writes_the_reference_bytes(tmp_path, mtx):
    out = tmp_path / "ours.egr"
    convert(mtx, out)
    assert out.read_bytes() == mtx.with_suffix(".egr").read_bytes()


@pytest.mark.parametrize("mtx", GOOD, ids=[p.stem for p in GOOD])
def test_cli_prints_and_writes_like_the_reference(tmp_path, mtx):
    out = tmp_path / "cli.egr"
    r = subprocess.run([str(N.BIN_DIR / "mm2ecl"), str(mtx), str(out)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == mtx.with_suffix(".egr").read_bytes()
    lines = [ln for ln in r.stdout.splitlines() if "\t#" in ln]
    raw = np.fromfile(mtx.with_suffix(".egr"), dtype="<i4")
    n, m = int(raw[0]), int(raw[1])
    kind = {0: "no", 1: "int", 2: "float"}
    has_w = raw.size == 2 + n + 1 + 2 * m and m > 0
    head = mtx.read_text().splitlines()[0]
    expect_kind = "int" if "integer" in head else "float" if "real" in head else "no"
    assert lines == [f"{mtx}\t#name", f"{n}\t#nodes", f"{m}\t#edges", f"{expect_kind}\t#weights"]
    assert has_w == (expect_kind != "no" and m > 0) and expect_kind in kind.values()
    if REF.exists():
        r2 = subprocess.run([str(REF), str(mtx), str(tmp_path / "ref.egr")], capture_output=True,
                            text=True, timeout=60)
        assert [ln for ln in r2.stdout.splitlines() if "\t#" in ln] == lines


@pytest.mark.parametrize("mtx", BAD, ids=[p.stem for p in BAD])
def test_converter_refuses_what_the_reference_refuses(tmp_path, mtx):
    expected = json.loads((MTX / "bad" / "expected_errors.json").read_text())[mtx.name]
    with pytest.raises(EclgraphError) as e:
        Graph.mm2ecl(str(mtx))
    msg = str(e.value).split(": ", 1)[1]
    assert expected == f"ERROR: {msg}"
    out = tmp_path / "cli.egr"
    r = subprocess.run([str(N.BIN_DIR / "mm2ecl"), str(mtx), str(out)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 255 and r.stderr.splitlines()[0] == expected and not out.exists()


def test_weights_are_scaled_and_truncated_like_c():
    # (int)(w * 100) in double arithmetic: 0.29 * 100 = 28.999999999999996
    g, kind = Graph.mm2ecl(str(MTX / "real_general.mtx"))
    assert kind == "float"
    src = np.repeat(np.arange(g.nodes), np.diff(g.nindex))
    w = {(int(a), int(b)): [] for a, b in zip(src, g.nlist)}
    for a, b, x in zip(src, g.nlist, g.eweight):
        w[(int(a), int(b))].append(int(x))
    assert w[(0, 1)] == [28, 28]          # 0.28 and 0.29, ordered by the real value
    assert w[(0, 2)] == [-150]            # -1.5
    assert w[(2, 0)] == [0]               # 1e-3
    assert w[(1, 3)] == [267]             # 2.675
    assert w[(3, 1)] == [12345]           # 123.456
    assert w[(2, 4)] == [0]               # -0.009 truncates toward zero
    assert w[(4, 3)] == [114]             # 1.15


def test_symmetric_files_are_not_mirrored_and_flags_make_them_cc_ready():
    g, kind = Graph.mm2ecl(str(MTX / "pattern_symmetric.mtx"))
    assert kind == "no" and g.edges == 5  # the lower triangle as given
    assert g.nlist.tolist() == [0, 0, 2, 3, 4] and g.nindex.tolist() == [0, 0, 1, 2, 3, 5]
    h = Graph.read_mtx(str(MTX / "pattern_symmetric.mtx"), SYMMETRIZE | DROP_SELF | DEDUPE)
    h.check()
    assert oracle.cc_labels(h.nindex, h.nlist).tolist() == [0, 0, 0, 0, 0]


def _random_mtx(rng: random.Random, path: Path) -> None:
    """A valid coordinate file in one of the forms the reference reads:
    pattern / integer / real, general / symmetric, entries unsorted with
    duplicates and self-loops, several entries per line or blank lines
    between them, comment lines before the size line."""
    field = rng.choice(["pattern", "integer", "real"])
    sym = rng.choice(["general", "symmetric"])
    n = rng.randint(1, 40)
    m = rng.randint(0, 3 * n)
    out = [f"%%MatrixMarket matrix coordinate {field} {sym}\n"]
    for _ in range(rng.randint(0, 3)):
        out.append(rng.choice(["% comment\n", "  % indented\n", "x % percent inside\n", "%\n"]))
    out.append(f"{n} {n} {m}\n")
    for k in range(m):
        i, j = rng.randint(1, n), rng.randint(1, n)
        if field == "pattern":
            tok = f"{i} {j}"
        elif field == "integer":
            tok = f"{i} {j} {rng.randint(-10**6, 10**6)}"
        else:
            x = rng.uniform(-1000, 1000)
            form = rng.choice(["{:.2f}", "{:.3f}", "{:.6g}", "{:e}", "{:.1f}", "{:.0f}"])
            tok = f"{i} {j} " + form.format(x)
        sep = rng.choice(["\n", "\n", " ", "\n\n", "\t"])
        out.append(tok + (sep if k + 1 < m else "\n"))
    path.write_text("".join(out))


@pytest.mark.skipif(not REF.exists(), reason="reference converter not built (make -C oracle ref)")
def test_random_files_convert_to_the_reference_bytes(tmp_path):
    rng = random.Random(20241204)
    for t in range(300):
        mtx = tmp_path / f"r{t}.mtx"
        _random_mtx(rng, mtx)
        ref_out = tmp_path / f"r{t}.ref.egr"
        r = subprocess.run([str(REF), str(mtx), str(ref_out)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, (mtx.read_text(), r.stderr)
        ours = tmp_path / f"r{t}.egr"
        convert(mtx, ours)
        assert ours.read_bytes() == ref_out.read_bytes(), mtx.read_text()
