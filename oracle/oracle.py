"""ctypes face of liboracle.so -- the CPU oracle of the ECL-CC and ECL-MST paths.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker (or as the CPU baseline
timed beside the GPU).  The product path (ecl-suite_amd/eclcc, libeclcc.so)
never imports it.  What each function restates and how it is pinned is in
oracle.h.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_ROUNDTRIP = HERE / "_ref" / "egr_roundtrip"

VERIFY_CODES = {
    0: "ok",
    -1: "found adjacent nodes in different components",
    -2: "found negative component number",
    -3: "found incorrect ID value",
    -4: "number of components do not match",
    -5: "component IDs are not unique",
    -6: "label out of range",
    -7: "label is not the component minimum",
    -8: "allocation failed",
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} not built; run `make -C {HERE}`")
        L = C.CDLL(str(LIB))
        vp, i32 = C.c_void_p, C.c_int32
        for f in ("oracle_cc_unionfind", "oracle_cc_bfs", "oracle_eclcc_serial"):
            getattr(L, f).argtypes = [i32, vp, vp, vp]
            getattr(L, f).restype = C.c_int
        L.oracle_init.argtypes = [i32, vp, vp, vp]
        L.oracle_init.restype = None
        L.oracle_verify.argtypes = [i32, vp, vp, vp, C.POINTER(C.c_int64)]
        L.oracle_verify.restype = C.c_int
        L.oracle_eclcc_threads.argtypes = [i32, vp, vp, vp, C.c_int]
        L.oracle_eclcc_threads.restype = C.c_int
        L.oracle_count_components.argtypes = [i32, vp]
        L.oracle_count_components.restype = C.c_int64
        L.oracle_mst_weights.argtypes = [i32, vp, vp, vp]
        L.oracle_mst_weights.restype = None
        L.oracle_mst.argtypes = [i32, vp, vp, vp, vp]
        L.oracle_mst.restype = C.c_int
        _lib = L
    return _lib


def _csr(nindex, nlist):
    ni = np.ascontiguousarray(nindex, dtype=np.int32)
    nl = np.ascontiguousarray(nlist, dtype=np.int32)
    return ni, nl, ni.shape[0] - 1


def _call(fname, nindex, nlist, *extra):
    ni, nl, n = _csr(nindex, nlist)
    out = np.empty(n, dtype=np.int32)
    rc = getattr(lib(), fname)(n, ni.ctypes.data, nl.ctypes.data if nl.size else None,
                               out.ctypes.data, *extra)
    if rc is not None and rc < 0:
        raise RuntimeError(f"{fname} failed with {rc}")
    return out


def cc_labels(nindex, nlist) -> np.ndarray:
    """Min-ID component labels by serial union-find."""
    return _call("oracle_cc_unionfind", nindex, nlist)


def cc_bfs(nindex, nlist) -> np.ndarray:
    """Min-ID component labels by BFS in increasing start order."""
    return _call("oracle_cc_bfs", nindex, nlist)


def init_labels(nindex, nlist) -> np.ndarray:
    """The reference's init kernel output (ECL-CC-atomic.cu:69-88)."""
    return _call("oracle_init", nindex, nlist)


def eclcc_serial(nindex, nlist) -> np.ndarray:
    return _call("oracle_eclcc_serial", nindex, nlist)


def eclcc_threads(nindex, nlist, nthreads: int = 0):
    ni, nl, n = _csr(nindex, nlist)
    out = np.empty(n, dtype=np.int32)
    used = lib().oracle_eclcc_threads(n, ni.ctypes.data, nl.ctypes.data if nl.size else None,
                                      out.ctypes.data, int(nthreads))
    if used < 0:
        raise RuntimeError(f"oracle_eclcc_threads failed with {used}")
    return out, used


def verify(nindex, nlist, labels):
    """(code, ncomp): code 0 iff labels are the min-ID component labels."""
    ni, nl, n = _csr(nindex, nlist)
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    nc = C.c_int64(-1)
    rc = lib().oracle_verify(n, ni.ctypes.data, nl.ctypes.data if nl.size else None,
                             lab.ctypes.data, C.byref(nc))
    return rc, nc.value


def count_components(labels) -> int:
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    return int(lib().oracle_count_components(lab.shape[0], lab.ctypes.data))


def mst_weights(nindex, nlist) -> np.ndarray:
    """The reference's default edge weights (ECL-MST_10-atomic.cu:428-437)."""
    ni, nl, n = _csr(nindex, nlist)
    w = np.empty(nl.shape[0], dtype=np.int32)
    lib().oracle_mst_weights(n, ni.ctypes.data, nl.ctypes.data if nl.size else None,
                             w.ctypes.data if w.size else None)
    return w


def mst(nindex, nlist, weights) -> np.ndarray:
    """in_mst[j] (uint8) of the minimum spanning forest under (weight, j), by
    Kruskal (ECL-MST_10-atomic.cu:112-154)."""
    ni, nl, n = _csr(nindex, nlist)
    w = np.ascontiguousarray(weights, dtype=np.int32)
    out = np.empty(max(nl.shape[0], 1), dtype=np.uint8)
    rc = lib().oracle_mst(n, ni.ctypes.data, nl.ctypes.data if nl.size else None,
                          w.ctypes.data if w.size else None, out.ctypes.data)
    if rc < 0:
        raise RuntimeError(f"oracle_mst failed with {rc}")
    return out[: nl.shape[0]]
