"""ctypes bindings of the in-tree libraries.

libeclcc.so    -- CC: HIP kernels + C ABI (include/eclcc.h)
libeclmst.so   -- MST: HIP kernels + C ABI (include/eclmst.h)
libeclgraph.so -- host graph I/O and generators (include/eclgraph.h)

All are built in-tree by ``make -C ecl-suite_amd`` (``__graft_entry__.build()``).
There is no fallback: if a library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent      # .../ecl-suite_amd
LIB_DIR = PKG_ROOT / "lib"
BIN_DIR = PKG_ROOT / "bin"
REPO_ROOT = PKG_ROOT.parent
INCLUDE_DIR = REPO_ROOT / "include"

i32 = C.c_int32
i64 = C.c_int64
u64 = C.c_uint64
vp = C.c_void_p
P_i32 = C.POINTER(C.c_int32)


class EclccParams(C.Structure):
    """Mirror of ``eclcc_params`` (include/eclcc.h)."""

    _fields_ = [
        ("thread_max_deg", i32),
        ("wave_max_deg", i32),
        ("hub_chunk", i32),
        ("hub_blocks", i32),
        ("block_threads", i32),
        ("giant_filter", i32),
        ("path_mode", i32),
        ("sample_shift", i32),
        ("sample_min_deg", i32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class EclccServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_size_t),
        ("max_inflight", i32),
        ("streams", i32),
        ("max_batch", i32),
        ("batch_wait_us", i32),
        ("reserved", i32 * 4),
    ]


class NativeLibraryMissing(RuntimeError):
    pass


def _load(name: str) -> C.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise NativeLibraryMissing(
            f"{path} is not built; run `make -C {PKG_ROOT}` (or __graft_entry__.build())")
    return C.CDLL(str(path))


_cc = None
_graph = None


def graph_lib() -> C.CDLL:
    global _graph
    if _graph is None:
        L = _load("libeclgraph.so")
        H = vp  # opaque eclgraph_csr*
        L.eclgraph_last_error.restype = C.c_char_p
        L.eclgraph_read.argtypes = [C.c_char_p, C.POINTER(H)]
        L.eclgraph_write.argtypes = [H, C.c_char_p]
        L.eclgraph_free.argtypes = [H]
        L.eclgraph_free.restype = None
        L.eclgraph_from_arrays.argtypes = [i32, i64, vp, vp, vp, C.POINTER(H)]
        L.eclgraph_from_edges.argtypes = [i32, i64, vp, vp, C.c_int, C.POINTER(H)]
        L.eclgraph_check.argtypes = [i32, i64, vp, vp, C.c_int, C.POINTER(i64)]
        L.eclgraph_read_mtx.argtypes = [C.c_char_p, C.c_int, C.POINTER(H)]
        L.eclgraph_mm2ecl.argtypes = [C.c_char_p, C.POINTER(H), C.POINTER(C.c_int)]
        L.eclgraph_nodes.argtypes = [H]
        L.eclgraph_nodes.restype = i32
        L.eclgraph_edges.argtypes = [H]
        L.eclgraph_edges.restype = i64
        for f in ("eclgraph_nindex", "eclgraph_nlist", "eclgraph_eweight"):
            getattr(L, f).argtypes = [H]
            getattr(L, f).restype = vp
        L.eclgraph_gen_er.argtypes = [i32, i64, u64, C.POINTER(H)]
        L.eclgraph_gen_rmat.argtypes = [C.c_int, i32, i64, C.c_double, C.c_double, C.c_double,
                                        u64, C.c_int, C.POINTER(H)]
        L.eclgraph_gen_grid3d.argtypes = [i32, i32, i32, C.c_double, u64, C.c_int, C.POINTER(H)]
        L.eclgraph_rand64.argtypes = [u64, u64]
        L.eclgraph_rand64.restype = u64
        L.eclgraph_set_threads.argtypes = [C.c_int]
        L.eclgraph_set_threads.restype = None
        L.eclgraph_mst_weights.argtypes = [i32, vp, vp, vp]
        _graph = L
    return _graph


def _declare_cc(L: C.CDLL) -> C.CDLL:
    PP = C.POINTER(EclccParams)
    L.eclcc_last_error.restype = C.c_char_p
    L.eclcc_status_string.argtypes = [C.c_int]
    L.eclcc_status_string.restype = C.c_char_p
    L.eclcc_version.restype = C.c_char_p
    L.eclcc_kernel_name.argtypes = [C.c_int]
    L.eclcc_kernel_name.restype = C.c_char_p
    L.eclcc_params_default.argtypes = [PP]
    L.eclcc_params_default.restype = None
    L.eclcc_params_for_graph.argtypes = [i32, i64, vp, PP]
    L.eclcc_compute.argtypes = [i32, i32, vp, vp, vp]
    L.eclcc_compute_timed.argtypes = [i32, i32, vp, vp, vp, PP, C.POINTER(C.c_double)]
    L.eclcc_workspace_bytes.argtypes = [i32, i64, PP]
    L.eclcc_workspace_bytes.restype = C.c_size_t
    L.eclcc_run_device.argtypes = [i32, i32, vp, vp, vp, vp, C.c_size_t, PP, vp]
    L.eclcc_run_device_noalloc.argtypes = [i32, i32, vp, vp, vp, PP, vp]
    L.eclcc_run_device_profiled.argtypes = [i32, i32, vp, vp, vp, vp, C.c_size_t, PP, vp,
                                            C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.eclcc_run_init_only.argtypes = [i32, i32, vp, vp, vp, vp, C.c_size_t, vp]
    L.eclcc_workspace_stats.argtypes = [vp, C.c_size_t, vp, C.POINTER(i64), C.POINTER(i64)]
    L.eclcc_batch_create.argtypes = [i32, vp, vp, vp, vp, PP, C.POINTER(vp)]
    L.eclcc_batch_run.argtypes = [vp, vp, C.c_int]
    L.eclcc_batch_run_single.argtypes = [vp, i32, vp]
    L.eclcc_batch_set_params.argtypes = [vp, PP]
    L.eclcc_batch_run_profiled.argtypes = [vp, vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.eclcc_batch_labels.argtypes = [vp, i32, vp]
    L.eclcc_batch_device_labels.argtypes = [vp, i32]
    L.eclcc_batch_device_labels.restype = vp
    L.eclcc_batch_device_bytes.argtypes = [vp]
    L.eclcc_batch_device_bytes.restype = C.c_size_t
    L.eclcc_batch_free.argtypes = [vp]
    L.eclcc_batch_free.restype = None
    L.eclcc_server_create.argtypes = [C.POINTER(EclccServerConfig), PP, C.POINTER(vp)]
    L.eclcc_server_submit.argtypes = [vp, i32, i32, vp, vp, vp]
    L.eclcc_server_submit.restype = i64
    L.eclcc_server_submit_device.argtypes = [vp, i32, i32, vp, vp, vp]
    L.eclcc_server_submit_device.restype = i64
    L.eclcc_server_submit_device_many.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp]
    L.eclcc_server_wait.argtypes = [vp, i64, i32, C.POINTER(C.c_double)]
    L.eclcc_server_stats.argtypes = [vp, C.POINTER(i64), C.POINTER(i64),
                                     C.POINTER(C.c_size_t), C.POINTER(i32)]
    L.eclcc_server_free.argtypes = [vp]
    L.eclcc_server_free.restype = None
    return L


def cc_lib() -> C.CDLL:
    global _cc
    if _cc is None:
        _cc = _declare_cc(_load("libeclcc.so"))
    return _cc


_plain = None


def plain_lib() -> C.CDLL:
    """Research copy of the library whose label loads and stores are plain
    (racy) memory accesses, built only to measure what race freedom costs.
    Nothing else in the package uses it."""
    global _plain
    if _plain is None:
        _plain = _declare_cc(_load("libeclcc_plain.so"))
    return _plain


_mst = None


def mst_lib() -> C.CDLL:
    global _mst
    if _mst is None:
        L = _load("libeclmst.so")
        L.eclmst_last_error.restype = C.c_char_p
        L.eclmst_version.restype = C.c_char_p
        L.eclmst_threshold.argtypes = [i32, i32, vp]
        L.eclmst_threshold.restype = i32
        L.eclmst_compute.argtypes = [i32, i32, vp, vp, vp, vp]
        L.eclmst_compute_timed.argtypes = [i32, i32, vp, vp, vp, vp, C.POINTER(C.c_double)]
        L.eclmst_workspace_bytes.argtypes = [i32, i32]
        L.eclmst_workspace_bytes.restype = C.c_size_t
        L.eclmst_run_device.argtypes = [i32, i32, vp, vp, vp, vp, vp, C.c_size_t, i32, vp,
                                        C.POINTER(C.c_float)]
        _mst = L
    return _mst


class EclmstError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"eclmst status {status}: {message}")
        self.status = status


def check_mst(rc: int) -> int:
    if rc < 0:
        raise EclmstError(rc, (mst_lib().eclmst_last_error() or b"").decode())
    return rc


class EclccError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"eclcc status {status}: {message}")
        self.status = status


class EclgraphError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"eclgraph status {status}: {message}")
        self.status = status


def check_cc(rc: int) -> int:
    if rc < 0:
        L = cc_lib()
        raise EclccError(rc, (L.eclcc_last_error() or b"").decode())
    return rc


def check_graph(rc: int) -> int:
    if rc != 0:
        L = graph_lib()
        raise EclgraphError(rc, (L.eclgraph_last_error() or b"").decode())
    return rc


def header_functions(header: str) -> list[str]:
    """Names of the functions an include/*.h header declares."""
    import re

    text = (INCLUDE_DIR / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b((?:eclcc|eclmst|eclgraph)_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


os.environ.setdefault("ECLGRAPH_THREADS", str(min(16, os.cpu_count() or 1)))
