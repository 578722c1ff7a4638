"""Race-free ECL-MST minimum spanning forest (libeclmst.so, include/eclmst.h).

Mirror of the reference's host interface for the MST path
(src/racefree/egr-input/MST/ECL-MST_10-atomic.cu): ``compute_mst`` is
``gpuMST`` with main's threshold choice (:285-382, :450-462), ``mst_weights``
is main's default weights (:428-437).  The output is the reference's
``inMST`` array: one byte per CSR entry, 1 for the entry j of vertex i with
nlist[j] > i of every forest edge.  No CPU fallback: calling without the
built library raises NativeLibraryMissing.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _native as N
from .graph import Graph


def mst_weights(g: Graph) -> np.ndarray:
    """The reference's default weights for g's entries."""
    ni = np.ascontiguousarray(g.nindex, dtype=np.int32)
    nl = np.ascontiguousarray(g.nlist, dtype=np.int32)
    w = np.empty(nl.shape[0], dtype=np.int32)
    N.check_graph(N.graph_lib().eclgraph_mst_weights(
        g.nodes, ni.ctypes.data, nl.ctypes.data if nl.size else None,
        w.ctypes.data if w.size else None))
    return w


def threshold(g: Graph, weights: np.ndarray) -> int:
    w = np.ascontiguousarray(weights, dtype=np.int32)
    return int(N.mst_lib().eclmst_threshold(g.nodes, g.edges, w.ctypes.data if w.size else None))


def compute_mst(g: Graph, weights: Optional[np.ndarray] = None, timed: bool = False):
    """in_mst (uint8 per entry); with timed=True also the device seconds."""
    if weights is None:
        weights = mst_weights(g)
    ni = np.ascontiguousarray(g.nindex, dtype=np.int32)
    nl = np.ascontiguousarray(g.nlist, dtype=np.int32)
    w = np.ascontiguousarray(weights, dtype=np.int32)
    if w.shape[0] != nl.shape[0]:
        raise ValueError("one weight per CSR entry")
    out = np.zeros(max(nl.shape[0], 1), dtype=np.uint8)
    L = N.mst_lib()
    args = (g.nodes, g.edges, ni.ctypes.data, nl.ctypes.data if nl.size else None,
            w.ctypes.data if w.size else None, out.ctypes.data)
    if timed:
        sec = C.c_double(0.0)
        N.check_mst(L.eclmst_compute_timed(*args, C.byref(sec)))
        return out[: nl.shape[0]], sec.value
    N.check_mst(L.eclmst_compute(*args))
    return out[: nl.shape[0]]


class DeviceMST:
    """A weighted graph resident in device memory, with its workspace; run()
    computes the forest into a device byte array (eclmst_run_device)."""

    def __init__(self, g: Graph, weights: np.ndarray, device="cuda:0"):
        import torch

        self.g = g
        self.dev = torch.device(device)
        self.nindex = torch.from_numpy(np.ascontiguousarray(g.nindex, dtype=np.int32)).to(self.dev)
        self.nlist = torch.from_numpy(np.ascontiguousarray(g.nlist, dtype=np.int32)).to(self.dev)
        self.weights = torch.from_numpy(np.ascontiguousarray(weights, dtype=np.int32)).to(self.dev)
        self.in_mst = torch.zeros(max(g.edges, 1), dtype=torch.uint8, device=self.dev)
        L = N.mst_lib()
        self.work_bytes = int(L.eclmst_workspace_bytes(g.nodes, g.edges))
        self.work = torch.empty(self.work_bytes, dtype=torch.uint8, device=self.dev)

    def run(self, threshold: int = -1, stream=None) -> float:
        """Computes the forest; returns the device milliseconds of the kernels."""
        import torch

        s = stream if stream is not None else torch.cuda.current_stream(self.dev)
        ms = C.c_float(0.0)
        N.check_mst(N.mst_lib().eclmst_run_device(
            self.g.nodes, self.g.edges, self.nindex.data_ptr(), self.nlist.data_ptr(),
            self.weights.data_ptr(), self.in_mst.data_ptr(), self.work.data_ptr(),
            self.work_bytes, int(threshold), s.cuda_stream, C.byref(ms)))
        return ms.value

    def result(self) -> np.ndarray:
        return self.in_mst.cpu().numpy()[: self.g.edges]


def version() -> str:
    return N.mst_lib().eclmst_version().decode()
