"""Python face of the CC path, mirroring the reference's host interface.

``compute_cc(nodes, edges, nidx, nlist, nstat)`` has the argument order and
meaning of the reference's ``computeCC`` (src/racefree/egr-input/CC/
ECL-CC-atomic.cu:275): host CSR in, labels written into ``nstat``.  Errors
raise ``EclccError`` instead of ``exit(-1)``.

``DeviceGraph`` keeps the CSR, labels and workspace resident in HBM (torch
tensors are only the allocator) and launches the kernel sequence on a stream.
Every call goes through libeclcc.so; nothing here computes labels itself.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from . import _native as N
from .graph import Graph

KERNELS = ("init", "sample", "pick", "mark", "hubscan", "compute1", "compute2", "compute3",
           "flatten")


def default_params(**overrides) -> N.EclccParams:
    p = N.EclccParams()
    N.cc_lib().eclcc_params_default(C.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, int(v))
    return p


def params_for_graph(g: Graph, **overrides) -> N.EclccParams:
    """The defaults tuned for g's class (eclcc_params_for_graph: degree
    statistics of the host CSR), then `overrides`."""
    p = N.EclccParams()
    ni = np.ascontiguousarray(g.nindex, dtype=np.int32)
    N.check_cc(N.cc_lib().eclcc_params_for_graph(g.nodes, g.edges, ni.ctypes.data, C.byref(p)))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, int(v))
    return p


def _pp(params: Optional[N.EclccParams]):
    return None if params is None else C.byref(params)


def version() -> str:
    return N.cc_lib().eclcc_version().decode()


def compute_cc(nodes: int, edges: int, nidx: np.ndarray, nlist: np.ndarray, nstat: np.ndarray,
               params: Optional[N.EclccParams] = None) -> float:
    """computeCC (ECL-CC-atomic.cu:275-325): fills nstat; returns the
    device time of the kernel sequence in seconds."""
    for a in (nidx, nlist, nstat):
        if a.dtype != np.int32 or not a.flags.c_contiguous:
            raise TypeError("CSR and label arrays must be contiguous int32")
    if nidx.shape[0] != nodes + 1 or nlist.shape[0] < edges or nstat.shape[0] < nodes:
        raise ValueError("array lengths do not match nodes/edges")
    secs = C.c_double(0.0)
    N.check_cc(N.cc_lib().eclcc_compute_timed(
        int(nodes), int(edges), nidx.ctypes.data, nlist.ctypes.data if edges else None,
        nstat.ctypes.data, _pp(params), C.byref(secs)))
    return secs.value


def compute(g: Graph, params: Optional[N.EclccParams] = None) -> np.ndarray:
    labels = np.empty(g.nodes, dtype=np.int32)
    compute_cc(g.nodes, g.edges, g.nindex, g.nlist, labels, params)
    return labels


def compute_timed(g: Graph, params: Optional[N.EclccParams] = None) -> Tuple[np.ndarray, float]:
    labels = np.empty(g.nodes, dtype=np.int32)
    secs = compute_cc(g.nodes, g.edges, g.nindex, g.nlist, labels, params)
    return labels, secs


class DeviceGraph:
    """CSR + labels + workspace resident on one GPU."""

    def __init__(self, g: Graph, device: str = "cuda:0", params: Optional[N.EclccParams] = None):
        import torch

        self.torch = torch
        self.device = torch.device(device)
        self.params = params if params is not None else params_for_graph(g)
        self.nodes, self.edges = g.nodes, g.edges
        self.nindex = torch.from_numpy(np.ascontiguousarray(g.nindex)).to(self.device)
        self.nlist = torch.from_numpy(np.ascontiguousarray(g.nlist)).to(self.device) if g.edges else \
            torch.zeros(1, dtype=torch.int32, device=self.device)
        self.nstat = torch.empty(self.nodes, dtype=torch.int32, device=self.device)
        wb = N.cc_lib().eclcc_workspace_bytes(self.nodes, self.edges, C.byref(self.params))
        if wb == 0:
            raise N.EclccError(-1, "bad graph size or parameters for the workspace")
        self.work = torch.empty(int(wb), dtype=torch.uint8, device=self.device)
        self.work_bytes = int(wb)
        torch.cuda.synchronize(self.device)

    def set_params(self, params: N.EclccParams) -> None:
        """Switches to other thresholds and modes for later runs; the CSR
        stays where it is, and the workspace grows if they need more."""
        wb = N.cc_lib().eclcc_workspace_bytes(self.nodes, self.edges, C.byref(params))
        if wb == 0:
            raise N.EclccError(-1, "bad graph size or parameters for the workspace")
        if wb > self.work_bytes:
            self.torch.cuda.synchronize(self.device)
            self.work = self.torch.empty(int(wb), dtype=self.torch.uint8, device=self.device)
            self.work_bytes = int(wb)
        self.params = params

    def _stream(self, stream) -> Optional[int]:
        if stream is None:
            stream = self.torch.cuda.current_stream(self.device)
        return stream.cuda_stream

    def run(self, stream=None) -> None:
        """Enqueues init..flatten on `stream` (default: torch's current)."""
        with self.torch.cuda.device(self.device):
            N.check_cc(N.cc_lib().eclcc_run_device(
                self.nodes, self.edges, self.nindex.data_ptr(), self.nlist.data_ptr(),
                self.nstat.data_ptr(), self.work.data_ptr(), self.work_bytes,
                C.byref(self.params), self._stream(stream)))

    def run_profiled(self, stream=None) -> Tuple[list, float]:
        """Per-kernel milliseconds (in KERNELS order) and the whole window,
        from HIP events on the launch stream."""
        k = (C.c_float * len(KERNELS))()
        tot = C.c_float(0.0)
        with self.torch.cuda.device(self.device):
            N.check_cc(N.cc_lib().eclcc_run_device_profiled(
                self.nodes, self.edges, self.nindex.data_ptr(), self.nlist.data_ptr(),
                self.nstat.data_ptr(), self.work.data_ptr(), self.work_bytes,
                C.byref(self.params), self._stream(stream), k, C.byref(tot)))
        return list(k), tot.value

    def run_init_only(self, stream=None) -> None:
        with self.torch.cuda.device(self.device):
            N.check_cc(N.cc_lib().eclcc_run_init_only(
                self.nodes, self.edges, self.nindex.data_ptr(), self.nlist.data_ptr(),
                self.nstat.data_ptr(), self.work.data_ptr(), self.work_bytes, self._stream(stream)))

    def hub_stats(self, stream=None) -> Tuple[int, int]:
        h, t = C.c_int64(0), C.c_int64(0)
        with self.torch.cuda.device(self.device):
            N.check_cc(N.cc_lib().eclcc_workspace_stats(self.work.data_ptr(), self.work_bytes,
                                                        self._stream(stream), C.byref(h), C.byref(t)))
        return h.value, t.value

    def labels(self) -> np.ndarray:
        self.torch.cuda.synchronize(self.device)
        return self.nstat.cpu().numpy()
