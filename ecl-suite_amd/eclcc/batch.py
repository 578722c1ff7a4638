"""Batched CC: many graphs packed into one device allocation and processed by
one launch sequence (optionally a replayed hipGraph).  Thin wrapper over
``eclcc_batch_*`` in libeclcc.so."""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import _native as N
from .graph import Graph


class Batch:
    def __init__(self, graphs: Sequence[Graph], params: Optional[N.EclccParams] = None):
        L = N.cc_lib()
        G = len(graphs)
        if G < 1:
            raise ValueError("empty batch")
        self.graphs = list(graphs)
        nodes = (N.i32 * G)(*[g.nodes for g in graphs])
        edges = (N.i32 * G)(*[g.edges for g in graphs])
        # keep contiguous int32 views alive during the upload
        self._ni = [np.ascontiguousarray(g.nindex, dtype=np.int32) for g in graphs]
        self._nl = [np.ascontiguousarray(g.nlist, dtype=np.int32) for g in graphs]
        pni = (C.c_void_p * G)(*[a.ctypes.data for a in self._ni])
        pnl = (C.c_void_p * G)(*[a.ctypes.data if a.size else None for a in self._nl])
        self.params = params
        h = C.c_void_p()
        N.check_cc(L.eclcc_batch_create(G, nodes, edges, pni, pnl,
                                        None if params is None else C.byref(params), C.byref(h)))
        self._h = h
        self._ni = self._nl = None

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise RuntimeError("batch freed")
        return self._h

    def run(self, stream=None, use_graph: bool = True) -> None:
        s = None if stream is None else stream.cuda_stream
        N.check_cc(N.cc_lib().eclcc_batch_run(self.handle, s, int(use_graph)))

    def run_single(self, g: int, stream=None) -> None:
        s = None if stream is None else stream.cuda_stream
        N.check_cc(N.cc_lib().eclcc_batch_run_single(self.handle, int(g), s))

    def run_profiled(self, stream=None):
        """Per-kernel milliseconds (eclcc.cc.KERNELS order) and the total."""
        from .cc import KERNELS

        k = (C.c_float * len(KERNELS))()
        tot = C.c_float(0.0)
        s = None if stream is None else stream.cuda_stream
        N.check_cc(N.cc_lib().eclcc_batch_run_profiled(self.handle, s, k, C.byref(tot)))
        return dict(zip(KERNELS, list(k))), tot.value

    def set_params(self, params: N.EclccParams) -> None:
        N.check_cc(N.cc_lib().eclcc_batch_set_params(self.handle, C.byref(params)))
        self.params = params

    def labels(self, g: int) -> np.ndarray:
        out = np.empty(self.graphs[g].nodes, dtype=np.int32)
        N.check_cc(N.cc_lib().eclcc_batch_labels(self.handle, int(g), out.ctypes.data))
        return out

    def all_labels(self) -> List[np.ndarray]:
        return [self.labels(g) for g in range(len(self.graphs))]

    def device_bytes(self) -> int:
        return int(N.cc_lib().eclcc_batch_device_bytes(self.handle))

    def close(self) -> None:
        if self._h is not None:
            N.cc_lib().eclcc_batch_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
