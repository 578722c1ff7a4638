"""Continuous-batching CC server (``eclcc_server_*`` in libeclcc.so).

Requests are host CSR arrays (copied in, labels copied back) or graphs that
already live in device memory (labels written in place, nothing crosses
PCIe).  The native worker admits them by HBM budget, batches whatever has
arrived into one launch sequence per stream and overlaps batches on several
streams.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Tuple

import numpy as np

from . import _native as N
from .graph import Graph


class Server:
    def __init__(self, hbm_budget_bytes: int = 0, max_inflight: int = 0, streams: int = 0,
                 max_batch: int = 0, params: Optional[N.EclccParams] = None,
                 batch_wait_us: int = 0):
        cfg = N.EclccServerConfig()
        cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        cfg.max_inflight = int(max_inflight)
        cfg.streams = int(streams)
        cfg.max_batch = int(max_batch)
        cfg.batch_wait_us = int(batch_wait_us)
        h = C.c_void_p()
        N.check_cc(N.cc_lib().eclcc_server_create(C.byref(cfg),
                                                  None if params is None else C.byref(params),
                                                  C.byref(h)))
        self._h = h
        # request id -> (graph arrays kept alive, output array)
        self._live: Dict[int, Tuple] = {}

    def submit(self, g: Graph) -> int:
        ni = np.ascontiguousarray(g.nindex, dtype=np.int32)
        nl = np.ascontiguousarray(g.nlist, dtype=np.int32)
        out = np.empty(g.nodes, dtype=np.int32)
        rid = N.cc_lib().eclcc_server_submit(self._h, g.nodes, g.edges, ni.ctypes.data,
                                             nl.ctypes.data if g.edges else None, out.ctypes.data)
        N.check_cc(int(rid))
        self._live[int(rid)] = (ni, nl, out)
        return int(rid)

    def submit_device(self, nodes: int, edges: int, nindex, nlist, labels) -> int:
        """Request on a device-resident graph: `nindex`, `nlist` and `labels`
        are contiguous int32 tensors on the server's device (labels gets
        `nodes` entries).  They are kept alive until the request is waited
        for; wait() then returns `labels`."""
        for t, need in ((nindex, nodes + 1), (nlist, edges), (labels, nodes)):
            if not t.is_cuda or not t.is_contiguous() or str(t.dtype) != "torch.int32" \
                    or t.numel() < need:
                raise ValueError("nindex/nlist/labels must be contiguous int32 device tensors "
                                 "of at least nodes+1 / edges / nodes entries")
        rid = N.cc_lib().eclcc_server_submit_device(self._h, int(nodes), int(edges),
                                                    nindex.data_ptr(),
                                                    nlist.data_ptr() if edges else None,
                                                    labels.data_ptr())
        N.check_cc(int(rid))
        self._live[int(rid)] = (nindex, nlist, labels)
        return int(rid)

    def submit_device_many(self, reqs) -> list:
        """Submits device-resident requests together: `reqs` is a sequence of
        (nodes, edges, nindex, nlist, labels) as for submit_device."""
        import torch

        reqs = list(reqs)
        k = len(reqs)
        i32 = torch.int32
        ptrs = np.zeros((3, k), dtype=np.uint64)
        # one pass, the checks in the order of their cost (a bulk submit of
        # a thousand requests spent about a millisecond here)
        for i, (nodes, edges, ni, nl, lab) in enumerate(reqs):
            for col, (t, need) in enumerate(((ni, nodes + 1), (nl, edges), (lab, nodes))):
                if t.dtype is not i32 or not t.is_cuda or t.numel() < need or not t.is_contiguous():
                    raise ValueError("nindex/nlist/labels must be contiguous int32 device tensors "
                                     "of at least nodes+1 / edges / nodes entries")
                ptrs[col, i] = t.data_ptr() if (col != 1 or edges) else 0
        nodes = np.array([r[0] for r in reqs], dtype=np.int32)
        edges = np.array([r[1] for r in reqs], dtype=np.int32)
        ni, nl, lb = ptrs[0].copy(), ptrs[1].copy(), ptrs[2].copy()
        ids = np.zeros(k, dtype=np.int64)
        N.check_cc(N.cc_lib().eclcc_server_submit_device_many(
            self._h, k, nodes.ctypes.data, edges.ctypes.data, ni.ctypes.data, nl.ctypes.data,
            lb.ctypes.data, ids.ctypes.data))
        out = [int(x) for x in ids]
        for rid, (_, _, a, b, c) in zip(out, reqs):
            self._live[rid] = (a, b, c)
        return out

    def wait(self, rid: int, timeout_ms: int = -1) -> Tuple[np.ndarray, float]:
        lat = C.c_double(0.0)
        rc = N.cc_lib().eclcc_server_wait(self._h, int(rid), int(timeout_ms), C.byref(lat))
        if rc == -9:  # timed out: the request stays live
            N.check_cc(rc)
        _, _, out = self._live.pop(int(rid))
        N.check_cc(rc)
        return out, lat.value

    def stats(self) -> dict:
        c, b, pk, mb = C.c_int64(0), C.c_int64(0), C.c_size_t(0), C.c_int32(0)
        N.check_cc(N.cc_lib().eclcc_server_stats(self._h, C.byref(c), C.byref(b), C.byref(pk),
                                                 C.byref(mb)))
        return {"completed": c.value, "batches": b.value, "peak_device_bytes": pk.value,
                "max_batch": mb.value}

    def close(self) -> None:
        if self._h is not None:
            N.cc_lib().eclcc_server_free(self._h)
            self._h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
