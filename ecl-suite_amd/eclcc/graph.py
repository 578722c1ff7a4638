"""Host-side ECLgraph CSR (``struct ECLgraph``, library/ECLgraph.h:41-47).

Arrays are numpy int32: ``nindex[nodes+1]``, ``nlist[edges]``, optional
``eweight[edges]``.  ``edges`` counts CSR entries (each undirected edge twice).
I/O and the seeded generators run in libeclgraph.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _native as N

SYMMETRIZE, DROP_SELF, DEDUPE = 1, 2, 4
CHECK_SYMMETRY = 1


class GraphFormatError(ValueError):
    """The CSR breaks a precondition; `where` is the offending position."""

    def __init__(self, msg: str, where: int):
        super().__init__(msg)
        self.where = where


class _CsrOwner:
    """Holds an eclgraph_csr handle; frees it when garbage collected."""

    def __init__(self, h: C.c_void_p):
        self.h = h

    def __del__(self):
        try:
            if self.h:
                N.graph_lib().eclgraph_free(self.h)
                self.h = None
        except Exception:
            pass


def _view(addr: int, count: int, owner: _CsrOwner) -> np.ndarray:
    buf = (C.c_int32 * count).from_address(addr)
    buf._owner = owner  # keeps the library memory alive as long as the array
    return np.frombuffer(buf, dtype=np.int32)


@dataclass
class Graph:
    nindex: np.ndarray
    nlist: np.ndarray
    eweight: Optional[np.ndarray] = None
    name: str = ""

    @property
    def nodes(self) -> int:
        return int(self.nindex.shape[0] - 1)

    @property
    def edges(self) -> int:
        return int(self.nlist.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.nindex)

    def algorithmic_bytes(self) -> int:
        """Compulsory bytes of one CC pass: read nindex and nlist once,
        write each label once: 4(n+1) + 4m + 4n."""
        return 4 * (self.nodes + 1) + 4 * self.edges + 4 * self.nodes

    def check(self, symmetry: bool = True) -> None:
        """Raises GraphFormatError if the CSR breaks a precondition of the CC
        path (index order, IDs in range, and with `symmetry` every entry's
        reverse present; ECL-CC-atomic.cu:132 hooks each edge only from its
        larger end)."""
        ni = np.ascontiguousarray(self.nindex, dtype=np.int32)
        nl = np.ascontiguousarray(self.nlist, dtype=np.int32)
        where = C.c_int64(-1)
        st = N.graph_lib().eclgraph_check(self.nodes, self.edges, ni.ctypes.data,
                                          nl.ctypes.data if self.edges else None,
                                          CHECK_SYMMETRY if symmetry else 0, C.byref(where))
        if st != 0:
            msg = N.graph_lib().eclgraph_last_error().decode()
            raise GraphFormatError(f"{msg} (position {where.value})", where.value)

    # -- construction through libeclgraph.so ---------------------------------

    @staticmethod
    def _from_handle(h: C.c_void_p, name: str = "") -> "Graph":
        """Wraps a library-owned CSR without copying; the library frees it
        when the last array that views it is gone."""
        L = N.graph_lib()
        owner = _CsrOwner(h)
        n = L.eclgraph_nodes(h)
        m = L.eclgraph_edges(h)
        ni = _view(L.eclgraph_nindex(h), n + 1, owner)
        nl = _view(L.eclgraph_nlist(h), m, owner) if m > 0 else np.zeros(0, dtype=np.int32)
        wp = L.eclgraph_eweight(h)
        w = _view(wp, m, owner) if (wp and m > 0) else None
        return Graph(ni, nl, w, name)

    @staticmethod
    def read(path: str) -> "Graph":
        """readECLgraph (library/ECLgraph.h:49-76)."""
        h = C.c_void_p()
        N.check_graph(N.graph_lib().eclgraph_read(str(path).encode(), C.byref(h)))
        return Graph._from_handle(h, str(path))

    def write(self, path: str) -> None:
        """writeECLgraph (library/ECLgraph.h:78-92)."""
        L = N.graph_lib()
        h = self._handle()
        try:
            N.check_graph(L.eclgraph_write(h, str(path).encode()))
        finally:
            L.eclgraph_free(h)

    def _handle(self) -> C.c_void_p:
        L = N.graph_lib()
        ni = np.ascontiguousarray(self.nindex, dtype=np.int32)
        nl = np.ascontiguousarray(self.nlist, dtype=np.int32)
        w = None if self.eweight is None else np.ascontiguousarray(self.eweight, dtype=np.int32)
        h = C.c_void_p()
        N.check_graph(L.eclgraph_from_arrays(self.nodes, self.edges, ni.ctypes.data,
                                             nl.ctypes.data if self.edges else None,
                                             None if w is None else w.ctypes.data, C.byref(h)))
        return h

    @staticmethod
    def from_edges(nodes: int, src, dst, flags: int = SYMMETRIZE | DROP_SELF | DEDUPE,
                   name: str = "") -> "Graph":
        s = np.ascontiguousarray(src, dtype=np.int32)
        d = np.ascontiguousarray(dst, dtype=np.int32)
        if s.shape != d.shape:
            raise ValueError("src and dst differ in length")
        h = C.c_void_p()
        N.check_graph(N.graph_lib().eclgraph_from_edges(
            int(nodes), int(s.shape[0]), s.ctypes.data if s.size else None,
            d.ctypes.data if d.size else None, int(flags), C.byref(h)))
        return Graph._from_handle(h, name)

    @staticmethod
    def mm2ecl(path: str):
        """MatrixMarket coordinate file -> the CSR the reference's converter
        writes (scripts/mm2ecl_real.cpp:50-184): entries as given, ordered by
        (row, column, weight), integer weights kept, real weights x 100
        truncated.  Returns (graph, kind) with kind "no", "int" or "float"
        as the reference prints it."""
        h = C.c_void_p()
        kind = C.c_int(0)
        N.check_graph(N.graph_lib().eclgraph_mm2ecl(str(path).encode(), C.byref(h), C.byref(kind)))
        return Graph._from_handle(h, str(path)), ("no", "int", "float")[kind.value]

    @staticmethod
    def read_mtx(path: str, flags: int = 0) -> "Graph":
        """mm2ecl followed by the from_edges flags on the weighted entries:
        flags = 0 is the reference's conversion; SYMMETRIZE | DROP_SELF |
        DEDUPE gives the symmetric CSR the CC path needs."""
        h = C.c_void_p()
        N.check_graph(N.graph_lib().eclgraph_read_mtx(str(path).encode(), int(flags), C.byref(h)))
        return Graph._from_handle(h, str(path))

    @staticmethod
    def er(n: int, pairs: int, seed: int) -> "Graph":
        h = C.c_void_p()
        N.check_graph(N.graph_lib().eclgraph_gen_er(int(n), int(pairs), int(seed), C.byref(h)))
        return Graph._from_handle(h, f"er-n{n}-M{pairs}-s{seed}")

    @staticmethod
    def rmat(scale: int, samples: int, seed: int, n: int = 0, a: float = 0.57, b: float = 0.19,
             c: float = 0.19, permute: bool = True) -> "Graph":
        h = C.c_void_p()
        N.check_graph(N.graph_lib().eclgraph_gen_rmat(int(scale), int(n), int(samples), a, b, c,
                                                      int(seed), int(permute), C.byref(h)))
        return Graph._from_handle(h, f"rmat-s{scale}-n{n or (1 << scale)}-S{samples}-seed{seed}")

    @staticmethod
    def grid3d(nx: int, ny: int, nz: int, keep_p: float, seed: int, permute: bool = True) -> "Graph":
        h = C.c_void_p()
        N.check_graph(N.graph_lib().eclgraph_gen_grid3d(int(nx), int(ny), int(nz), float(keep_p),
                                                        int(seed), int(permute), C.byref(h)))
        return Graph._from_handle(h, f"grid{nx}x{ny}x{nz}-p{keep_p}-s{seed}")


def rand64(seed: int, counter: int) -> int:
    """Output number `counter` of splitmix64 started at `seed` (libeclgraph)."""
    return int(N.graph_lib().eclgraph_rand64(int(seed) & (2**64 - 1), int(counter)))


def set_generator_threads(n: int) -> None:
    """Thread count of the generators for the calling thread (0 = default)."""
    N.graph_lib().eclgraph_set_threads(int(n))
