"""Graph generation and the .config/.idx/.adj(/.wgt) triple, over ctypes to
libipregel_graph.so (C, OpenMP). Arrays come back as numpy arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "lib" / "libipregel_graph.so"
_lib = None


class _Csr(C.Structure):
    _fields_ = [("V", C.c_uint64), ("E", C.c_uint64), ("idx", C.POINTER(C.c_uint64)),
                ("adj", C.POINTER(C.c_uint32)), ("wgt", C.POINTER(C.c_uint32))]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"{_LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(str(_LIB_PATH))
        P = C.POINTER(_Csr)
        L.ipr_graph_rmat.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                     C.c_uint64, C.c_int, P]
        L.ipr_graph_grid.argtypes = [C.c_uint32, C.c_uint32, P]
        L.ipr_graph_weights.argtypes = [P, C.c_uint64, C.c_uint32, C.c_uint32]
        L.ipr_graph_from_arcs.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_int, P]
        L.ipr_graph_transpose.argtypes = [P, P]
        L.ipr_graph_write_triple.argtypes = [C.c_char_p, P]
        L.ipr_graph_read_triple.argtypes = [C.c_char_p, C.c_int, P]
        L.ipr_graph_free.argtypes = [P]
        L.ipr_graph_free.restype = None
        L.ipr_graph_max_degree_vertex.argtypes = [P]
        L.ipr_graph_max_degree_vertex.restype = C.c_uint64
        _lib = L
    return _lib


@dataclass
class Graph:
    """Out-edge CSR: idx (V+1, uint64), adj (E, uint32), wgt (E, uint32) or None."""
    idx: np.ndarray
    adj: np.ndarray
    wgt: np.ndarray | None = None

    @property
    def V(self) -> int:
        return int(self.idx.shape[0] - 1)

    @property
    def E(self) -> int:
        return int(self.adj.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.idx)

    def max_degree_vertex(self) -> int:
        d = self.degrees()
        return int(np.argmax(d)) if d.size else 0

    def _as_c(self) -> _Csr:
        c = _Csr()
        c.V, c.E = self.V, self.E
        c.idx = self.idx.ctypes.data_as(C.POINTER(C.c_uint64))
        c.adj = self.adj.ctypes.data_as(C.POINTER(C.c_uint32))
        c.wgt = self.wgt.ctypes.data_as(C.POINTER(C.c_uint32)) if self.wgt is not None else None
        return c

    def with_weights(self, seed: int, wmin: int = 1, wmax: int = 255) -> "Graph":
        c = self._as_c()
        c.wgt = None
        rc = lib().ipr_graph_weights(C.byref(c), seed, wmin, wmax)
        if rc:
            raise RuntimeError(f"ipr_graph_weights failed ({rc})")
        w = np.ctypeslib.as_array(c.wgt, shape=(self.E,)).copy() if self.E else np.zeros(0, np.uint32)
        # release only the C-allocated weight array
        c.idx = None
        c.adj = None
        lib().ipr_graph_free(C.byref(c))
        return Graph(self.idx, self.adj, w)

    def transpose(self) -> "Graph":
        c = self._as_c()
        out = _Csr()
        rc = lib().ipr_graph_transpose(C.byref(c), C.byref(out))
        if rc:
            raise RuntimeError(f"ipr_graph_transpose failed ({rc})")
        return _take(out)

    def write(self, root: str | Path) -> None:
        c = self._as_c()
        rc = lib().ipr_graph_write_triple(str(root).encode(), C.byref(c))
        if rc:
            raise RuntimeError(f"ipr_graph_write_triple failed ({rc})")


def _take(c: _Csr) -> Graph:
    V, E = int(c.V), int(c.E)
    idx = np.ctypeslib.as_array(c.idx, shape=(V + 1,)).copy()
    adj = np.ctypeslib.as_array(c.adj, shape=(E,)).copy() if E else np.zeros(0, np.uint32)
    wgt = None
    if c.wgt:
        wgt = np.ctypeslib.as_array(c.wgt, shape=(E,)).copy() if E else np.zeros(0, np.uint32)
    lib().ipr_graph_free(C.byref(c))
    return Graph(idx, adj, wgt)


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 1, symmetrise: bool = False) -> Graph:
    out = _Csr()
    rc = lib().ipr_graph_rmat(scale, edge_factor, a, b, c, seed, int(symmetrise), C.byref(out))
    if rc:
        raise RuntimeError(f"ipr_graph_rmat failed ({rc})")
    return _take(out)


def grid(rows: int, cols: int) -> Graph:
    out = _Csr()
    rc = lib().ipr_graph_grid(rows, cols, C.byref(out))
    if rc:
        raise RuntimeError(f"ipr_graph_grid failed ({rc})")
    return _take(out)


def from_arcs(V: int, src, dst, keep_self_loops: bool = False) -> Graph:
    s = np.ascontiguousarray(src, dtype=np.uint32)
    d = np.ascontiguousarray(dst, dtype=np.uint32)
    out = _Csr()
    rc = lib().ipr_graph_from_arcs(V, s.size, s.ctypes.data, d.ctypes.data, int(keep_self_loops),
                                   C.byref(out))
    if rc:
        raise RuntimeError(f"ipr_graph_from_arcs failed ({rc})")
    return _take(out)


def read(root: str | Path, with_weights: bool = False) -> Graph:
    out = _Csr()
    rc = lib().ipr_graph_read_triple(str(root).encode(), int(with_weights), C.byref(out))
    if rc:
        raise RuntimeError(f"ipr_graph_read_triple failed ({rc})")
    return _take(out)
