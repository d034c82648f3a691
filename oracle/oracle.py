"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
PUSH, PULL, SPREAD, SPREAD_PULL = 0, 1, 2, 3
PR, CC, SSSP, SSSP_W = 0, 1, 2, 3
VARIANTS = {"": PUSH, "_single_broadcast": PULL, "_spread": SPREAD,
            "_single_broadcast_spread": SPREAD_PULL}


class _Csr(C.Structure):
    _fields_ = [("V", C.c_uint64), ("E", C.c_uint64), ("idx", C.c_void_p),
                ("adj", C.c_void_p), ("wgt", C.c_void_p)]


class _Params(C.Structure):
    _fields_ = [("source", C.c_uint32), ("rounds", C.c_uint32), ("damping", C.c_double),
                ("max_supersteps", C.c_uint64), ("stop_after", C.c_uint64),
                ("stop_seconds", C.c_double)]


class _Result(C.Structure):
    _fields_ = [("supersteps", C.c_uint64), ("work_edges", C.c_uint64),
                ("work_vertices", C.c_uint64), ("messages", C.c_uint64)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        p = HERE / "liboracle.so"
        if not p.exists():
            raise RuntimeError(f"{p} is not built; run __graft_entry__.build()")
        L = C.CDLL(str(p))
        L.ipr_oracle_run.argtypes = [C.POINTER(_Csr), C.c_int, C.c_int, C.c_int, C.POINTER(_Params),
                                     C.c_void_p, C.c_void_p, C.POINTER(_Result)]
        L.ipr_oracle_dump.argtypes = [C.c_int, C.c_uint64, C.c_void_p, C.c_char_p]
        for f in ("ipr_oracle_min_ancestor",):
            getattr(L, f).argtypes = [C.POINTER(_Csr), C.c_void_p]
        L.ipr_oracle_ppr.argtypes = [C.POINTER(_Csr), C.c_uint32, C.c_uint32, C.c_void_p]
        L.ipr_oracle_ppr_many.argtypes = [C.POINTER(_Csr), C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        for f in ("ipr_oracle_bfs", "ipr_oracle_dijkstra"):
            getattr(L, f).argtypes = [C.POINTER(_Csr), C.c_uint32, C.c_void_p]
        _lib = L
    return _lib


def _csr(g) -> tuple[_Csr, tuple]:
    idx = np.ascontiguousarray(g.idx, dtype=np.uint64)
    adj = np.ascontiguousarray(g.adj, dtype=np.uint32)
    wgt = np.ascontiguousarray(g.wgt, dtype=np.uint32) if g.wgt is not None else None
    c = _Csr(g.V, g.E, idx.ctypes.data, adj.ctypes.data, wgt.ctypes.data if wgt is not None else None)
    return c, (idx, adj, wgt)


@dataclass
class OracleResult:
    values: np.ndarray
    active: np.ndarray
    supersteps: int
    work_edges: int
    work_vertices: int


def run(g, app: int, variant: int, directed: bool = False, source: int = 0,
        rounds: int = 10, max_supersteps: int = 1 << 20, stop_after: int = 0,
        stop_seconds: float = 0.0) -> OracleResult:
    """A bounded sample: stop_after > 0 ends the run after that many
    supersteps, stop_seconds > 0 at the first superstep boundary after that
    much run time."""
    c, keep = _csr(g)
    p = _Params(source, rounds, 0.85, max_supersteps, stop_after, stop_seconds)
    vals = np.zeros(g.V, dtype=np.float64 if app == PR else np.uint32)
    trace = np.zeros(max_supersteps, dtype=np.uint64)
    r = _Result()
    rc = lib().ipr_oracle_run(C.byref(c), int(directed), variant, app, C.byref(p),
                              vals.ctypes.data, trace.ctypes.data, C.byref(r))
    if rc:
        raise RuntimeError(f"oracle run failed ({rc})")
    n = int(r.supersteps)
    return OracleResult(vals, trace[:min(n, max_supersteps)].copy(), n, int(r.work_edges),
                        int(r.work_vertices))


def dump(app: int, values: np.ndarray, path: str) -> None:
    rc = lib().ipr_oracle_dump(app, values.size, values.ctypes.data, path.encode())
    if rc:
        raise RuntimeError(f"oracle dump failed ({rc})")


def min_ancestor(g) -> np.ndarray:
    c, keep = _csr(g)
    out = np.zeros(g.V, np.uint32)
    rc = lib().ipr_oracle_min_ancestor(C.byref(c), out.ctypes.data)
    if rc:
        raise RuntimeError(f"min_ancestor failed ({rc})")
    return out


def bfs(g, source: int) -> np.ndarray:
    c, keep = _csr(g)
    out = np.zeros(g.V, np.uint32)
    rc = lib().ipr_oracle_bfs(C.byref(c), source, out.ctypes.data)
    if rc:
        raise RuntimeError(f"bfs failed ({rc})")
    return out


def dijkstra(g, source: int) -> np.ndarray:
    c, keep = _csr(g)
    out = np.zeros(g.V, np.uint32)
    rc = lib().ipr_oracle_dijkstra(C.byref(c), source, out.ctypes.data)
    if rc:
        raise RuntimeError(f"dijkstra failed ({rc})")
    return out


def ppr(g, seed: int, iters: int = 20) -> np.ndarray:
    c, keep = _csr(g)
    out = np.zeros(g.V, np.float64)
    rc = lib().ipr_oracle_ppr(C.byref(c), seed, iters, out.ctypes.data)
    if rc:
        raise RuntimeError(f"ppr failed ({rc})")
    return out


def ppr_many(g, seeds, iters: int = 20) -> np.ndarray:
    """ppr for every seed (rows), bit for bit, with the host's OpenMP threads"""
    c, keep = _csr(g)
    s = np.ascontiguousarray(seeds, dtype=np.uint32)
    out = np.zeros((s.size, g.V), np.float64)
    rc = lib().ipr_oracle_ppr_many(C.byref(c), s.ctypes.data, s.size, iters, out.ctypes.data)
    if rc:
        raise RuntimeError(f"ppr_many failed ({rc})")
    return out
