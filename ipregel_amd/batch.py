"""Python host side of libipregel_batch.so: graphs resident in HBM, batched
SSSP / personalised-PageRank queries (64 per batch, one wavefront lane per
query), CC, their one-query counterparts, and the serving loop with Poisson
traces. No compute happens here; a missing library is an error.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .graph import Graph

LIB = Path(__file__).resolve().parent / "lib" / "libipregel_batch.so"
SLOTS = 64
SSSP, PPR, CC = 0, 1, 2
KIND_NAMES = {SSSP: "sssp", PPR: "ppr", CC: "cc"}
INF = 0xFFFFFFFF


class BatchStats(C.Structure):
    _fields_ = [("supersteps", C.c_uint64), ("relaxations", C.c_uint64), ("ms", C.c_double)]


class Query(C.Structure):
    _fields_ = [("t_arrival", C.c_double), ("kind", C.c_int32), ("graph", C.c_int32),
                ("param", C.c_uint32), ("pad", C.c_uint32)]


class ServeOpts(C.Structure):
    _fields_ = [("sequential", C.c_int32), ("ppr_iters", C.c_int32), ("hbm_budget", C.c_uint64),
                ("use_graph", C.c_int32), ("pad", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("large", C.c_uint32), ("chunk", C.c_uint32), ("grid", C.c_uint32), ("delta", C.c_uint32),
                ("delta1", C.c_uint32), ("cc_mode", C.c_uint32), ("sssp_states", C.c_uint32)]


class ServeStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("ticks", C.c_uint64), ("admitted_waits", C.c_uint64),
                ("peak_slots", C.c_uint64), ("hbm_peak", C.c_uint64)]


_lib = None

EXPORTS = ("ipr_batch_last_error", "ipr_batch_graph_create", "ipr_batch_graph_destroy",
           "ipr_batch_graph_bytes", "ipr_batch_slot_bytes", "ipr_batch_sssp", "ipr_batch_ppr",
           "ipr_single_sssp", "ipr_single_ppr", "ipr_batch_cc", "ipr_serve", "ipr_batch_graph_get_tuning",
           "ipr_batch_graph_set_tuning")


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is not built; run __graft_entry__.build()")
        L = C.CDLL(str(LIB))
        L.ipr_batch_last_error.restype = C.c_char_p
        L.ipr_batch_graph_create.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_int, C.POINTER(C.c_void_p)]
        L.ipr_batch_graph_destroy.argtypes = [C.c_void_p]
        L.ipr_batch_graph_bytes.argtypes = [C.c_void_p]
        L.ipr_batch_graph_bytes.restype = C.c_uint64
        L.ipr_batch_slot_bytes.argtypes = [C.c_void_p, C.c_int]
        L.ipr_batch_slot_bytes.restype = C.c_uint64
        L.ipr_batch_graph_get_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.ipr_batch_graph_set_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        SP = C.POINTER(BatchStats)
        L.ipr_batch_sssp.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, SP]
        L.ipr_batch_ppr.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, SP]
        L.ipr_single_sssp.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, SP]
        L.ipr_single_ppr.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, SP]
        L.ipr_batch_cc.argtypes = [C.c_void_p, C.c_void_p, SP]
        L.ipr_serve.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(ServeOpts),
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(ServeStats)]
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().ipr_batch_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _stats(st: BatchStats) -> dict:
    return {"supersteps": int(st.supersteps), "relaxations": int(st.relaxations), "ms": float(st.ms)}


class ResidentGraph:
    """A graph uploaded once (CSR, its transpose, inverse degrees, hub table)."""

    def __init__(self, g: Graph, device: int = 0):
        self.V, self.E = g.V, g.E
        idx = np.ascontiguousarray(g.idx, dtype=np.uint64)
        adj = np.ascontiguousarray(g.adj, dtype=np.uint32)
        wgt = np.ascontiguousarray(g.wgt, dtype=np.uint32) if g.wgt is not None else None
        h = C.c_void_p()
        _check(lib().ipr_batch_graph_create(self.V, self.E, idx.ctypes.data, adj.ctypes.data,
                                            wgt.ctypes.data if wgt is not None else None, device,
                                            C.byref(h)), "graph upload")
        self.h = h

    def close(self) -> None:
        if self.h:
            lib().ipr_batch_graph_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def bytes(self) -> int:
        return int(lib().ipr_batch_graph_bytes(self.h))

    def slot_bytes(self, kind: int) -> int:
        return int(lib().ipr_batch_slot_bytes(self.h, kind))

    @property
    def tuning(self) -> dict:
        t = Tuning()
        _check(lib().ipr_batch_graph_get_tuning(self.h, C.byref(t)), "get tuning")
        return {f: int(getattr(t, f)) for f, _ in Tuning._fields_}

    def tune(self, **kw) -> None:
        cur = self.tuning
        cur.update(kw)
        t = Tuning(*[cur[f] for f, _ in Tuning._fields_])
        _check(lib().ipr_batch_graph_set_tuning(self.h, C.byref(t)), "set tuning")

    def batch_sssp(self, sources) -> tuple[np.ndarray, dict]:
        s = np.ascontiguousarray(sources, dtype=np.uint32)
        out = np.empty((s.size, self.V), np.uint32)
        st = BatchStats()
        _check(lib().ipr_batch_sssp(self.h, s.ctypes.data, s.size, out.ctypes.data, C.byref(st)), "batched sssp")
        return out, _stats(st)

    def sssp(self, source: int) -> tuple[np.ndarray, dict]:
        out = np.empty(self.V, np.uint32)
        st = BatchStats()
        _check(lib().ipr_single_sssp(self.h, source, out.ctypes.data, C.byref(st)), "sssp")
        return out, _stats(st)

    def cc(self) -> tuple[np.ndarray, dict]:
        out = np.empty(self.V, np.uint32)
        st = BatchStats()
        _check(lib().ipr_batch_cc(self.h, out.ctypes.data, C.byref(st)), "cc")
        return out, _stats(st)

    def batch_ppr(self, seeds, iters: int = 20) -> tuple[np.ndarray, np.ndarray, dict]:
        s = np.ascontiguousarray(seeds, dtype=np.uint32)
        out = np.empty((s.size, self.V), np.float32)
        res = np.empty(s.size, np.float64)
        st = BatchStats()
        _check(lib().ipr_batch_ppr(self.h, s.ctypes.data, s.size, iters, out.ctypes.data, res.ctypes.data,
                                   C.byref(st)), "batched ppr")
        return out, res, _stats(st)

    def ppr(self, seed: int, iters: int = 20) -> tuple[np.ndarray, float, dict]:
        out = np.empty(self.V, np.float32)
        res = np.empty(1, np.float64)
        st = BatchStats()
        _check(lib().ipr_single_ppr(self.h, seed, iters, out.ctypes.data, res.ctypes.data, C.byref(st)), "ppr")
        return out, float(res[0]), _stats(st)


# ---------------------------------------------------------------- digests


def _mix(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64, copy=True)
    z ^= z >> np.uint64(30)
    z *= np.uint64(0xBF58476D1CE4E5B9)
    z ^= z >> np.uint64(27)
    z *= np.uint64(0x94D049BB133111EB)
    z ^= z >> np.uint64(31)
    return z


def digest(values: np.ndarray) -> int:
    """The result digest the serving loop reports: sum over v of
    mix(v << 32 | bits(value_v)) modulo 2^64 (uint32 or float32 values)."""
    bits = np.ascontiguousarray(values).view(np.uint32).astype(np.uint64)
    v = np.arange(bits.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int(_mix((v << np.uint64(32)) | bits).sum(dtype=np.uint64))


# ---------------------------------------------------------------- PPR accuracy


def ppr_agreement(x32: np.ndarray, x64: np.ndarray, top: int = 100) -> dict:
    """L1 error, max relative error (over entries above 1e-12 of the total
    mass) and top-`top` rank agreement of an FP32 result against FP64."""
    x = x32.astype(np.float64)
    l1 = float(np.abs(x - x64).sum())
    m = x64 > 1e-12
    rel = float((np.abs(x[m] - x64[m]) / x64[m]).max()) if m.any() else 0.0
    k = min(top, x64.size)
    t64 = np.argsort(-x64, kind="stable")[:k]
    t32 = np.argsort(-x, kind="stable")[:k]
    return {"l1": l1, "max_rel": rel, "top_overlap": int(np.intersect1d(t64, t32).size), "top": k}


# ---------------------------------------------------------------- serving


@dataclass
class ServeResult:
    t_arrival: np.ndarray
    t_admit: np.ndarray
    t_done: np.ndarray
    digest: np.ndarray
    wall_s: float
    ticks: int
    admitted_waits: int
    peak_slots: int
    hbm_peak: int

    @property
    def latency(self) -> np.ndarray:
        return self.t_done - self.t_arrival

    def summary(self, trace: np.ndarray | None = None, names: list[str] | None = None) -> dict:
        lat = self.latency
        n = lat.size
        out = {"queries": n, "qps": n / self.wall_s if self.wall_s > 0 else 0.0,
               "p50_ms": float(np.percentile(lat, 50) * 1e3) if n else 0.0,
               "p99_ms": float(np.percentile(lat, 99) * 1e3) if n else 0.0,
               "wall_s": self.wall_s, "ticks": self.ticks, "admitted_waits": self.admitted_waits,
               "peak_slots": self.peak_slots, "hbm_peak_bytes": self.hbm_peak}
        if trace is not None:
            per = {}
            for k, kn in KIND_NAMES.items():
                m = trace["kind"] == k
                if m.any():
                    per[kn] = {"n": int(m.sum()), "p50_ms": float(np.percentile(lat[m], 50) * 1e3),
                               "p99_ms": float(np.percentile(lat[m], 99) * 1e3)}
            out["by_kind"] = per
            if names is not None:
                pg = {}
                for gi in np.unique(trace["graph"]):
                    m = trace["graph"] == gi
                    pg[names[int(gi)]] = {"n": int(m.sum()), "p50_ms": float(np.percentile(lat[m], 50) * 1e3),
                                          "max_ms": float(lat[m].max() * 1e3)}
                out["by_graph"] = pg
        return out


def poisson_trace(n: int, rate: float, mix: dict[int, float], graphs_by_kind: dict[int, list[int]],
                  V_of: list[int], seed: int = 5) -> np.ndarray:
    """Structured array of n queries: exponential inter-arrival times at `rate`
    per second, kinds drawn from `mix`, a graph uniformly among those of the
    kind, parameter uniform over the graph's vertices."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    t -= t[0]
    kinds = np.array(sorted(mix))
    p = np.array([mix[k] for k in kinds], dtype=np.float64)
    k = rng.choice(kinds, size=n, p=p / p.sum())
    q = np.zeros(n, dtype=[("t_arrival", np.float64), ("kind", np.int32), ("graph", np.int32),
                           ("param", np.uint32), ("pad", np.uint32)])
    q["t_arrival"] = t
    q["kind"] = k
    for i in range(n):
        gs = graphs_by_kind[int(k[i])]
        gi = gs[int(rng.integers(len(gs)))]
        q["graph"][i] = gi
        q["param"][i] = int(rng.integers(V_of[gi]))
    return q


def serve(graphs: list[ResidentGraph], trace: np.ndarray, *, sequential: bool = False, ppr_iters: int = 20,
          hbm_budget: int = 0, use_graph: int = 1, results: bool = False) -> tuple[ServeResult, list]:
    """Run a trace through ipr_serve. With results=True, every query's full
    result vector is returned too (for tests; costs a copy per query)."""
    n = trace.size
    q = np.ascontiguousarray(trace)
    assert q.dtype.itemsize == C.sizeof(Query)
    hs = (C.c_void_p * len(graphs))(*[g.h.value for g in graphs])
    t_admit = np.zeros(n, np.float64)
    t_done = np.zeros(n, np.float64)
    dig = np.zeros(n, np.uint64)
    bufs = []
    rp = None
    if results:
        ptrs = (C.c_void_p * n)()
        for i in range(n):
            g = graphs[int(q["graph"][i])]
            b = np.empty(g.V, np.float32 if int(q["kind"][i]) == PPR else np.uint32)
            bufs.append(b)
            ptrs[i] = b.ctypes.data
        rp = ptrs
    o = ServeOpts(int(sequential), ppr_iters, hbm_budget, use_graph, 0)
    st = ServeStats()
    _check(lib().ipr_serve(hs, len(graphs), q.ctypes.data, n, C.byref(o), t_admit.ctypes.data,
                           t_done.ctypes.data, dig.ctypes.data, rp, C.byref(st)), "serve")
    r = ServeResult(q["t_arrival"].copy(), t_admit, t_done, dig, st.wall_s, int(st.ticks),
                    int(st.admitted_waits), int(st.peak_slots), int(st.hbm_peak))
    return r, bufs
