"""Python host side of the MI355X iPregel engine: loads an application library
(libipr_<app>_32.so = the app's translation unit + the C host driver) over
ctypes and drives it through the app's own main() with the harness entry
points of include/ipregel_driver.h (graph from host arrays, repeated runs,
values/trace readback). No compute happens in Python; a missing library is an
error, never a fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .graph import Graph

PKG = Path(__file__).resolve().parent
LIB = PKG / "lib"
# the driver's push-or-gather threshold when none is given: gather when the
# broadcasters' out-arcs times this exceed E
DIR_ALPHA_DEFAULT = 6
BIN = PKG / "bin"

# application name -> (value dtype, takes a source argument, needs weights)
APP_INFO = {
    "pagerank": (np.float64, False, False),
    "cc": (np.uint32, False, False),
    "sssp": (np.uint32, True, False),
    "sssp_weighted": (np.uint32, True, True),
}


def app_kind(name: str) -> str:
    for k in ("sssp_weighted", "pagerank", "sssp", "cc"):
        if name.startswith(k):
            return k
    raise ValueError(f"unknown application {name}")


class HarnessCfg(C.Structure):
    _fields_ = [
        ("V", C.c_uint64), ("E", C.c_uint64),
        ("idx", C.c_void_p), ("adj", C.c_void_p), ("wgt", C.c_void_p),
        ("skip_dump", C.c_int32), ("repeat", C.c_int32), ("quiet", C.c_int32), ("device", C.c_int32),
        ("stream", C.c_void_p),
        ("pull_ordered", C.c_int32), ("use_graph", C.c_int32), ("dir_opt", C.c_int32),
        ("dir_alpha", C.c_uint32), ("bin_small_max", C.c_uint32), ("bin_medium_max", C.c_uint32),
        ("grid_blocks", C.c_uint32), ("pull_small_max", C.c_uint32), ("pull_medium_max", C.c_uint32),
        ("max_chunk", C.c_uint32), ("defer_run", C.c_int32), ("kernel_events", C.c_int32),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("V", C.c_uint64), ("E", C.c_uint64), ("supersteps", C.c_uint64),
        ("work_edges", C.c_uint64), ("work_vertices", C.c_uint64), ("gathered", C.c_uint64),
        ("runs", C.c_uint64), ("launches", C.c_uint64),
        ("last_ms", C.c_double), ("total_ms", C.c_double), ("min_ms", C.c_double),
        ("device_ms", C.c_double), ("load_s", C.c_double),
        ("kernel_ms", C.c_double), ("kernel_launches", C.c_uint64),
        ("device_name", C.c_char * 128), ("fused_steps", C.c_uint64), ("gather_tiles", C.c_uint64),
    ]


@dataclass
class RunResult:
    values: np.ndarray
    supersteps: int
    work_edges: int
    work_vertices: int
    active: np.ndarray
    step_edges: np.ndarray
    step_ns: np.ndarray
    last_ms: float
    min_ms: float
    total_ms: float
    device_ms: float
    runs: int
    gathered: int
    launches: int
    device_name: str
    fused_steps: int = 0
    extra: dict = field(default_factory=dict)


_loaded: dict[str, C.CDLL] = {}


def load_app(name: str) -> C.CDLL:
    """Load libipr_<name>_32.so; raises if it is not built."""
    if name in _loaded:
        return _loaded[name]
    path = LIB / f"libipr_{name}_32.so"
    if not path.exists():
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
    L = C.CDLL(str(path), mode=os.RTLD_LOCAL | os.RTLD_NOW)
    L.main.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
    L.main.restype = C.c_int
    L.ipr_harness_configure.argtypes = [C.POINTER(HarnessCfg)]
    L.ipr_harness_configure.restype = None
    L.ipr_harness_stats.argtypes = [C.POINTER(RunStats)]
    L.ipr_harness_values.argtypes = [C.c_void_p, C.c_uint64]
    L.ipr_harness_trace.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
    L.ipr_harness_release.argtypes = []
    L.ipr_harness_release.restype = None
    L.ipr_harness_run.argtypes = [C.c_int]
    L.ipr_harness_stream.argtypes = []
    L.ipr_harness_stream.restype = C.c_void_p
    L.ipr_harness_kernel_events.argtypes = [C.c_int]
    L.ipr_harness_kernel_times.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
    _loaded[name] = L
    return L


def run_app(name: str, graph: Graph, *, source: int = 0, out_path: str | None = None,
            repeat: int = 1, quiet: bool = True, device: int = 0, stream: int | None = None,
            pull_ordered: int = 0, use_graph: int = 0, dir_opt: int = 0, dir_alpha: int = 0,
            bin_small_max: int = 0, bin_medium_max: int = 0, grid_blocks: int = 0,
            pull_small_max: int = 0, pull_medium_max: int = 0, max_chunk: int = 0,
            schedule: str = "static", release: bool = False) -> RunResult:
    """Run application `name` (e.g. "cc_spread") on `graph` through its main()."""
    kind = app_kind(name)
    dtype, takes_source, needs_w = APP_INFO[kind]
    if needs_w and graph.wgt is None:
        raise ValueError(f"{name} needs edge weights")
    L = load_app(name)
    idx = np.ascontiguousarray(graph.idx, dtype=np.uint64)
    adj = np.ascontiguousarray(graph.adj, dtype=np.uint32)
    wgt = np.ascontiguousarray(graph.wgt, dtype=np.uint32) if graph.wgt is not None else None
    cfg = HarnessCfg()
    cfg.V, cfg.E = graph.V, graph.E
    cfg.idx, cfg.adj = idx.ctypes.data, adj.ctypes.data
    cfg.wgt = wgt.ctypes.data if wgt is not None else None
    cfg.skip_dump = 0 if out_path else 1
    cfg.repeat = repeat
    cfg.quiet = int(quiet)
    cfg.device = device
    cfg.stream = stream
    cfg.pull_ordered, cfg.use_graph, cfg.dir_opt, cfg.dir_alpha = pull_ordered, use_graph, dir_opt, dir_alpha
    cfg.bin_small_max, cfg.bin_medium_max, cfg.grid_blocks = bin_small_max, bin_medium_max, grid_blocks
    cfg.pull_small_max, cfg.pull_medium_max, cfg.max_chunk = pull_small_max, pull_medium_max, max_chunk
    L.ipr_harness_configure(C.byref(cfg))
    argv = [f"{name}_32".encode(), b"harness", (out_path or "/dev/null").encode(), b"1",
            schedule.encode(), b"0"]
    if takes_source:
        argv.append(str(source).encode())
    arr = (C.c_char_p * len(argv))(*argv)
    rc = L.main(len(argv), arr)
    if rc != 0:
        raise RuntimeError(f"{name} main returned {rc}")
    st = RunStats()
    L.ipr_harness_stats(C.byref(st))
    values = np.empty(graph.V, dtype=dtype)
    if graph.V:
        r = L.ipr_harness_values(values.ctypes.data, values.nbytes)
        if r:
            raise RuntimeError(f"value readback failed ({r})")
    n = int(st.supersteps)
    active = np.zeros(n, np.uint64)
    sedges = np.zeros(n, np.uint64)
    sns = np.zeros(n, np.float64)
    if n:
        got = L.ipr_harness_trace(active.ctypes.data, sedges.ctypes.data, sns.ctypes.data, n)
        if got < 0:
            raise RuntimeError(f"trace readback failed ({got})")
    res = RunResult(values=values, supersteps=n, work_edges=int(st.work_edges),
                    work_vertices=int(st.work_vertices), active=active, step_edges=sedges,
                    step_ns=sns, last_ms=st.last_ms, min_ms=st.min_ms, total_ms=st.total_ms,
                    device_ms=st.device_ms, runs=int(st.runs), gathered=int(st.gathered),
                    launches=int(st.launches), device_name=st.device_name.decode(errors="replace"),
                    fused_steps=int(st.fused_steps))
    # the C driver keeps pointers to the host arrays until the next ip_init
    L.ipr_harness_configure(None)
    if release:
        L.ipr_harness_release()
    return res


def release_app(name: str) -> None:
    if name in _loaded:
        _loaded[name].ipr_harness_release()


class Session:
    """One application with its graph resident on the device: ip_init runs
    once (through main, with ip_run deferred); run(k) executes the superstep
    loop k more times. Keeps the host arrays alive while the driver holds
    pointers to them."""

    def __init__(self, name: str, graph: Graph, *, source: int = 0, kernel_events: bool = False,
                 **tuning):
        kind = app_kind(name)
        self.name = name
        self.dtype, takes_source, needs_w = APP_INFO[kind]
        if needs_w and graph.wgt is None:
            raise ValueError(f"{name} needs edge weights")
        self.L = load_app(name)
        self.graph = graph
        self.idx = np.ascontiguousarray(graph.idx, dtype=np.uint64)
        self.adj = np.ascontiguousarray(graph.adj, dtype=np.uint32)
        self.wgt = np.ascontiguousarray(graph.wgt, dtype=np.uint32) if graph.wgt is not None else None
        cfg = HarnessCfg()
        cfg.V, cfg.E = graph.V, graph.E
        cfg.idx, cfg.adj = self.idx.ctypes.data, self.adj.ctypes.data
        cfg.wgt = self.wgt.ctypes.data if self.wgt is not None else None
        cfg.skip_dump, cfg.repeat, cfg.quiet, cfg.defer_run = 1, 1, 1, 1
        cfg.kernel_events = int(kernel_events)
        for k, v in tuning.items():
            setattr(cfg, k, v)
        self.cfg = cfg
        self.L.ipr_harness_configure(C.byref(cfg))
        argv = [f"{name}_32".encode(), b"harness", b"/dev/null", b"1", b"static", b"0"]
        if takes_source:
            argv.append(str(source).encode())
        arr = (C.c_char_p * len(argv))(*argv)
        rc = self.L.main(len(argv), arr)
        if rc != 0:
            raise RuntimeError(f"{name} main returned {rc}")

    def run(self, repeat: int = 1) -> RunStats:
        rc = self.L.ipr_harness_run(repeat)
        if rc:
            raise RuntimeError(f"ipr_harness_run failed ({rc})")
        return self.stats()

    def stats(self) -> RunStats:
        st = RunStats()
        self.L.ipr_harness_stats(C.byref(st))
        return st

    def values(self) -> np.ndarray:
        out = np.empty(self.graph.V, dtype=self.dtype)
        if self.graph.V:
            rc = self.L.ipr_harness_values(out.ctypes.data, out.nbytes)
            if rc:
                raise RuntimeError(f"value readback failed ({rc})")
        return out

    def trace(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        n = int(self.stats().supersteps)
        a, e, t = np.zeros(n, np.uint64), np.zeros(n, np.uint64), np.zeros(n, np.float64)
        if n:
            self.L.ipr_harness_trace(a.ctypes.data, e.ctypes.data, t.ctypes.data, n)
        return a, e, t

    def kernel_times(self) -> tuple[np.ndarray, np.ndarray]:
        """With kernel events on: per superstep of the last run, milliseconds of
        the gather (or fused pull kernel) and of the expansion after it."""
        n = int(self.stats().supersteps)
        g, x = np.zeros(n, np.float64), np.zeros(n, np.float64)
        got = self.L.ipr_harness_kernel_times(g.ctypes.data, x.ctypes.data, n) if n else 0
        if got < 0:
            raise RuntimeError("kernel events were not on")
        return g[:got], x[:got]

    def kernel_events(self, on: bool) -> None:
        if self.L.ipr_harness_kernel_events(int(on)):
            raise RuntimeError("could not create kernel events")

    def stream(self) -> int:
        return int(self.L.ipr_harness_stream() or 0)

    def close(self) -> None:
        self.L.ipr_harness_release()
        self.L.ipr_harness_configure(None)
