"""Benchmark of the MI355X iPregel engine on one GPU.

A step is one complete superstep loop (ip_run: every superstep until the
engine halts) of an application over its graph, which stays resident in HBM
across steps. Six workloads are measured; the headline line is the first,
hashmin connected components (cc_spread: push with an active frontier, min
combiner) on a symmetrised R-MAT graph, scale 22, edge factor 16, a/b/c/d =
.57/.19/.19/.05, seed 2:

  cc22   cc_spread, symmetrised R-MAT 22 (seed 2)
  pr16   pagerank_single_broadcast FP64, R-MAT 16 (seed 1), 10 supersteps
  ss24a  sssp_spread unit weights, R-MAT 24 (seed 3), source = highest out-degree
  ss24b  sssp_weighted_spread, the same graph with weights in [1,255] (seed 3)
  grid   sssp_weighted_spread, 4096x4096 grid, weights in [1,255] (seed 3), source 4097
  cc24   cc_spread, symmetrised R-MAT 24 (seed 2)

For each:
  GTEPS      W / t, W = sum over supersteps of the out-degrees of the vertices
             that broadcast (push) or E per message-bearing superstep (pull),
             deterministic under BSP; t = device time of the timed runs (HIP
             events on the engine's stream)
  roofline   the dominant kernel: the gather of the dense supersteps (8 B per
             arc relaxed in them, 12 B with weights), the pull kernel for
             PageRank (12 B per in-arc + 40 B per vertex), or, without a dense
             superstep, the fused sparse kernel (arcs as above + 24 B per
             computed vertex) or the expansion; over its HIP-event-timed launch
             durations (plain launches), against the 8 TB/s HBM peak of MI355X.
             traffic = the rocprofv3-measured HBM bytes per launch where a
             profile exists (profiles/)
  cpu_baseline  the reference's own OpenMP build (oracle/_ref, compiled from
             the reference sources) on the same graph file with the host's
             threads (OMP_NUM_THREADS, else the CPUs this process may use), run
             again until about 10 s of supersteps (at most 12 runs), the median
             of its "Total time of supersteps"; its first run's dump is
             compared with the GPU's results. The reference has no weighted
             application: ss24b and the grid take the oracle's single-thread
             C restatement on a bounded sample (its supersteps up to the first
             boundary after 12 s), kind "port".

Two more sections go into the same JSON line (libipregel_batch.so):

  batched    64 SSSP queries (sources uniform, seed 4, weights [1,255]) and 64
             personalised-PageRank queries (restart 0.15, 20 iterations) on the
             directed R-MAT scale 22 ef 16 seed 2 graph, each as one 64-lane
             batch, against the plain loop (the same queries one at a time on
             the one-query kernels); results checked identical; device time by
             HIP events. Every PPR query's L1 / max-rel / top-100 agreement
             with the FP64 oracle (all 64, computed by the oracle's multi-seed
             restatement on the host's threads, which is also the section's
             cpu_baseline).
  served     Poisson arrivals (seed 5) at several rates of a 40/40/20 SSSP /
             PPR / CC mix over six resident graphs (R-MAT scales 18, 20, 22,
             23, 24, ef 16, seed 2, weights seed 4; CC on their symmetrised
             versions; and the 4096x4096 grid, weights seed 3), served with
             continuous batching under an HBM budget for the batch states and,
             on the same trace, by the sequential plain loop; p50/p99 latency,
             queries/s, and whether every result digest matched.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--apps cc22,pr16,...]
                       [--no-cpu] [--no-batched] [--no-served] [--serve-rates 10,100,500,2000]
       python bench.py --copy-probe MiB   (a plain device copy of that size, to
                       calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters)
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from ipregel_amd import autotune as AT  # noqa: E402
from ipregel_amd import batch as BA  # noqa: E402
from ipregel_amd import engine as EN  # noqa: E402
from ipregel_amd import graph as G  # noqa: E402

METRIC = "GTEPS & % HBM peak per app; served p50/p99 + queries/s vs plain loop, 1 GPU"
HBM_PEAK_GBS = 8000.0
REF_BIN = ROOT / "oracle" / "_ref" / "bin"
DEFAULT_APPS = "cc22,pr16,ss24a,ss24b,grid,cc24"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads() -> int:
    """The host threads a CPU baseline uses: OMP_NUM_THREADS when set (the
    share of the machine this job is given), else the CPUs it may run on."""
    cpus = len(os.sched_getaffinity(0))
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(n, cpus) if n > 0 else cpus


# graphs, built once per process (the two SSSP workloads share theirs)
_GRAPHS: dict[str, G.Graph] = {}


def graph(key: str) -> G.Graph:
    if key not in _GRAPHS:
        make = {
            "rmat22sym": lambda: G.rmat(22, 16, seed=2, symmetrise=True),
            "rmat16": lambda: G.rmat(16, 16, seed=1),
            "rmat24": lambda: G.rmat(24, 16, seed=3),
            "rmat24w": lambda: graph("rmat24").with_weights(3),
            "grid4096w": lambda: G.grid(4096, 4096).with_weights(3),
            "rmat24sym": lambda: G.rmat(24, 16, seed=2, symmetrise=True),
        }[key]
        _GRAPHS[key] = make()
    return _GRAPHS[key]


# workload -> app, graph, description, reference binary (None: port), source,
# engine options, step cap (the grid's run is 0.4 s)
WORKLOADS = {
    "cc22": dict(app="cc_spread", graph="rmat22sym", ref="cc_spread_32", source=0,
                 desc="cc_spread hashmin CC, symmetrised R-MAT scale 22 ef 16 a/b/c/d .57/.19/.19/.05 seed 2"),
    "pr16": dict(app="pagerank_single_broadcast", graph="rmat16", ref="pagerank_single_broadcast_32", source=0,
                 tuning={"pull_ordered": 2},
                 desc="pagerank_single_broadcast FP64 10 supersteps, R-MAT scale 16 ef 16 seed 1 (pull over own "
                      "list, directed=false), reassociating fold (per-vertex relative error <= 1e-12 vs the "
                      "in-order fold)"),
    "ss24a": dict(app="sssp_spread", graph="rmat24", ref="sssp_spread_32", source="maxdeg",
                  desc="sssp_spread unit weights, R-MAT scale 24 ef 16 seed 3, source = highest out-degree vertex"),
    "ss24b": dict(app="sssp_weighted_spread", graph="rmat24w", ref=None, source="maxdeg",
                  desc="sssp_weighted_spread uint32 weights [1,255], R-MAT scale 24 ef 16 seed 3"),
    "grid": dict(app="sssp_weighted_spread", graph="grid4096w", ref=None, source=4097,
                 steps_cap=5, warmup_cap=1,
                 desc="sssp_weighted_spread 4096x4096 grid, weights [1,255] seed 3, source 4097"),
    "cc24": dict(app="cc_spread", graph="rmat24sym", ref="cc_spread_32", source=0,
                 desc="cc_spread hashmin CC, symmetrised R-MAT scale 24 ef 16 a/b/c/d .57/.19/.19/.05 seed 2"),
}

# HBM bytes per launch of each workload's dominant kernel, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE passes over `bench.py --apps <workload>` with every
# kernel launched directly (FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0, the
# factors a plain 1 GiB device copy gives; KiB, mean over the launches that do
# work: the dense gathers, the pull launches of supersteps 0-11, every fused
# sparse launch); commands and per-launch figures are in profiles/
TRAFFIC = {("cc22", "ipr_k_gather_tiled"): (382_816 * 2 + 17_428) * 1024,
           ("cc24", "ipr_k_gather_tiled"): (2_617_108 * 2 + 67_282) * 1024,
           ("ss24a", "ipr_k_gather_tiled"): (1_088_746 * 2 + 58_912) * 1024,
           ("ss24b", "ipr_k_gather_tiled"): (1_459_258 * 2 + 59_058) * 1024,
           ("cc22", "ipr_k_gather"): (767_103 * 2 + 56_781) * 1024,
           ("cc24", "ipr_k_gather"): (7_919_949 * 2 + 245_533) * 1024,
           ("ss24a", "ipr_k_gather"): (4_066_499 * 2 + 181_607) * 1024,
           ("ss24b", "ipr_k_gather"): (4_645_812 * 2 + 205_246) * 1024,
           ("pr16", "ipr_k_pull_step"): (5_975 * 2 + 1_385) * 1024,
           ("grid", "ipr_k_sparse_step"): (34_113 * 2 + 31_488) * 1024}


def model_bytes(app: str, g: G.Graph, W: int, F: int, supersteps_with_bcast: int) -> int:
    """Algorithmic bytes of a whole run: push 8 B/relaxed arc (+4 B weight) +
    28 B/computed vertex; pull PR 12 B/in-arc + 40 B/vertex per superstep."""
    if app.startswith("pagerank"):
        return supersteps_with_bcast * (12 * g.E + 40 * g.V)
    per_arc = 12 if "weighted" in app else 8
    return per_arc * W + 28 * F


def parse_dump_ints(path: str) -> np.ndarray | None:
    """the values of a "%u: %u" dump (benchmarks/cc.c:67, sssp.c:83), None
    unless its ids are 0..V-1 in order"""
    a = np.fromstring(Path(path).read_text().replace(":", " "), dtype=np.uint64, sep=" ")
    if a.size % 2 or not np.array_equal(a[0::2], np.arange(a.size // 2, dtype=np.uint64)):
        return None
    return a[1::2]


def pr_text(values: np.ndarray) -> str:
    """PageRank's ip_dump text (benchmarks/pagerank.c:75, "%lu: %0.20f")"""
    return "".join(f"{i}: {v:.20f}\n" for i, v in enumerate(values.tolist()))


def cpu_reference(binname: str, g: G.Graph, W: int, source: int, budget_s: float = 10.0,
                  max_runs: int = 12) -> tuple[dict, str | None, tempfile.TemporaryDirectory]:
    """The reference's own OpenMP binary on the same graph file, run again
    until its supersteps have taken about `budget_s` seconds (at most
    `max_runs` runs, each a full process: load, supersteps, dump); the median
    of its 'Total time of supersteps' is the baseline. Its first run writes
    its dump to a file (returned) for comparison with the GPU's results."""
    threads = host_threads()
    td = tempfile.TemporaryDirectory(dir="/tmp")
    root = os.path.join(td.name, "g")
    g.write(root)
    dump = os.path.join(td.name, "ref_dump.txt")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="TRUE", OMP_PLACES="cores")
    times, trace, t0 = [], None, time.time()
    while len(times) < max_runs and (not times or sum(times) < budget_s):
        args = [str(REF_BIN / binname), root, dump if not times else "/dev/null", str(threads), "static", "0"]
        if binname.startswith("sssp"):
            args.append(str(source))
        try:
            p = subprocess.run(args, env=env, capture_output=True, text=True, timeout=150)
        except subprocess.TimeoutExpired:
            return ({"value": None, "unit": "GTEPS", "cores": threads, "kind": "reference",
                     "sample": f"{binname} did not finish within 150 s"}, None, td)
        m = re.search(r"Total time of supersteps: ([0-9.]+)s", p.stdout)
        if p.returncode != 0 or not m:
            return ({"value": None, "unit": "GTEPS", "cores": threads, "kind": "reference",
                     "sample": f"{binname} failed rc={p.returncode}"}, None, td)
        times.append(float(m.group(1)))
        if trace is None:
            trace = [int(x.split(":")[1]) for x in p.stdout.splitlines() if "ActiveVertexCount" in x]
    wall = time.time() - t0
    t = float(np.median(times))
    out = {"value": W / t / 1e9, "unit": "GTEPS", "cores": threads, "kind": "reference",
           "sample": f"{len(times)} full runs of the reference {binname} (OpenMP, {threads} threads, static "
                     f"schedule) on the same graph file; median 'Total time of supersteps' {t:.3f} s (min "
                     f"{min(times):.3f}, max {max(times):.3f}; {sum(times):.1f} s of supersteps, {wall:.1f} s of "
                     f"process wall time)",
           "seconds": t, "runs": len(times), "trace": trace}
    return out, dump, td


def cpu_port(g: G.Graph, source: int, n_super: int, budget_s: float = 12.0) -> dict:
    """The oracle's single-thread C restatement of the spread variant running
    weighted SSSP, a bounded sample: its supersteps up to the first boundary
    after `budget_s` seconds; GTEPS by the sample's own W."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    r = O.run(g, O.SSSP_W, O.SPREAD, source=source, stop_seconds=budget_s, max_supersteps=1)
    dt = time.perf_counter() - t0
    whole = r.supersteps >= n_super
    return {"value": r.work_edges / dt / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port",
            "sample": f"the oracle's single-thread C restatement (oracle/ipr_oracle.c, spread variant, weighted "
                      f"SSSP), {'the whole run' if whole else f'its first {r.supersteps} of {n_super} supersteps'}"
                      f": W = {r.work_edges} relaxed arcs in {dt:.2f} s (graph conversion included)",
            "seconds": dt, "supersteps": r.supersteps, "W": r.work_edges}


def run_workload(key: str, steps: int, warmup: int, with_cpu: bool) -> dict:
    w = WORKLOADS[key]
    t0 = time.time()
    g = graph(w["graph"])
    app = w["app"]
    source = g.max_degree_vertex() if w["source"] == "maxdeg" else int(w["source"])
    tuning = w.get("tuning", {})
    steps = min(steps, w.get("steps_cap", steps))
    warmup = min(warmup, w.get("warmup_cap", warmup))
    log(f"{key}: graph V={g.V} E={g.E} ready in {time.time() - t0:.1f}s")
    sess = EN.Session(app, g, source=source, **tuning)
    log(f"{key}: resident on {sess.stats().device_name.decode()} (load {sess.stats().load_s:.2f}s)")
    sess.run(warmup)
    st = sess.stats()
    before_ms = st.total_ms
    # timed region: exactly `steps` runs, each bracketed by HIP events on the
    # engine's stream and followed by a stream synchronisation
    w0 = time.perf_counter()
    sess.run(steps)
    wall_ms = (time.perf_counter() - w0) * 1e3 / steps
    st = sess.stats()
    ms = (st.total_ms - before_ms) / steps
    W = int(st.work_edges)
    F = int(st.work_vertices)
    fused = int(st.fused_steps)
    active, sedges, sns = sess.trace()
    values = sess.values()
    n_super = int(st.supersteps)
    n_bcast_steps = int(np.count_nonzero(sedges))
    gteps = W / (ms * 1e-3) / 1e9
    step_bytes = model_bytes(app, g, W, F, n_bcast_steps)
    log(f"{key}: {n_super} supersteps ({fused} fused), W={W}, {ms:.3f} ms/step, {gteps:.2f} GTEPS")
    # the dominant kernel, timed by HIP events around each of its launches
    # (plain launches; per superstep: the gather or the pull kernel, then the
    # expansion or the fused sparse kernel)
    sess.kernel_events(True)
    k0_ms, k0_n = st.kernel_ms, st.kernel_launches
    ksteps = min(steps, 3)
    sess.run(ksteps)
    st2 = sess.stats()
    t_first, t_second = sess.kernel_times()
    sess.kernel_events(False)
    phase_ms = (st2.kernel_ms - k0_ms) / ksteps
    phase_launch = (st2.kernel_launches - k0_n) / ksteps
    per_arc = 12 if "weighted" in app else 8
    alpha = tuning.get("dir_alpha", EN.DIR_ALPHA_DEFAULT)
    gathered = [i for i in range(len(sedges)) if int(sedges[i]) * alpha > g.E]
    if app.startswith("pagerank"):
        # 4 B in-index + 8 B gathered record per in-arc; 8 B offset + 8 B
        # out-offset + 8 B value read + 8 B value write + 8 B record per vertex
        k_name, k_launch, k_ms = "ipr_k_pull_step", phase_launch, phase_ms
        k_bytes = n_bcast_steps * (12 * g.E + 40 * g.V)
    elif gathered and int(st.gathered) == len(gathered):
        k_name = "ipr_k_gather_tiled" if int(st.gather_tiles) else "ipr_k_gather"
        k_launch = len(gathered)
        k_ms = float(sum(t_first[i] for i in gathered))
        k_bytes = per_arc * int(sum(int(sedges[i]) for i in gathered))
    elif fused:
        # compute + expansion: arcs as above, 24 B per computed vertex
        # (frontier id, two offsets, value read and write, message)
        k_name, k_launch = "ipr_k_sparse_step", phase_launch
        k_ms = float(t_second.sum())
        k_bytes = per_arc * W + 24 * F
    else:
        k_name, k_launch = "ipr_k_expand", phase_launch
        k_ms = float(t_second.sum())
        k_bytes = per_arc * W + 4 * int(active.sum())
    achieved = k_bytes / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    traffic = TRAFFIC.get((key, k_name))
    ordered_values = None
    if app.startswith("pagerank"):
        # the bit-exact ordered fold, for the comparison with the reference's dump
        sess.close()
        so = EN.Session(app, g, source=source, pull_ordered=1)
        so.run(1)
        ordered_values = so.values()
        so.close()
    else:
        sess.close()
    out = {
        "app": app, "workload": w["desc"], "V": g.V, "E": g.E, "supersteps": n_super, "fused_supersteps": fused,
        "W": W, "vertices_computed": F, "steps": steps, "ms_per_step": ms, "wall_ms_per_step": wall_ms,
        "GTEPS": gteps, "model_bytes_per_step": step_bytes,
        "model_GBps": step_bytes / (ms * 1e-3) / 1e9,
        "hbm_frac_model": step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": k_name,
                     "avg_launch_ms": k_ms / k_launch if k_launch else None,
                     "algorithmic_bytes_per_step": k_bytes, "ms_per_step": k_ms,
                     "launches_per_step": k_launch},
        "kernel_events": {"delivery_phase_ms_per_step": phase_ms,
                          "per_superstep_ms": {"first": [round(float(x), 4) for x in t_first[:64]],
                                               "second": [round(float(x), 4) for x in t_second[:64]]},
                          "ms_per_step_plain_launches": (st2.total_ms - st.total_ms) / ksteps},
        "active_trace": [int(x) for x in active[:64]],
    }
    if with_cpu:
        if w["ref"]:
            cb, dump, td = cpu_reference(w["ref"], g, W, source)
            if cb.get("trace") is not None:
                cb["trace_matches_gpu"] = cb["trace"] == [int(x) for x in active]
                cb.pop("trace")
            if dump and os.path.exists(dump):
                if app.startswith("pagerank"):
                    ref_text = Path(dump).read_text()
                    cb["ref_dump_identical_ordered"] = pr_text(ordered_values) == ref_text
                    ref_vals = np.array([float(x.split(":")[1]) for x in ref_text.splitlines()])
                    rel = np.abs(values - ref_vals) / np.maximum(np.abs(ref_vals), 1e-300)
                    cb["timed_fold_max_rel_vs_ref_dump"] = float(rel.max())
                else:
                    ref_vals = parse_dump_ints(dump)
                    cb["ref_dump_identical"] = ref_vals is not None and bool(
                        np.array_equal(ref_vals, values.astype(np.uint64)))
            td.cleanup()
        else:
            cb = cpu_port(g, source, n_super)
        out["cpu_baseline"] = cb
        log(f"{key}: cpu baseline {cb}")
    return out


def run_batched(with_cpu: bool) -> dict:
    t0 = time.time()
    g = G.rmat(22, 16, seed=2).with_weights(4)
    rg = BA.ResidentGraph(g)
    rng = np.random.default_rng(4)
    src = rng.integers(0, g.V, 64).astype(np.uint32)
    seeds = rng.integers(0, g.V, 64).astype(np.uint32)
    log(f"batched: graph V={g.V} E={g.E} resident ({rg.bytes / 2**30:.2f} GiB) in {time.time() - t0:.1f}s")
    tune = AT.tune_resident(rg, kinds=("sssp",), mean_weight=float(g.wgt.mean()))
    log(f"batched: tuned {tune['final']} in {tune['seconds']:.1f}s")
    rg.batch_sssp(src[:2])
    rg.batch_ppr(seeds[:2], 2)
    dist, sb = rg.batch_sssp(src)
    plain_ms, same = 0.0, True
    for q in range(64):
        one, st = rg.sssp(int(src[q]))
        plain_ms += st["ms"]
        same &= bool(np.array_equal(one, dist[q]))
    del dist
    x, res, pb = rg.batch_ppr(seeds, 20)
    pplain_ms, psame = 0.0, True
    for q in range(64):
        one, r1, st = rg.ppr(int(seeds[q]), 20)
        pplain_ms += st["ms"]
        psame &= bool(np.array_equal(one.view(np.uint32), x[q].view(np.uint32))) and bool(r1 == res[q])
    E = g.E
    # pull iteration, 64 live slots: 4 B id + 4 B weight + 256 B row per in-arc, 256 B row written per vertex
    ppr_bytes_iter = E * (8 + 64 * 4) + g.V * 64 * 4
    # the plain loop's one-query pull iteration: 4 B id + 4 B weight + 4 B value per in-arc, 8 B per vertex
    plain_bytes_iter = E * 12 + g.V * 8
    plain_ms_iter = pplain_ms / 64 / 20
    out = {
        "workload": "directed R-MAT scale 22 ef 16 seed 2, weights [1,255] seed 4; 64 queries per kind "
                    "(seed 4)", "V": g.V, "E": E,
        "sssp": {"batch_ms": sb["ms"], "plain_loop_ms": plain_ms, "speedup": plain_ms / sb["ms"],
                 "queries_per_s": 64 / (sb["ms"] * 1e-3), "plain_queries_per_s": 64 / (plain_ms * 1e-3),
                 "supersteps": sb["supersteps"], "relaxations": sb["relaxations"],
                 "G_relaxations_per_s": sb["relaxations"] / (sb["ms"] * 1e-3) / 1e9,
                 "identical_to_plain": same},
        "ppr": {"batch_ms": pb["ms"], "plain_loop_ms": pplain_ms, "speedup": pplain_ms / pb["ms"],
                "queries_per_s": 64 / (pb["ms"] * 1e-3), "plain_queries_per_s": 64 / (pplain_ms * 1e-3),
                "iterations": 20, "ms_per_iteration": pb["ms"] / 20,
                "shared_bytes_per_iteration": ppr_bytes_iter,
                "shared_GBps": ppr_bytes_iter / (pb["ms"] / 20 * 1e-3) / 1e9,
                "hbm_frac_shared": ppr_bytes_iter / (pb["ms"] / 20 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "plain_ms_per_iteration": plain_ms_iter,
                "plain_bytes_per_iteration": plain_bytes_iter,
                "plain_hbm_frac": plain_bytes_iter / (plain_ms_iter * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "identical_to_plain": psame,
                "residual_max": float(res.max())},
        "mix_speedup": (plain_ms + pplain_ms) / (sb["ms"] + pb["ms"]),
        "tuning": tune["final"],
    }
    log(f"batched: sssp {sb['ms']:.1f} ms vs plain {plain_ms:.1f} ms; ppr {pb['ms']:.1f} ms vs "
        f"plain {pplain_ms:.1f} ms; identical {same} {psame}")
    if with_cpu:
        # every PPR query against the FP64 oracle (multi-seed, the host's threads)
        from oracle import oracle as O
        t1 = time.perf_counter()
        x64 = O.ppr_many(g, seeds, 20)
        dt = time.perf_counter() - t1
        acc = [BA.ppr_agreement(x[q], x64[q]) for q in range(64)]
        out["ppr"]["fp32_vs_fp64"] = {
            "l1": [a["l1"] for a in acc], "max_rel": [a["max_rel"] for a in acc],
            "top100_overlap": [a["top_overlap"] for a in acc],
            "worst": {"l1": max(a["l1"] for a in acc), "max_rel": max(a["max_rel"] for a in acc),
                      "top100_overlap": min(a["top_overlap"] for a in acc)}}
        out["cpu_baseline"] = {"value": 64 / dt, "unit": "queries/s", "cores": host_threads(), "kind": "port",
                               "sample": f"FP64 PPR restatement (oracle ppr_many, eight seeds per pass, "
                                         f"OpenMP over vertices), all 64 seeds x 20 iterations, {dt:.1f} s"}
        log(f"batched: PPR accuracy worst {out['ppr']['fp32_vs_fp64']['worst']}; cpu {dt:.1f}s")
    rg.close()
    return out


def build_served_graphs() -> tuple[list, list, dict]:
    """Six logical graphs; SSSP/PPR on the directed weighted ones (and the
    grid), CC on the symmetrised ones (and the grid)."""
    t0 = time.time()
    res, names, tuned = [], [], {}

    def add(name: str, g: G.Graph, kinds) -> None:
        rg = BA.ResidentGraph(g)
        mw = float(g.wgt.mean()) if g.wgt is not None and g.E else None
        rep = AT.tune_resident(rg, kinds=kinds, mean_weight=mw)
        tuned[name] = {k: rep[k]["choice"] for k in ("delta", "delta1", "large", "grid", "cc_mode", "sssp_states")
                       if k in rep}
        tuned[name]["seconds"] = round(rep["seconds"], 2)
        log(f"served: {name} V={g.V} E={g.E} tuned {tuned[name]}")
        res.append(rg)
        names.append(name)

    for sc in (18, 20, 22, 23, 24):
        add(f"rmat{sc}", G.rmat(sc, 16, seed=2).with_weights(4), ("sssp",))
    add("grid4096", G.grid(4096, 4096).with_weights(3), ("sssp", "cc"))
    for sc in (18, 20, 22, 23, 24):
        add(f"rmat{sc}sym", G.rmat(sc, 16, seed=2, symmetrise=True), ("cc",))
    by_kind = {BA.SSSP: [0, 1, 2, 3, 4, 5], BA.PPR: [0, 1, 2, 3, 4, 5], BA.CC: [6, 7, 8, 9, 10, 5]}
    log(f"served: 11 resident graphs ({sum(r.bytes for r in res) / 2**30:.1f} GiB) built and tuned in "
        f"{time.time() - t0:.1f}s")
    return res, names, by_kind, tuned


def run_served(rates: list[float], counts: list[int], budget: int) -> dict:
    res, names, by_kind, tuned = build_served_graphs()
    V = [r.V for r in res]
    mix = {BA.SSSP: 0.4, BA.PPR: 0.4, BA.CC: 0.2}
    warm = BA.poisson_trace(48, 2000.0, mix, by_kind, V, seed=99)
    BA.serve(res, warm, hbm_budget=budget)
    BA.serve(res, warm[:12], sequential=True)
    sweep = []
    for rate, n in zip(rates, counts):
        tr = BA.poisson_trace(n, rate, mix, by_kind, V, seed=5)
        rb, _ = BA.serve(res, tr, hbm_budget=budget)
        rs, _ = BA.serve(res, tr, sequential=True)
        b, s_ = rb.summary(tr, names), rs.summary(tr, names)
        row = {"rate": rate, "queries": n, "batched": b, "plain": s_,
               "throughput_ratio": b["qps"] / s_["qps"] if s_["qps"] else None,
               "identical": bool(np.array_equal(rb.digest, rs.digest))}
        sweep.append(row)
        log(f"served @ {rate} q/s: batched {b['qps']:.1f} q/s p50 {b['p50_ms']:.1f} p99 {b['p99_ms']:.1f} ms; "
            f"plain {s_['qps']:.1f} q/s p50 {s_['p50_ms']:.1f} p99 {s_['p99_ms']:.1f} ms; identical {row['identical']}")
    slot = {n: {"sssp": r.slot_bytes(BA.SSSP), "ppr": r.slot_bytes(BA.PPR), "cc": r.slot_bytes(BA.CC)}
            for n, r in zip(names, res)}
    for r in res:
        r.close()
    return {"graphs": names, "tuning": tuned,
            "mix": "40/40/20 SSSP/PPR/CC, graph uniform over six, parameter uniform",
            "hbm_budget_bytes": budget, "ppr_iterations": 20, "sweep": sweep, "state_bytes": slot}


def copy_probe(mib: int) -> None:
    """A device-to-device copy of `mib` MiB through libipregel_hip.so: under
    rocprofv3 --pmc FETCH_SIZE (or WRITE_SIZE) it gives the bytes those
    counters report per byte copied, the factor the gather's traffic is
    scaled by."""
    import ctypes as C
    L = C.CDLL(str(ROOT / "ipregel_amd" / "lib" / "libipregel_hip.so"))
    L.ipr_hip_init.argtypes = [C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_uint64)]
    L.ipr_hip_malloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.ipr_hip_memcpy_d2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    L.ipr_hip_fill.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
    L.ipr_hip_stream_create.argtypes = [C.POINTER(C.c_void_p)]
    L.ipr_hip_stream_sync.argtypes = [C.c_void_p]
    L.ipr_hip_free.argtypes = [C.c_void_p]
    name, hbm = C.create_string_buffer(128), C.c_uint64()
    if L.ipr_hip_init(0, name, 128, C.byref(hbm)):
        raise SystemExit("no device")
    n = mib << 20
    a, b, st = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if L.ipr_hip_malloc(C.byref(p), n):
            raise SystemExit("allocation failed")
    L.ipr_hip_stream_create(C.byref(st))
    L.ipr_hip_fill(a, 1, n, st)
    L.ipr_hip_memcpy_d2d(b, a, n, st)
    L.ipr_hip_stream_sync(st)
    L.ipr_hip_free(a)
    L.ipr_hip_free(b)
    print(json.dumps({"copy_probe_bytes": n, "device": name.value.decode()}))


def _plain(o):
    """numpy scalars in the JSON line as plain numbers"""
    if isinstance(o, np.generic):
        return o.item()
    raise TypeError(f"not serialisable: {type(o)}")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--apps", default=DEFAULT_APPS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--serve-rates", default="10,100,500,2000")
    ap.add_argument("--serve-counts", default="40,200,300,400")
    ap.add_argument("--serve-budget-gib", type=float, default=192.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MiB")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design "
                                   "(see the design notes), --gpus must be 1"}))
        return 2
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return 0
    keys = [k for k in a.apps.split(",") if k]
    results = {}
    for k in keys:
        results[k] = run_workload(k, a.steps, a.warmup, not a.no_cpu)
    _GRAPHS.clear()
    batched = None if a.no_batched else run_batched(not a.no_cpu)
    served = None
    if not a.no_served:
        rates = [float(x) for x in a.serve_rates.split(",") if x]
        counts = [int(x) for x in a.serve_counts.split(",") if x]
        served = run_served(rates, counts, int(a.serve_budget_gib * 2**30))
    head = results[keys[0]]
    line = {
        "metric": METRIC,
        "value": head["GTEPS"],
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32" if not head["app"].startswith("pagerank") else "f64",
        "data": "synthetic",
        "config": {"workload": head["workload"], "V": head["V"], "E": head["E"],
                   "supersteps": head["supersteps"], "W": head["W"]},
        "roofline": {x: head["roofline"][x] for x in ("bound", "achieved", "peak", "unit", "frac", "traffic",
                                                      "kernel", "avg_launch_ms")},
        "cpu_baseline": head.get("cpu_baseline"),
        "hbm_frac_model": head["hbm_frac_model"],
        "apps": {kk: {x: v for x, v in r.items() if x not in ("active_trace",)}
                 for kk, r in results.items()},
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line, default=_plain), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
