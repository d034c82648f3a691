"""The batched and served configurations at full size (libipregel_batch.so):

- batched: the directed R-MAT scale 22 ef 16 seed 2 graph with weights in
  [1,255] (seed 4), 64 SSSP sources and 64 personalised-PageRank seeds drawn
  as the benchmark draws them (seed 4). Eight of the SSSP results are checked
  bit for bit against Dijkstra and eight PPR results against the FP64 oracle
  (L1 <= 2e-6, max relative error <= 1e-4 on entries above 1e-12, at least 95
  of the top 100 shared); every PPR query's agreement is computed from the
  same batch.
- served: one Poisson trace (500 queries/s, seed 5, 40/40/20 SSSP / PPR / CC)
  over the served configuration's resident graphs (R-MAT scales 18, 20, 22,
  23, 24 ef 16 seed 2 with weights seed 4, their symmetrised versions for CC,
  and the 4096x4096 grid with weights seed 3), served with continuous
  batching and by the sequential plain loop: every digest is the same in both
  modes and equals the digest of the returned result; a sample of each kind
  is checked against the oracle (Dijkstra, min-ancestor labels, FP64 PPR on
  the graphs of scale 22 and below, where it takes seconds).
The oracle's routines release the GIL, so the checks run in a thread pool.
"""
import concurrent.futures as cf

import numpy as np
import pytest

from ipregel_amd import batch as BA
from ipregel_amd import graph as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PPR_L1 = 2e-6
PPR_MAXREL = 1e-4
PPR_TOP = 95
POOL = 8


def check_ppr(x32, ref):
    m = BA.ppr_agreement(x32, ref)
    assert m["l1"] <= PPR_L1 and m["max_rel"] <= PPR_MAXREL and m["top_overlap"] >= PPR_TOP, m
    return m


@pytest.mark.timeout(900)
def test_batched_config_full_size():
    print("generating directed R-MAT scale 22 with weights (seed 4)", flush=True)
    g = G.rmat(22, 16, seed=2).with_weights(4)
    rng = np.random.default_rng(4)
    src = rng.integers(0, g.V, 64).astype(np.uint32)
    seeds = rng.integers(0, g.V, 64).astype(np.uint32)
    rg = BA.ResidentGraph(g)
    try:
        print("64 SSSP queries in one batch", flush=True)
        dist, _ = rg.batch_sssp(src)
        pick = list(range(0, 64, 8))
        with cf.ThreadPoolExecutor(POOL) as ex:
            refs = list(ex.map(lambda q: O.dijkstra(g, int(src[q])), pick))
        for q, ref in zip(pick, refs):
            assert np.array_equal(dist[q], ref), f"sssp query {q} source {src[q]}"
        del dist
        print("64 PPR queries in one batch", flush=True)
        x, res, _ = rg.batch_ppr(seeds, 20)
        assert np.isfinite(x).all()
        assert np.allclose(x.astype(np.float64).sum(axis=1), 1.0, atol=1e-5)
        with cf.ThreadPoolExecutor(POOL) as ex:
            prefs = list(ex.map(lambda q: O.ppr(g, int(seeds[q]), 20), pick))
        for q, ref in zip(pick, prefs):
            m = check_ppr(x[q], ref)
            print(f"ppr query {q} seed {seeds[q]}: {m}", flush=True)
    finally:
        rg.close()


@pytest.fixture(scope="module")
def served():
    graphs, names = [], []

    def add(name, g):
        print(f"resident {name} V={g.V} E={g.E}", flush=True)
        graphs.append((g, BA.ResidentGraph(g)))
        names.append(name)

    for sc in (18, 20, 22, 23, 24):
        add(f"rmat{sc}", G.rmat(sc, 16, seed=2).with_weights(4))
    add("grid4096", G.grid(4096, 4096).with_weights(3))
    for sc in (18, 20, 22, 23, 24):
        add(f"rmat{sc}sym", G.rmat(sc, 16, seed=2, symmetrise=True))
    by_kind = {BA.SSSP: [0, 1, 2, 3, 4, 5], BA.PPR: [0, 1, 2, 3, 4, 5], BA.CC: [6, 7, 8, 9, 10, 5]}
    yield graphs, names, by_kind
    for _, r in graphs:
        r.close()


@pytest.mark.timeout(900)
def test_served_trace_full_size(served):
    graphs, names, by_kind = served
    res = [r for _, r in graphs]
    V = [r.V for r in res]
    tr = BA.poisson_trace(60, 500.0, {BA.SSSP: 0.4, BA.PPR: 0.4, BA.CC: 0.2}, by_kind, V, seed=5)
    print("serving the trace with continuous batching", flush=True)
    rb, bufs = BA.serve(res, tr, hbm_budget=192 << 30, results=True)
    print("serving the trace with the sequential plain loop", flush=True)
    rs, _ = BA.serve(res, tr, sequential=True)
    assert np.array_equal(rb.digest, rs.digest)
    for i in range(tr.size):
        assert BA.digest(bufs[i]) == int(rb.digest[i]), f"query {i}"
    # a sample of each kind: the first query of the kind on each graph
    first = {}
    for i in range(tr.size):
        first.setdefault((int(tr["kind"][i]), int(tr["graph"][i])), i)
    jobs = []
    for (kind, gi), i in sorted(first.items()):
        g = graphs[gi][0]
        p = int(tr["param"][i])
        if kind == BA.SSSP:
            jobs.append((i, "sssp", lambda g=g, p=p: O.dijkstra(g, p)))
        elif kind == BA.CC:
            jobs.append((i, "cc", lambda g=g: O.min_ancestor(g)))
        elif kind == BA.PPR and g.E <= 70_000_000 and not names[gi].startswith("grid"):
            jobs.append((i, "ppr", lambda g=g, p=p: O.ppr(g, p, 20)))
    kinds = {k for _, k, _ in jobs}
    assert {"sssp", "cc", "ppr"} <= kinds, kinds
    print(f"checking {len(jobs)} results against the oracle", flush=True)
    with cf.ThreadPoolExecutor(POOL) as ex:
        refs = list(ex.map(lambda j: j[2](), jobs))
    for (i, kind, _), ref in zip(jobs, refs):
        what = f"query {i} {kind} on {names[int(tr['graph'][i])]} param {int(tr['param'][i])}"
        if kind == "ppr":
            check_ppr(bufs[i], ref)
        else:
            assert np.array_equal(bufs[i], ref), what
        print(f"ok {what}", flush=True)
