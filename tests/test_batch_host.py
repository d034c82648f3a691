"""CPU checks of the batched/serving host side: the ctypes layouts match the
C structs, the Poisson trace generator, the digest and the PPR agreement
metrics (no GPU calls)."""
import ctypes as C

import numpy as np

from ipregel_amd import batch as BA
from oracle import oracle as O
from ipregel_amd import graph as G


def test_struct_sizes():
    assert C.sizeof(BA.Query) == 24
    assert C.sizeof(BA.ServeOpts) == 24
    assert C.sizeof(BA.ServeStats) == 40
    assert C.sizeof(BA.BatchStats) == 24


def test_poisson_trace_shape_and_rate():
    V = [100, 200, 300]
    q = BA.poisson_trace(4000, 500.0, {BA.SSSP: 0.4, BA.PPR: 0.4, BA.CC: 0.2},
                         {BA.SSSP: [0, 1], BA.PPR: [0, 1], BA.CC: [2]}, V, seed=5)
    assert q.dtype.itemsize == 24
    assert q["t_arrival"][0] == 0.0 and np.all(np.diff(q["t_arrival"]) >= 0)
    rate = (q.size - 1) / q["t_arrival"][-1]
    assert 450 < rate < 550
    frac = np.bincount(q["kind"], minlength=3) / q.size
    assert abs(frac[0] - 0.4) < 0.03 and abs(frac[2] - 0.2) < 0.03
    assert set(q["graph"][q["kind"] == BA.CC]) == {2}
    assert np.all(q["param"] < np.array(V)[q["graph"]])
    q2 = BA.poisson_trace(4000, 500.0, {BA.SSSP: 0.4, BA.PPR: 0.4, BA.CC: 0.2},
                          {BA.SSSP: [0, 1], BA.PPR: [0, 1], BA.CC: [2]}, V, seed=5)
    assert np.array_equal(q, q2)


def test_digest_is_positional_and_type_blind():
    a = np.arange(10, dtype=np.uint32)
    b = a.copy()
    b[[3, 4]] = b[[4, 3]]
    assert BA.digest(a) != BA.digest(b)
    f = np.linspace(0, 1, 10).astype(np.float32)
    assert BA.digest(f) == BA.digest(f.view(np.uint32))
    # additive over positions: the sum of per-vertex terms
    assert (BA.digest(a[:1]) + 0) == BA.digest(a[:1])


def test_ppr_agreement_metrics():
    x64 = np.array([0.5, 0.3, 0.2, 0.0])
    m = BA.ppr_agreement(x64.astype(np.float32), x64, top=3)
    assert m["top_overlap"] == 3 and m["l1"] < 1e-7 and m["max_rel"] < 1e-7
    m2 = BA.ppr_agreement(np.array([0.2, 0.3, 0.5, 0.0], np.float32), x64, top=1)
    assert m2["top_overlap"] == 0 and abs(m2["l1"] - 0.6) < 1e-7


def test_oracle_ppr_closed_forms():
    """The FP64 PPR restatement: total mass 1, dangling mass returns to the
    seed, and on a directed cycle it is the closed form of a restarted walk."""
    n = 7
    src = np.arange(n)
    g = G.from_arcs(n, src, (src + 1) % n)
    x = O.ppr(g, 0, iters=200)
    assert abs(x.sum() - 1.0) < 1e-12
    a = 0.85
    want = (1 - a) * a ** np.arange(n) / (1 - a ** n)
    assert np.allclose(x, want, rtol=1e-12, atol=1e-15)
    # a chain into a sink: the sink's mass is sent back to the seed
    g2 = G.from_arcs(3, [0, 1], [1, 2])
    y = O.ppr(g2, 0, iters=500)
    # stationary: y0 = 0.15 + 0.85*y2, y1 = 0.85*y0, y2 = 0.85*y1
    y0 = 0.15 / (1 - 0.85 ** 3)
    assert np.allclose(y, [y0, 0.85 * y0, 0.85 ** 2 * y0], rtol=1e-12)
    # one iteration from a seed with two out-arcs
    g3 = G.from_arcs(3, [0, 0], [1, 2])
    z = O.ppr(g3, 0, iters=1)
    assert np.allclose(z, [0.15, 0.425, 0.425], rtol=1e-15)


def test_oracle_ppr_many_is_ppr_bit_for_bit():
    """The multi-seed FP64 PPR (eight seeds per cache line, OpenMP over
    vertices) gives every seed's ppr bits, with a repeated seed, a sink and
    a seed count that is not a multiple of eight."""
    g = G.rmat(10, 16, seed=5)
    seeds = np.array([0, 3, 3, 700, 1023, 17, 512, 9, 100, 44, 2], np.uint32)
    many = O.ppr_many(g, seeds, 20)
    for i, s in enumerate(seeds):
        assert np.array_equal(many[i].view(np.uint64), O.ppr(g, int(s), 20).view(np.uint64)), i


def test_autotune_scan_stops_and_picks_the_fastest():
    """A one-dimensional scan keeps the fastest candidate and stops at the
    first one 1.5x slower than the best so far (later ones are not timed)."""
    from ipregel_amd import autotune as AT
    times = {1: 10.0, 2: 8.0, 4: 9.0, 8: 12.5, 16: 1.0}
    timed = []

    def measure(v):
        timed.append(v)
        return times[v]

    best, seen = AT._scan([1, 2, 4, 8, 16], measure)
    assert best == 2 and timed == [1, 2, 4, 8] and set(seen) == {"1", "2", "4", "8"}
    best, seen = AT._scan([16, 1], lambda v: times[v])
    assert best == 16 and list(seen) == ["16", "1"]


def test_autotune_best_of_repeats_short_runs():
    from ipregel_amd import autotune as AT
    calls = []

    def run(ts):
        it = iter(ts)
        return lambda: (calls.append(1), next(it))[1]

    assert AT._best_of(run([100.0, 80.0])) == 80.0 and len(calls) == 2
    calls.clear()
    assert AT._best_of(run([900.0, 1.0])) == 900.0 and len(calls) == 1
