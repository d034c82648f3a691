"""The oracle pinned: against the reference's own binaries (golden fixtures),
against closed forms (min-ancestor labels, BFS, Dijkstra) and against a dense
matrix restatement of the PageRank recurrence."""
import numpy as np
import pytest

from conftest import GOLDEN, golden_runs
from ipregel_amd import graph as G
from oracle import oracle as O

APP = {"pagerank": O.PR, "cc": O.CC, "sssp": O.SSSP}


def parse_bin(binname):
    """reference binary name -> (app, variant, directed); a "_directed" infix
    is benchmarks/pagerank.c built with directed=true (oracle/ref.mk)"""
    app = binname.split("_")[0]
    variant = binname[len(app):-3]
    directed = variant.startswith("_directed")
    if directed:
        variant = variant[len("_directed"):]
    return APP[app], O.VARIANTS[variant], directed


@pytest.mark.parametrize("case,binname", golden_runs())
def test_oracle_matches_reference_binary(case, binname, tmp_path):
    g = G.read(GOLDEN / case)
    app, variant, directed = parse_bin(binname)
    r = O.run(g, app, variant, directed=directed, source=0)
    out = tmp_path / "o.txt"
    O.dump(app, r.values, str(out))
    assert out.read_text() == (GOLDEN / f"{case}__{binname}.out").read_text()
    trace = [int(x) for x in (GOLDEN / f"{case}__{binname}.trace").read_text().split()]
    assert list(r.active[: r.supersteps]) == trace


@pytest.mark.parametrize("case,binname", golden_runs(wide=True))
def test_reference_64bit_build_matches_32bit(case, binname):
    """The reference's _64 builds on the graph written with 64-bit ids print
    what its _32 builds print on the 32-bit copy (so the _64 drop-in runs are
    pinned by the same oracle)."""
    narrow = case[: -len("_w64")]
    wide = np.fromfile(GOLDEN / f"{case}.adj", dtype=np.uint64)
    assert np.array_equal(wide, np.fromfile(GOLDEN / f"{narrow}.adj", dtype=np.uint32).astype(np.uint64))
    for ext in (".idx", ".config"):
        assert (GOLDEN / f"{case}{ext}").read_bytes() == (GOLDEN / f"{narrow}{ext}").read_bytes()
    b32 = binname[: -len("_64")] + "_32"
    for ext in (".out", ".trace"):
        assert (GOLDEN / f"{case}__{binname}{ext}").read_text() == (GOLDEN / f"{narrow}__{b32}{ext}").read_text()


@pytest.mark.parametrize("seed", [3, 4])
@pytest.mark.parametrize("variant", [O.PUSH, O.PULL, O.SPREAD, O.SPREAD_PULL])
def test_cc_is_component_minimum(seed, variant):
    g = G.rmat(11, 8, seed=seed, symmetrise=True)
    r = O.run(g, O.CC, variant)
    assert np.array_equal(r.values, O.min_ancestor(g))


def test_min_ancestor_closed_form_small():
    # 0 -> 1 -> 2, 3 -> 2: labels = min ancestor
    g = G.from_arcs(4, [0, 1, 3], [1, 2, 2])
    assert O.min_ancestor(g).tolist() == [0, 0, 0, 3]


@pytest.mark.parametrize("variant", [O.PUSH, O.SPREAD])
def test_sssp_is_bfs(variant):
    g = G.rmat(11, 16, seed=5)
    for s in (0, 7):
        r = O.run(g, O.SSSP, variant, source=s)
        assert np.array_equal(r.values, O.bfs(g, s))


def test_bfs_closed_form_chain():
    g = G.from_arcs(5, [0, 1, 2, 3], [1, 2, 3, 4])
    assert O.bfs(g, 1).tolist() == [2**32 - 1, 0, 1, 2, 3]


@pytest.mark.parametrize("variant", [O.PUSH, O.SPREAD])
def test_weighted_sssp_is_dijkstra(variant):
    g = G.rmat(11, 16, seed=6).with_weights(3)
    r = O.run(g, O.SSSP_W, variant, source=0)
    assert np.array_equal(r.values, O.dijkstra(g, 0))
    gg = G.grid(20, 30).with_weights(3)
    r = O.run(gg, O.SSSP_W, variant, source=31)
    assert np.array_equal(r.values, O.dijkstra(gg, 31))


def test_dijkstra_closed_form():
    g = G.from_arcs(4, [0, 0, 1, 2], [1, 2, 3, 3])
    g.wgt = np.array([5, 1, 1, 10], np.uint32)  # arcs in CSR order: 0->1, 0->2, 1->3, 2->3
    assert O.dijkstra(g, 0).tolist() == [0, 5, 1, 6]


def dense_pagerank(g, rounds=10, transpose=False):
    """x_s = 0.15/V + 0.85 * A (x_{s-1}/outdeg), A[v, w] = 1 for w in adj(v):
    pull over the vertex's own list, as the reference does with directed=false.
    transpose=True uses A^T (v sums over the u with v in adj(u)): PageRank of
    the directed graph, what a push of x/outdeg along the out-arcs computes
    and what a pull over mirrored in-lists (directed=true) computes."""
    V = g.V
    deg = np.diff(g.idx).astype(np.float64)
    A = np.zeros((V, V))
    for v in range(V):
        A[v, g.adj[g.idx[v]:g.idx[v + 1]]] = 1.0
    if transpose:
        A = A.T.copy()
    inv = np.where(deg > 0, 1.0 / np.where(deg > 0, deg, 1), 0.0)
    x = np.full(V, 1.0 / V)
    for _ in range(rounds):
        x = 0.15 / V + 0.85 * (A @ (x * inv))
    return x


def test_dense_pagerank_transpose_on_a_hand_case():
    # 0 -> 1, 0 -> 2, 1 -> 2: with A^T vertex 2 gathers x0/2 + x1/1
    g = G.from_arcs(3, [0, 0, 1], [1, 2, 2])
    x = dense_pagerank(g, rounds=1, transpose=True)
    v = 1.0 / 3
    assert x[0] == 0.15 / 3
    assert abs(x[1] - (0.15 / 3 + 0.85 * v / 2)) < 1e-15
    assert abs(x[2] - (0.15 / 3 + 0.85 * (v / 2 + v))) < 1e-15


def test_push_pagerank_matches_dense_transposed_recurrence():
    """The oracle's push variant (messages x/outdeg along the out-arcs, summed
    in program order) is PageRank of G: the A^T recurrence."""
    g = G.rmat(9, 16, seed=7)
    r = O.run(g, O.PR, O.PUSH)
    np.testing.assert_allclose(r.values, dense_pagerank(g, transpose=True), rtol=1e-12, atol=0)
    assert r.supersteps == 11


def test_directed_pull_pagerank_matches_dense_transposed_recurrence():
    """With directed=true the pull reads mirrored in-lists: the same A^T
    recurrence as the push, and a result that differs from directed=false."""
    g = G.rmat(9, 16, seed=7)
    r = O.run(g, O.PR, O.PULL, directed=True)
    ref = dense_pagerank(g, transpose=True)
    np.testing.assert_allclose(r.values, ref, rtol=1e-12, atol=0)
    assert not np.allclose(O.run(g, O.PR, O.PULL).values, ref, rtol=1e-6)


def test_pagerank_matches_dense_recurrence():
    g = G.rmat(9, 16, seed=7)
    r = O.run(g, O.PR, O.PULL)
    ref = dense_pagerank(g)
    np.testing.assert_allclose(r.values, ref, rtol=1e-12, atol=0)
    assert r.supersteps == 11
    assert list(r.active[:11]) == [g.V] * 10 + [0]


def test_pagerank_no_inneighbours_gets_teleport_only():
    # vertex 2 has an empty list: no message, value = 0.15/V after superstep 1
    g = G.from_arcs(3, [0, 1], [1, 0])
    r = O.run(g, O.PR, O.PULL)
    assert r.values[2] == 0.15 / 3


@pytest.mark.parametrize("variant", [O.PUSH, O.SPREAD])
def test_bounded_sample_is_a_prefix_of_the_run(variant):
    # the CPU baseline's bounded samples: the first k supersteps of the run,
    # with the same per-superstep active counts
    g = G.rmat(12, 8, seed=3).with_weights(3)
    src = g.max_degree_vertex()
    full = O.run(g, O.SSSP_W, variant, source=src)
    assert full.supersteps > 3
    part = O.run(g, O.SSSP_W, variant, source=src, stop_after=3)
    assert part.supersteps == 3
    assert np.array_equal(part.active, full.active[:3])
    assert 0 < part.work_edges <= full.work_edges
    timed = O.run(g, O.SSSP_W, variant, source=src, stop_seconds=1e-9)
    assert timed.supersteps == 1
    assert np.array_equal(timed.active, full.active[:1])
