"""GPU parity: the HIP engine through the C-ABI application libraries against
the reference binaries' golden outputs and against the oracle.
Bit-exact for CC/SSSP (labels, distances, ip_dump text, active-count trace);
FP64 PageRank bit-exact in ordered mode and within 1e-12 relative per vertex
in the reassociating mode."""
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, golden_runs
from ipregel_amd import engine as EN
from ipregel_amd import graph as G
from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]

pytestmark = pytest.mark.gpu

BUILT = {p.name[len("libipr_"):-len("_32.so")] for p in EN.LIB.glob("libipr_*_32.so")}
PR_RTOL = 1e-12


def ours(binname):
    return binname[:-3]


def engine_name_for(binname):
    name = ours(binname)
    assert name in BUILT, f"{name} missing from {sorted(BUILT)}"
    return name


@pytest.mark.parametrize("case,binname", golden_runs())
def test_golden_text_and_trace(case, binname, tmp_path):
    name = engine_name_for(binname)
    g = G.read(GOLDEN / case)
    out = tmp_path / "out.txt"
    print(f"case={case} app={name}", flush=True)
    r = EN.run_app(name, g, source=0, out_path=str(out), pull_ordered=1)
    assert out.read_text() == (GOLDEN / f"{case}__{binname}.out").read_text()
    trace = [int(x) for x in (GOLDEN / f"{case}__{binname}.trace").read_text().split()]
    assert list(r.active) == trace
    assert r.supersteps == len(trace)


@pytest.mark.parametrize("case", ["rmat10", "rmat8", "star65", "chain10"])
def test_pagerank_reassociated_within_tolerance(case):
    g = G.read(GOLDEN / case)
    ref = O.run(g, O.PR, O.PULL).values
    r = EN.run_app("pagerank_single_broadcast", g, pull_ordered=2, pull_small_max=2,
                   pull_medium_max=16)
    rel = np.abs(r.values - ref) / np.maximum(np.abs(ref), 1e-300)
    assert rel.max() <= PR_RTOL, rel.max()


@pytest.mark.parametrize("case", ["rmat10", "rmat8", "star65", "chain10", "rmat14"])
def test_push_pagerank_within_tolerance(case):
    """The push PageRank build (pagerank_32: messages combined by FP64 atomic
    adds, in arrival order) against the oracle's push variant, which combines
    them in program order: within 1e-12 relative per vertex, active counts
    exact. (The reference's own push build truncates the messages to integers
    in its compare-and-swap, so its outputs are not the comparison.)"""
    g = G.rmat(14, 16, seed=7) if case == "rmat14" else G.read(GOLDEN / case)
    ref = O.run(g, O.PR, O.PUSH)
    r = EN.run_app("pagerank", g)
    rel = np.abs(r.values - ref.values) / np.maximum(np.abs(ref.values), 1e-300)
    assert rel.max() <= PR_RTOL, rel.max()
    assert list(r.active) == list(ref.active[: ref.supersteps])


def test_pagerank_scale16_config():
    """PageRank, 10 supersteps, FP64, pull, R-MAT scale 16 ef 16 seed 1."""
    g = G.rmat(16, 16, seed=1)
    ref = O.run(g, O.PR, O.PULL)
    r1 = EN.run_app("pagerank_single_broadcast", g, pull_ordered=1)
    assert np.array_equal(r1.values.view(np.uint64), ref.values.view(np.uint64))
    assert list(r1.active) == list(ref.active[: ref.supersteps])
    r2 = EN.run_app("pagerank_single_broadcast", g, pull_ordered=2)
    rel = np.abs(r2.values - ref.values) / np.abs(ref.values)
    assert rel.max() <= PR_RTOL, rel.max()
    assert r2.work_edges == ref.work_edges == 10 * g.E


@pytest.mark.parametrize("name", ["cc_spread", "cc", "cc_single_broadcast", "cc_generic_spread",
                                  "cc_single_broadcast_spread"])
def test_cc_medium(name):
    g = G.rmat(15, 16, seed=2, symmetrise=True)
    ref = O.run(g, O.CC, O.VARIANTS[name[2:].replace("_generic", "")])
    r = EN.run_app(name, g)
    assert np.array_equal(r.values, ref.values)
    assert np.array_equal(r.values, O.min_ancestor(g))
    assert list(r.active) == list(ref.active[: ref.supersteps])
    if name in ("cc_spread", "cc", "cc_generic_spread"):
        assert r.work_edges == ref.work_edges


@pytest.mark.parametrize("name", ["sssp_spread", "sssp", "sssp_single_broadcast",
                                  "sssp_single_broadcast_spread"])
def test_sssp_medium(name):
    g = G.rmat(15, 16, seed=3)
    src = g.max_degree_vertex()
    ref = O.run(g, O.SSSP, O.VARIANTS[name[4:]], source=src)
    r = EN.run_app(name, g, source=src)
    assert np.array_equal(r.values, ref.values)
    assert list(r.active) == list(ref.active[: ref.supersteps])


@pytest.mark.parametrize("name", ["sssp_weighted_spread", "sssp_weighted"])
def test_weighted_sssp_medium(name):
    g = G.rmat(14, 16, seed=3).with_weights(3)
    ref = O.dijkstra(g, 0)
    r = EN.run_app(name, g, source=0)
    assert np.array_equal(r.values, ref)
    gg = G.grid(64, 64).with_weights(3)
    r = EN.run_app(name, gg, source=65)
    assert np.array_equal(r.values, O.dijkstra(gg, 65))


@pytest.mark.parametrize("tune", [
    dict(use_graph=2), dict(max_chunk=1), dict(bin_small_max=1, bin_medium_max=2),
    dict(bin_small_max=4096, bin_medium_max=8192), dict(grid_blocks=7), dict(dir_opt=2),
    dict(dir_opt=2, bin_small_max=1, bin_medium_max=2)])
def test_cc_tuning_does_not_change_results(tune):
    g = G.rmat(14, 16, seed=9, symmetrise=True)
    ref = O.min_ancestor(g)
    r = EN.run_app("cc_spread", g, **tune)
    assert np.array_equal(r.values, ref)


@pytest.mark.parametrize("mode", ["0", "1", "2"])
def test_bitmap_launch_modes(mode, monkeypatch):
    """The broadcasters' bitmap kernel launched never (0: every gather that
    asks for the bitmap finds none built for its superstep and reads every
    record), always (1) or while runs use it (2, the default: CC's runs
    build none, so the later runs of its session leave the launch out).
    Three runs per session, each checked."""
    monkeypatch.setenv("IPR_GATHER_LIVE_LAUNCH", mode)
    monkeypatch.setenv("IPR_GATHER_LIVE", "100")
    monkeypatch.setenv("IPR_GATHER_LIVE_VERTICES", "100")
    g = G.rmat(14, 16, seed=3)
    src = int(np.argmax(np.diff(g.idx)))
    ref = O.bfs(g, src)
    s = EN.Session("sssp_spread", g, source=src, dir_alpha=1 << 30)
    for _ in range(3):
        st = s.run(1)
        assert np.array_equal(s.values(), ref)
    assert st.gathered > 0
    s.close()
    monkeypatch.setenv("IPR_GATHER_LIVE", "60")
    monkeypatch.setenv("IPR_GATHER_LIVE_VERTICES", "5")
    gs = G.rmat(14, 16, seed=9, symmetrise=True)
    ref = O.min_ancestor(gs)
    s = EN.Session("cc_spread", gs)
    for _ in range(3):
        s.run(1)
        assert np.array_equal(s.values(), ref)
    s.close()


@pytest.mark.parametrize("dense_flags,order", [("1", "1"), ("1", "0"), ("0", "1")])
def test_frontier_left_as_flags(dense_flags, order, monkeypatch):
    """After a superstep that gathered, the compaction only counts the raised
    flags, and the next superstep computes all V filtered by them and finds
    its broadcasters by a sweep of the records (by position with records in
    out-degree order, by vertex id without): CC, unit and weighted SSSP,
    where gathered and pushed supersteps alternate; the list-only mode (0)
    beside them."""
    monkeypatch.setenv("IPR_DENSE_FLAGS", dense_flags)
    monkeypatch.setenv("IPR_RECORD_ORDER", order)
    gs = G.rmat(14, 16, seed=9, symmetrise=True)
    r = EN.run_app("cc_spread", gs)
    assert np.array_equal(r.values, O.min_ancestor(gs))
    ref = O.run(gs, O.CC, O.VARIANTS["_spread"])
    assert list(r.active) == list(ref.active[: ref.supersteps])
    assert 0 < r.gathered < r.supersteps
    g = G.rmat(14, 16, seed=3)
    src = int(np.argmax(np.diff(g.idx)))
    r = EN.run_app("sssp_spread", g, source=src)
    assert np.array_equal(r.values, O.bfs(g, src))
    gw = g.with_weights(3)
    r = EN.run_app("sssp_weighted_spread", gw, source=src)
    assert np.array_equal(r.values, O.dijkstra(gw, src))


@pytest.mark.parametrize("plan", ["1", "2", "0"])
def test_launch_plan(plan, monkeypatch):
    """A run that repeats the previous one launches in each superstep only
    the delivery kernel that superstep used, gather or expansion (1, the
    default); 2 inverts the recorded plan, so every planned run stops at its
    first superstep and runs again without it; 0 launches both. Three runs
    per session, each checked, with gathered and pushed supersteps mixed."""
    monkeypatch.setenv("IPR_LAUNCH_PLAN", plan)
    gs = G.rmat(14, 16, seed=9, symmetrise=True)
    ref = O.min_ancestor(gs)
    s = EN.Session("cc_spread", gs)
    for _ in range(3):
        st = s.run(1)
        assert np.array_equal(s.values(), ref)
    assert 0 < st.gathered < st.supersteps
    s.close()
    g = G.rmat(14, 16, seed=3).with_weights(3)
    src = int(np.argmax(np.diff(g.idx)))
    ref = O.dijkstra(g, src)
    s = EN.Session("sssp_weighted_spread", g, source=src)
    for _ in range(3):
        s.run(1)
        assert np.array_equal(s.values(), ref)
    s.close()


@pytest.mark.parametrize("large_min,chunk", [(2, 64), (40, 100), (300, 64)])
def test_large_broadcasters_push(large_min, chunk):
    """Every superstep pushed (no gather), broadcasters above large_min
    out-arcs split into chunks of `chunk` arcs: they register in the 64
    counter slots and the expansion finds each chunk's slot and entry.
    Weighted SSSP, so each chunk also carries its arcs' weights."""
    g = G.rmat(14, 16, seed=3).with_weights(3)
    src = int(np.argmax(np.diff(g.idx)))
    r = EN.run_app("sssp_weighted_spread", g, source=src, dir_opt=2, bin_small_max=chunk,
                   bin_medium_max=large_min)
    assert np.array_equal(r.values, O.dijkstra(g, src))
    gs = G.rmat(14, 16, seed=9, symmetrise=True)
    r = EN.run_app("cc_spread", gs, dir_opt=2, bin_small_max=chunk, bin_medium_max=large_min)
    assert np.array_equal(r.values, O.min_ancestor(gs))


@pytest.mark.parametrize("blocks,order,live,ranges,record4", [
    ("1", "1", "60", "2048", "1"), ("3", "0", "0", "100000", "1"), ("8192", "1", "100", "512", "1"),
    ("8192", "0", "60", "4000000", "1"), ("", "1", "0", "", "1"), ("", "1", "60", "", "0"),
    ("5", "0", "100", "1024", "0")])
def test_gather_every_superstep(blocks, order, live, ranges, record4, monkeypatch):
    """Dense-superstep gather forced on every superstep (dir_alpha huge), with
    1, 3, 8192 or the default number of workgroups, and arc ranges of one
    pass (512 arcs) up to the whole graph, taken from the range counters:
    in-lists split across lanes, passes and ranges, hubs, isolated vertices;
    symmetric and directed transposes,
    with and without weights; broadcast records by vertex id or by degree;
    the broadcasters' bitmap off, below 60% live arcs, or always; 4-byte or
    8-byte broadcast records. The receiver-blocked gather is off here
    (IPR_GATHER_TILED=0); test_tiled_gather_every_superstep covers it."""
    monkeypatch.setenv("IPR_GATHER_TILED", "0")
    monkeypatch.setenv("IPR_GATHER_RECORD4", record4)
    for k, v in (("IPR_GATHER_BLOCKS", blocks), ("IPR_GATHER_RANGE_ARCS", ranges)):
        if v:
            monkeypatch.setenv(k, v)
        else:
            monkeypatch.delenv(k, raising=False)
    monkeypatch.setenv("IPR_RECORD_ORDER", order)
    monkeypatch.setenv("IPR_GATHER_LIVE", live)
    monkeypatch.setenv("IPR_GATHER_LIVE_VERTICES", "100")
    g = G.rmat(13, 16, seed=5, symmetrise=True)
    r = EN.run_app("cc_spread", g, dir_alpha=1 << 30)
    assert np.array_equal(r.values, O.min_ancestor(g))
    assert r.gathered > 0
    gd = G.rmat(12, 16, seed=6)
    r = EN.run_app("sssp_spread", gd, source=0, dir_alpha=1 << 30)
    assert np.array_equal(r.values, O.bfs(gd, 0))
    gw = gd.with_weights(2)
    r = EN.run_app("sssp_weighted_spread", gw, source=0, dir_alpha=1 << 30)
    assert np.array_equal(r.values, O.dijkstra(gw, 0))
    assert r.gathered > 0
    star = G.from_arcs(3100, np.r_[np.arange(1, 3000), np.zeros(5, int)],
                       np.r_[np.full(2999, 2999), np.arange(3000, 3005)], keep_self_loops=False)
    r = EN.run_app("cc_spread", star, dir_alpha=1 << 30)
    assert np.array_equal(r.values, O.min_ancestor(star))


@pytest.mark.parametrize("blocks,recv,arcs,order,record4,live", [
    ("", "", "", "1", "1", ""), ("1", "", "", "1", "1", "100"), ("3", "32", "100", "0", "1", ""),
    ("", "64", "1000", "1", "0", "100"), ("7", "16384", "5000", "1", "1", "0"), ("2", "33", "65536", "0", "0", ""),
    ("", "64", "1000", "1", "1", "100")])
def test_tiled_gather_every_superstep(blocks, recv, arcs, order, record4, live, monkeypatch):
    """The receiver-blocked gather (tiles of consecutive receivers, arcs sorted
    by source record, folds in LDS) forced on every superstep, with 1, 2, 3,
    7 or the default number of resident workgroups; blocks of 32, 33, 64,
    16384 or the default number of receivers; tiles of 100 arcs up to the
    default, so that hub in-lists split into many tiles that deliver by
    atomics; records by id or by degree, 4- or 8-byte; the broadcasters'
    bitmap by default, off, or read in every gathered superstep; SPREAD and
    PUSH, with and without weights; a hub star; sparse graphs of 160-200K vertices
    whose small blocks have slot groups spanning more than 16 bits of record
    positions (the escape rows: 155-294 groups with 32-64 receivers per block
    and 100-1000 arcs per tile, by a count over the same layout on the CPU)."""
    monkeypatch.setenv("IPR_GATHER_TILED", "1")
    monkeypatch.setenv("IPR_GATHER_RECORD4", record4)
    monkeypatch.setenv("IPR_RECORD_ORDER", order)
    for k, v in (("IPR_TILE_BLOCKS", blocks), ("IPR_TILE_RECV", recv), ("IPR_TILE_ARCS", arcs),
                 ("IPR_GATHER_LIVE", live), ("IPR_GATHER_LIVE_VERTICES", "100" if live else "")):
        if v:
            monkeypatch.setenv(k, v)
        else:
            monkeypatch.delenv(k, raising=False)
    g = G.rmat(13, 16, seed=5, symmetrise=True)
    for name in ("cc_spread", "cc"):
        print(f"app={name} rmat13 tiles {blocks}/{recv}/{arcs}", flush=True)
        r = EN.run_app(name, g, dir_alpha=1 << 30)
        assert np.array_equal(r.values, O.min_ancestor(g))
        assert r.gathered > 0
    gd = G.rmat(12, 16, seed=6)
    r = EN.run_app("sssp_spread", gd, source=0, dir_alpha=1 << 30)
    assert np.array_equal(r.values, O.bfs(gd, 0))
    gw = gd.with_weights(2)
    for name in ("sssp_weighted_spread", "sssp_weighted"):
        print(f"app={name} rmat12 tiles {blocks}/{recv}/{arcs}", flush=True)
        r = EN.run_app(name, gw, source=0, dir_alpha=1 << 30)
        assert np.array_equal(r.values, O.dijkstra(gw, 0))
        assert r.gathered > 0
    star = G.from_arcs(3100, np.r_[np.arange(1, 3000), np.zeros(5, int)],
                       np.r_[np.full(2999, 2999), np.arange(3000, 3005)], keep_self_loops=False)
    ref = O.run(star, O.CC, O.SPREAD)
    r = EN.run_app("cc_spread", star, dir_alpha=1 << 30)
    assert np.array_equal(r.values, ref.values)
    assert list(r.active) == list(ref.active[: ref.supersteps])
    rng = np.random.default_rng(1)
    gr = G.from_arcs(200000, rng.integers(0, 200000, 1000000), rng.integers(0, 200000, 1000000))
    print(f"app=cc_spread random 200K tiles {blocks}/{recv}/{arcs}", flush=True)
    r = EN.run_app("cc_spread", gr, dir_alpha=1 << 30)
    assert np.array_equal(r.values, O.min_ancestor(gr))
    grw = gr.with_weights(3)
    r = EN.run_app("sssp_weighted_spread", grw, source=0, dir_alpha=1 << 30)
    assert np.array_equal(r.values, O.dijkstra(grw, 0))
    gg = G.grid(400, 400)
    print(f"app=sssp_spread grid400 tiles {blocks}/{recv}/{arcs}", flush=True)
    r = EN.run_app("sssp_spread", gg, source=0, dir_alpha=1 << 30)
    assert np.array_equal(r.values, O.bfs(gg, 0))


@pytest.mark.parametrize("blocks,ranges", [("", ""), ("3", "512"), ("1", "100000")])
def test_gather_split_by_record_line(blocks, ranges, monkeypatch):
    """IPR_GATHER_XCD=1: the gather's arcs split into 8 classes by the line of
    their source's record, workgroups starting on their own class and then
    helping with the others (with 1 or 3 workgroups, fewer than the classes,
    every class is still taken); CC, BFS and weighted SSSP with the gather
    forced on every superstep, and a hub star."""
    monkeypatch.setenv("IPR_GATHER_XCD", "1")
    for k, v in (("IPR_GATHER_BLOCKS", blocks), ("IPR_GATHER_RANGE_ARCS", ranges)):
        if v:
            monkeypatch.setenv(k, v)
        else:
            monkeypatch.delenv(k, raising=False)
    g = G.rmat(13, 16, seed=5, symmetrise=True)
    r = EN.run_app("cc_spread", g, dir_alpha=1 << 30)
    assert np.array_equal(r.values, O.min_ancestor(g))
    assert r.gathered > 0
    gd = G.rmat(12, 16, seed=6)
    r = EN.run_app("sssp_spread", gd, source=0, dir_alpha=1 << 30)
    assert np.array_equal(r.values, O.bfs(gd, 0))
    gw = gd.with_weights(2)
    r = EN.run_app("sssp_weighted", gw, source=0, dir_alpha=1 << 30)
    assert np.array_equal(r.values, O.dijkstra(gw, 0))
    star = G.from_arcs(3100, np.r_[np.arange(1, 3000), np.zeros(5, int)],
                       np.r_[np.full(2999, 2999), np.arange(3000, 3005)], keep_self_loops=False)
    r = EN.run_app("cc_spread", star, dir_alpha=1 << 30)
    assert np.array_equal(r.values, O.min_ancestor(star))


@pytest.mark.parametrize("rounds", ["1", "2", "4", "8"])
def test_frontier_compaction_rounds(rounds, monkeypatch):
    """The frontier compaction with 1, 2, 4 and 8 rounds of 4096 flags per
    workgroup (the engine picks the count from V; small graphs get 1): CC on
    R-MAT 16 (dense and sparse frontiers over 2-16 workgroups) and BFS on a
    333x301 grid (V not a multiple of 16, 629 supersteps), with the
    active-count traces against the oracle's."""
    monkeypatch.setenv("IPR_COMPACT_ROUNDS", rounds)
    g = G.rmat(16, 16, seed=2, symmetrise=True)
    ref = O.run(g, O.CC, O.SPREAD)
    r = EN.run_app("cc_spread", g)
    assert np.array_equal(r.values, ref.values)
    assert list(r.active) == list(ref.active[: ref.supersteps])
    gg = G.grid(333, 301)
    r = EN.run_app("sssp_spread", gg, source=5)
    assert np.array_equal(r.values, O.bfs(gg, 5))
    ref = O.run(gg, O.SSSP, O.SPREAD, source=5)
    assert list(r.active) == list(ref.active[: ref.supersteps])


@pytest.mark.parametrize("record4", ["1", "0"])
def test_gather_stamp_window_and_escapes(record4, monkeypatch):
    """The gather forced on every superstep for more supersteps than the
    4-byte records' stamp window (63): CC and BFS on a 40x40 grid (78
    supersteps) in the SPREAD and PUSH variants; weighted SSSP whose distances
    pass 2^25, so its messages escape to the 8-byte records; with the 4-byte
    records and without them."""
    monkeypatch.setenv("IPR_GATHER_RECORD4", record4)
    gg = G.grid(40, 40)
    for name in ("cc_spread", "cc"):
        print(f"app={name} grid40", flush=True)
        r = EN.run_app(name, gg, dir_alpha=1 << 30)
        assert np.array_equal(r.values, O.min_ancestor(gg))
        assert r.supersteps > 63 and r.gathered > 63
    for name in ("sssp_spread", "sssp"):
        print(f"app={name} grid40", flush=True)
        r = EN.run_app(name, gg, source=0, dir_alpha=1 << 30)
        assert np.array_equal(r.values, O.bfs(gg, 0))
        assert r.gathered > 63
    gw = G.rmat(12, 16, seed=6).with_weights(2, 1 << 24, 1 << 26)
    ref = O.dijkstra(gw, 0)
    assert ref[ref != 2**32 - 1].max() >= 1 << 25
    for name in ("sssp_weighted_spread", "sssp_weighted"):
        print(f"app={name} rmat12 wide weights", flush=True)
        r = EN.run_app(name, gw, source=0, dir_alpha=1 << 30)
        assert np.array_equal(r.values, ref)
        assert r.gathered > 0


@pytest.mark.parametrize("fused,blocks,maxdeg,alpha", [
    ("1", "", "", 0), ("0", "", "", 0), ("1", "1", "", 0), ("1", "3", "", 0),
    ("1", "", "4", 0), ("1", "2", "2", 40), ("1", "", "", 40)])
def test_fused_sparse_supersteps(fused, blocks, maxdeg, alpha, monkeypatch):
    """Sparse supersteps of the spread variant as one fused kernel (frontier
    from the list or from the raised flags, compute, expansion, the next
    frontier counted as receivers are activated, the last workgroup ending the
    superstep), against the general kernels and the oracle: values and
    per-superstep active counts. With 1 or 3 workgroups (many scan passes and
    tiles each), with a tiny out-degree limit or a low gather threshold
    (supersteps alternate between the fused and the general kernels, the fused
    one stopping itself when its superstep turns dense and the host rebuilding
    the list from the flags), on CC (R-MAT 14), BFS and weighted SSSP on a
    57x61 grid (a hundred sparse supersteps, past the 4-byte records'
    63-superstep stamp window)."""
    monkeypatch.setenv("IPR_FUSED", fused)
    for k, v in (("IPR_FUSED_BLOCKS", blocks), ("IPR_FUSED_MAX_DEG", maxdeg)):
        if v:
            monkeypatch.setenv(k, v)
        else:
            monkeypatch.delenv(k, raising=False)
    tune = {"dir_alpha": alpha} if alpha else {}
    g = G.rmat(14, 16, seed=9, symmetrise=True)
    ref = O.run(g, O.CC, O.SPREAD)
    print(f"cc_spread rmat14 fused={fused} blocks={blocks} maxdeg={maxdeg} alpha={alpha}", flush=True)
    # one superstep per host check, so that the fused-or-general choice is
    # made before every superstep of this short run
    r = EN.run_app("cc_spread", g, max_chunk=1, **tune)
    assert np.array_equal(r.values, ref.values)
    assert list(r.active) == list(ref.active[: ref.supersteps])
    if fused == "0":
        assert r.fused_steps == 0
    elif not maxdeg and not alpha:
        assert 0 < r.fused_steps < r.supersteps
    gg = G.grid(57, 61)
    ref = O.run(gg, O.SSSP, O.SPREAD, source=3)
    print("sssp_spread grid57x61", flush=True)
    r = EN.run_app("sssp_spread", gg, source=3, **tune)
    assert np.array_equal(r.values, ref.values)
    assert list(r.active) == list(ref.active[: ref.supersteps])
    if fused == "1" and not maxdeg:
        assert r.fused_steps > 63
    gw = gg.with_weights(5)
    ref = O.run(gw, O.SSSP_W, O.SPREAD, source=3)
    print("sssp_weighted_spread grid57x61", flush=True)
    r = EN.run_app("sssp_weighted_spread", gw, source=3, **tune)
    assert np.array_equal(r.values, ref.values)
    assert np.array_equal(r.values, O.dijkstra(gw, 3))
    assert list(r.active) == list(ref.active[: ref.supersteps])


@pytest.mark.parametrize("fin,first", [("0", ""), ("1", ""), ("1", "1"), ("0", "64")])
def test_end_of_superstep_in_kernel_or_launch(fin, first, monkeypatch):
    """The end of a superstep done by the last workgroup of the pull kernel or
    of the frontier compaction (IPR_FIN_IN_KERNEL=1, the default) or by its
    own launch (0), with first chunks of 1, 16/4 (the defaults) or 64
    supersteps: PageRank ordered (bit-exact), CC and BFS with their traces."""
    monkeypatch.setenv("IPR_FIN_IN_KERNEL", fin)
    if first:
        monkeypatch.setenv("IPR_FIRST_CHUNK", first)
    else:
        monkeypatch.delenv("IPR_FIRST_CHUNK", raising=False)
    g = G.rmat(12, 16, seed=1)
    ref = O.run(g, O.PR, O.PULL)
    r = EN.run_app("pagerank_single_broadcast", g, pull_ordered=1)
    assert np.array_equal(r.values.view(np.uint64), ref.values.view(np.uint64))
    assert list(r.active) == list(ref.active[: ref.supersteps])
    gs = G.rmat(13, 16, seed=4, symmetrise=True)
    ref = O.run(gs, O.CC, O.SPREAD)
    r = EN.run_app("cc_spread", gs)
    assert np.array_equal(r.values, ref.values)
    assert list(r.active) == list(ref.active[: ref.supersteps])
    gg = G.grid(31, 47)
    ref = O.run(gg, O.SSSP, O.SPREAD, source=40)
    r = EN.run_app("sssp_spread", gg, source=40)
    assert np.array_equal(r.values, ref.values)
    assert list(r.active) == list(ref.active[: ref.supersteps])


def test_fused_sparse_supersteps_with_send_message(tmp_path, monkeypatch):
    """cc_send_spread (ip_send_message to each out-neighbour) with the fused
    kernel on and off: the reference cc_spread binary's output and trace."""
    import subprocess
    for fused in ("1", "0"):
        for case in ("rmat10", "grid8x8", "chain10"):
            env = dict(__import__("os").environ, IPR_FUSED=fused)
            root = tmp_path / case
            (tmp_path / f"{case}.config").write_text((GOLDEN / f"{case}.config").read_text())
            (tmp_path / f"{case}.adj").write_bytes((GOLDEN / f"{case}.adj").read_bytes())
            np.fromfile(GOLDEN / f"{case}.idx", dtype=np.uint64).astype(np.uint32).tofile(tmp_path / f"{case}.idx")
            out = tmp_path / "o.txt"
            print(f"cc_send_spread_32 {case} fused={fused}", flush=True)
            p = subprocess.run([str(ROOT / "ipregel_amd" / "bin" / "cc_send_spread_32"), str(root), str(out),
                                "1", "static", "0"], capture_output=True, text=True, timeout=120, env=env)
            assert p.returncode == 0, p.stdout[-2000:]
            assert out.read_text() == (GOLDEN / f"{case}__cc_spread_32.out").read_text()
            trace = [int(x.split(":")[1]) for x in p.stdout.splitlines() if "ActiveVertexCount" in x]
            assert trace == [int(x) for x in (GOLDEN / f"{case}__cc_spread_32.trace").read_text().split()]


def test_tune_engine_keeps_results():
    from ipregel_amd import autotune as AT
    g = G.rmat(12, 16, seed=5, symmetrise=True)
    rep = AT.tune_engine("cc_spread", g, repeat=1)
    assert {"bin_small_max", "bin_medium_max", "grid_blocks", "dir_alpha"} <= set(rep["choice"])
    assert len(rep["tried"]) == len(AT.ENGINE_GRID["push"]) + len(AT.ENGINE_GRID["grid_blocks"]) \
        + len(AT.ENGINE_GRID["dir_alpha"])
    r = EN.run_app("cc_spread", g, **rep["choice"])
    assert np.array_equal(r.values, O.min_ancestor(g))


@pytest.mark.parametrize("binname", ["cc_spread_32", "sssp_32", "pagerank_single_broadcast_32"])
def test_phase_profile_table(binname, tmp_path):
    """IPR_PHASE_PROFILE=1 (as IP_ENABLE_THREAD_PROFILING at build time) adds
    the per-superstep phase table to the log; the dump and the active counts
    do not change, and the tables' arcs add up to WorkEdges."""
    import os
    import re
    import subprocess
    case = "rmat10"
    exe = ROOT / "ipregel_amd" / "bin" / binname
    out = tmp_path / "out.txt"
    args = [str(exe), str(GOLDEN / case), str(out), "1", "static", "0"] + (["0"] if binname.startswith("sssp") else [])
    p = subprocess.run(args, capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, IPR_PHASE_PROFILE="1"))
    assert p.returncode == 0, p.stdout[-2000:]
    assert out.read_text() == (GOLDEN / f"{case}__{binname}.out").read_text()
    trace = [int(x.split(":")[1]) for x in p.stdout.splitlines() if "ActiveVertexCount" in x]
    assert trace == [int(x) for x in (GOLDEN / f"{case}__{binname}.trace").read_text().split()]
    rows = {k: [float(v) for v in re.findall(rf"\|\s+{k} \|\s+([0-9.]+)", p.stdout)]
            for k in ("Compute", "Merging", "Mailbox", "Total", "Arcs")}
    assert len(rows["Total"]) == len(trace) > 0
    for c, mg, mb, t in zip(rows["Compute"], rows["Merging"], rows["Mailbox"], rows["Total"]):
        assert min(c, mg, mb) >= 0 and abs(c + mg + mb - t) < 3e-6
    assert all(t > 0 for t in rows["Total"]) and sum(rows["Compute"]) > 0
    work = int(re.search(r"WorkEdges:(\d+)", p.stdout).group(1))
    assert sum(rows["Arcs"]) == work


def test_repeat_runs_identical():
    g = G.rmat(13, 16, seed=4)
    r = EN.run_app("sssp_spread", g, source=0, repeat=3)
    assert r.runs == 3
    assert np.array_equal(r.values, O.bfs(g, 0))


@pytest.mark.parametrize("adapt", ["1", "0"])
def test_adaptive_first_chunk_keeps_results(adapt, monkeypatch):
    """Runs that repeat a run of the same graph start with a chunk of exactly
    the launches it needed (IPR_ADAPT_CHUNK=1, the default): PageRank ordered
    (bit-exact, 12 launches for 11 supersteps), CC and push SSSP give the
    oracle's values and active counts on every run, as with the rule off."""
    monkeypatch.setenv("IPR_ADAPT_CHUNK", adapt)
    cases = [("pagerank_single_broadcast", G.rmat(12, 16, seed=1), O.PR, O.PULL, {"pull_ordered": 1}),
             ("cc_spread", G.rmat(13, 16, seed=4, symmetrise=True), O.CC, O.SPREAD, {}),
             ("sssp", G.rmat(12, 8, seed=5), O.SSSP, O.PUSH, {})]
    for app, g, oapp, ovar, opts in cases:
        ref = O.run(g, oapp, ovar, source=0)
        s = EN.Session(app, g, source=0, **opts)
        try:
            for k in range(3):
                print(f"{app} adapt={adapt} run {k}", flush=True)
                s.run(1)
                v = s.values()
                if oapp == O.PR:
                    assert np.array_equal(v.view(np.uint64), ref.values.view(np.uint64))
                else:
                    assert np.array_equal(v, ref.values)
                active, _, _ = s.trace()
                assert list(active) == list(ref.active[: ref.supersteps])
        finally:
            s.close()


def test_empty_and_edgeless_graphs():
    g = G.from_arcs(5, [], [])
    r = EN.run_app("cc_spread", g)
    assert r.values.tolist() == [0, 1, 2, 3, 4]
    r = EN.run_app("sssp_spread", g, source=2)
    assert r.values.tolist() == [2**32 - 1, 2**32 - 1, 0, 2**32 - 1, 2**32 - 1]
