"""Parity at the benchmark configurations' full sizes: the drop-in engine's
labels and distances against the oracle's closed forms (min-ancestor labels,
BFS, Dijkstra), bit for bit, on the symmetrised R-MAT scale 22 CC graph, the
directed R-MAT scale 24 SSSP graph (unit weights and weights in [1,255]) and
the 4096x4096 weighted grid; and the engine's ip_dump text against the
SHA-256 of the reference binaries' own dumps of the same configurations
(tests/golden/digests.json, written by tests/golden/make_digests.py), with
their per-superstep active counts. The graphs take tens of seconds to
generate on the host, so each is built once."""
import hashlib
import json

import numpy as np
import pytest

from conftest import GOLDEN
from ipregel_amd import engine as EN
from ipregel_amd import graph as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DIGESTS = json.loads((GOLDEN / "digests.json").read_text())


@pytest.fixture(scope="module")
def rmat24():
    print("generating R-MAT scale 24 (directed, seed 3)", flush=True)
    return G.rmat(24, 16, seed=3)


def sha256_of(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for block in iter(lambda: f.read(1 << 24), b""):
            h.update(block)
    return h.hexdigest()


# digest key -> (engine application, graph, options): PageRank in the ordered
# fold, which is bit-exact with the reference's in-list order
DUMP_CASES = {
    "pr16": ("pagerank_single_broadcast", lambda r24: G.rmat(16, 16, seed=1), {"pull_ordered": 1}),
    "pr16_directed": ("pagerank_directed_single_broadcast", lambda r24: G.rmat(16, 16, seed=1),
                      {"pull_ordered": 1}),
    "cc22": ("cc_spread", lambda r24: G.rmat(22, 16, seed=2, symmetrise=True), {}),
    "ss24a": ("sssp_spread", lambda r24: r24, {}),
    "cc24": ("cc_spread", lambda r24: G.rmat(24, 16, seed=2, symmetrise=True), {}),
}


@pytest.mark.timeout(600)
@pytest.mark.parametrize("key", list(DUMP_CASES))
def test_dump_matches_reference_binary_digest(key, rmat24, tmp_path):
    app, make, opts = DUMP_CASES[key]
    ref = DIGESTS[key]
    print(f"{key}: generating the graph", flush=True)
    g = make(rmat24)
    assert (g.V, g.E) == (ref["V"], ref["E"])
    out = tmp_path / "dump.txt"
    print(f"{key}: {app} on the GPU, dumping", flush=True)
    r = EN.run_app(app, g, source=0, out_path=str(out), **opts)
    assert list(r.active) == ref["trace"]
    assert out.stat().st_size == ref["dump_bytes"]
    assert sha256_of(out) == ref["dump_sha256"]


def test_cc_rmat22_symmetrised_full_size():
    print("generating symmetrised R-MAT scale 22 (seed 2)", flush=True)
    g = G.rmat(22, 16, seed=2, symmetrise=True)
    ref = O.min_ancestor(g)
    print("cc_spread on the GPU", flush=True)
    r = EN.run_app("cc_spread", g)
    assert r.gathered > 0
    assert np.array_equal(r.values, ref)


def test_sssp_rmat24_unit_full_size(rmat24):
    src = rmat24.max_degree_vertex()
    ref = O.bfs(rmat24, src)
    print("sssp_spread on the GPU", flush=True)
    r = EN.run_app("sssp_spread", rmat24, source=src)
    assert r.gathered > 0
    assert np.array_equal(r.values, ref)


def test_sssp_rmat24_weighted_full_size(rmat24):
    gw = rmat24.with_weights(3)
    src = rmat24.max_degree_vertex()
    print("Dijkstra on the host", flush=True)
    ref = O.dijkstra(gw, src)
    print("sssp_weighted_spread on the GPU", flush=True)
    r = EN.run_app("sssp_weighted_spread", gw, source=src)
    assert np.array_equal(r.values, ref)


def test_sssp_grid4096_weighted_full_size():
    g = G.grid(4096, 4096).with_weights(3)
    ref = O.dijkstra(g, 4097)
    print("sssp_weighted_spread on the grid", flush=True)
    r = EN.run_app("sssp_weighted_spread", g, source=4097)
    assert r.supersteps > 1000
    assert np.array_equal(r.values, ref)
