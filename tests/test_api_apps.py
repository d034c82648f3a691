"""The parts of the iPregel surface that the benchmark applications do not
reach, run through the engine's own API-coverage applications
(ipregel_amd/apps/cc_send.c, nbr_view.c) as command-line programs:

- ip_send_message to each out-neighbour, in the full-sweep push and the
  spread variants, in _32 and _64 builds: its labels and per-superstep active
  counts must equal the reference cc binaries' (tests/golden/), since sending
  to every out-neighbour is what ip_broadcast does. The _32 builds declare
  IP_NEIGHBOUR_COUNT_TYPE as the id type (benchmarks/template.c:2-3) and so
  read a graph whose .idx holds 4-byte offsets.
- ip_send_message in a single-broadcast build: the reference's abort text and
  exit status (src/combiner_single_broadcast_postamble.h:42-49).
- v->in_neighbours / in_neighbour_count in ip_compute, and v->out_neighbours
  through ip_get_vertex_by_id after ip_run, in _32 and _64 builds, with
  directed=false (in-lists are the out-lists) and directed=true (mirrored in
  ascending source order, src/iPregel_postamble.h:429-487): checksums that
  weigh each id by its position in the list, against numpy.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "ipregel_amd" / "bin"

CASES32 = ["rmat10", "rmat10sym", "twocomp10", "chain10", "star65", "emptyends8", "grid8x8"]
CASES64 = ["rmat10", "twocomp10", "chain10", "star65"]
ABORT_TEXT = ("The function send_message should not be used in the SINGLE_BROADCAST version; "
              "only broadcast() should be called, and once per superstep maximum.")


def read_triple(root: Path, id_bytes: int = 4):
    V, E = (int(x) for x in (root.parent / f"{root.name}.config").read_text().split())
    idx = np.fromfile(f"{root}.idx", dtype=np.uint64, count=V)
    adj = np.fromfile(f"{root}.adj", dtype=np.uint32 if id_bytes == 4 else np.uint64, count=E)
    return V, E, np.append(idx, np.uint64(E)), adj.astype(np.uint64)


def write_idx32_copy(case: str, dst: Path) -> Path:
    """<case> with its .idx offsets narrowed to 4 bytes"""
    root = dst / case
    (dst / f"{case}.config").write_text((GOLDEN / f"{case}.config").read_text())
    (dst / f"{case}.adj").write_bytes((GOLDEN / f"{case}.adj").read_bytes())
    idx = np.fromfile(GOLDEN / f"{case}.idx", dtype=np.uint64)
    assert idx.max(initial=0) < 2**32
    idx.astype(np.uint32).tofile(dst / f"{case}.idx")
    return root


def run(binname: str, root: Path, out: Path, *extra: str):
    args = [str(BIN / binname), str(root), str(out), "1", "static", "0", *extra]
    print(f"{binname} {root.name} {' '.join(extra)}", flush=True)
    return subprocess.run(args, capture_output=True, text=True, timeout=120)


def trace_of(stdout: str):
    return [int(x.split(":")[1]) for x in stdout.splitlines() if "ActiveVertexCount" in x]


def test_binaries_built():
    for name in ("cc_send_32", "cc_send_64", "cc_send_spread_32", "cc_send_spread_64",
                 "cc_send_single_broadcast_32", "nbr_view_32", "nbr_view_64",
                 "nbr_view_single_broadcast_32", "nbr_view_single_broadcast_64",
                 "nbr_view_single_broadcast_spread_32", "nbr_view_single_broadcast_spread_64"):
        assert (BIN / name).exists(), name


def test_idx32_copy_round_trips(tmp_path):
    root = write_idx32_copy("rmat10", tmp_path)
    a = np.fromfile(f"{root}.idx", dtype=np.uint32).astype(np.uint64)
    b = np.fromfile(GOLDEN / "rmat10.idx", dtype=np.uint64)
    assert a.dtype == np.uint64 and np.array_equal(a, b)


def checksum_lists(idx, adj):
    """sum over i of (list[i] + 1) * (i + 1) per list, modulo 2^64"""
    V = idx.size - 1
    out = np.zeros(V, np.uint64)
    deg = np.diff(idx).astype(np.int64)
    pos = np.arange(adj.size, dtype=np.uint64) - np.repeat(idx[:-1], deg) + np.uint64(1)
    terms = (adj + np.uint64(1)) * pos
    owner = np.repeat(np.arange(V), deg)
    np.add.at(out, owner, terms)
    return out


def in_lists(V, idx, adj, directed):
    if not directed:
        return idx, adj
    src = np.repeat(np.arange(V, dtype=np.uint64), np.diff(idx).astype(np.int64))
    order = np.lexsort((src, adj))            # by destination, then ascending source
    cnt = np.bincount(adj.astype(np.int64), minlength=V)
    iidx = np.zeros(V + 1, np.uint64)
    iidx[1:] = np.cumsum(cnt)
    return iidx, src[order]


def test_checksum_helpers_on_a_hand_case():
    # arcs 0->2, 1->2, 2->0: in-list of 2 is [0, 1] (ascending sources)
    idx = np.array([0, 1, 2, 3], np.uint64)
    adj = np.array([2, 2, 0], np.uint64)
    iidx, iadj = in_lists(3, idx, adj, True)
    assert list(iidx) == [0, 1, 1, 3] and list(iadj) == [2, 0, 1]
    # (0+1)*1 + (1+1)*2 = 5 for vertex 2; (2+1)*1 = 3 for vertex 0
    assert list(checksum_lists(iidx, iadj)) == [3, 0, 5]


@pytest.mark.gpu
@pytest.mark.parametrize("variant,ref", [("cc_send", "cc"), ("cc_send_spread", "cc_spread")])
@pytest.mark.parametrize("width", [32, 64])
def test_send_message_matches_reference_cc(variant, ref, width, tmp_path):
    cases = CASES32 if width == 32 else CASES64
    for case in cases:
        if width == 32:
            root = write_idx32_copy(case, tmp_path)
            golden = f"{case}__{ref}_32"
        else:
            root = GOLDEN / f"{case}_w64"
            golden = f"{case}_w64__{ref}_64"
        out = tmp_path / f"{case}.out"
        p = run(f"{variant}_{width}", root, out)
        assert p.returncode == 0, p.stdout[-2000:]
        assert out.read_text() == (GOLDEN / f"{golden}.out").read_text(), case
        expect = [int(x) for x in (GOLDEN / f"{golden}.trace").read_text().split()]
        assert trace_of(p.stdout) == expect, case


@pytest.mark.gpu
def test_send_message_in_single_broadcast_aborts(tmp_path):
    root = write_idx32_copy("chain10", tmp_path)
    p = run("cc_send_single_broadcast_32", root, tmp_path / "o.txt")
    assert p.returncode == 255
    assert ABORT_TEXT in p.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["nbr_view", "nbr_view_single_broadcast",
                                     "nbr_view_single_broadcast_spread"])
@pytest.mark.parametrize("width", [32, 64])
@pytest.mark.parametrize("directed", [0, 1])
def test_neighbour_lists_and_vertex_lookup(variant, width, directed, tmp_path):
    cases = CASES32 if width == 32 else CASES64
    for case in cases:
        root = GOLDEN / (case if width == 32 else f"{case}_w64")
        V, E, idx, adj = read_triple(root, 4 if width == 32 else 8)
        iidx, iadj = in_lists(V, idx, adj, directed)
        hin = checksum_lists(iidx, iadj)
        hout = checksum_lists(idx, adj)
        with np.errstate(over="ignore"):
            value = hin * np.uint64(31) + hout
        expect = "".join(f"{v}: {int(value[v])} {int(iidx[v + 1] - iidx[v])} {int(idx[v + 1] - idx[v])}\n"
                         for v in range(V))
        out = tmp_path / f"{case}.out"
        p = run(f"{variant}_{width}", root, out, str(directed))
        assert p.returncode == 0, p.stdout[-2000:]
        assert out.read_text() == expect, case
