"""Writes the golden fixtures of tests/golden/ by running the reference's own
applications (compiled from the reference tree into oracle/_ref/bin by
oracle/ref.mk) on small graphs.

For every case: the graph triple (<case>.config/.idx/.adj), and for every
reference binary that applies, its ip_dump output (<case>__<bin>.out) and its
per-superstep ActiveVertexCount trace (<case>__<bin>.trace, one number a line).
Four cases are also written with 64-bit ids (<case>_w64.*) and run by the
reference's _64 builds (<case>_w64__<bin>_64.out / .trace).
pagerank_directed_single_broadcast_32 is benchmarks/pagerank.c with
directed=true (oracle/ref.mk), run on every case.
SSSP runs from vertex 0: the reference cannot run from another source
(benchmarks/sssp.c:101 reads the schedule string).

Run in the build container only: python tests/golden/make_golden.py
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from ipregel_amd import graph as G  # noqa: E402

REF_BIN = ROOT / "oracle" / "_ref" / "bin"

BINS = {
    "pagerank": ["pagerank_single_broadcast_32", "pagerank_directed_single_broadcast_32"],
    "cc": ["cc_32", "cc_spread_32", "cc_single_broadcast_32", "cc_single_broadcast_spread_32"],
    "sssp": ["sssp_32", "sssp_spread_32", "sssp_single_broadcast_32", "sssp_single_broadcast_spread_32"],
}


def cases() -> dict[str, G.Graph]:
    c: dict[str, G.Graph] = {}
    # a directed chain 0 -> 1 -> ... -> 9 (long diameter, one broadcaster per superstep)
    c["chain10"] = G.from_arcs(10, np.arange(9), np.arange(1, 10))
    # hub 0 pointing at 1..63, plus a back arc 63 -> 0; isolated last vertex 64
    c["star65"] = G.from_arcs(65, [0] * 63 + [63], list(range(1, 64)) + [0])
    # two symmetric components {0..4} ring and {5..7} path, isolated 8, 9;
    # duplicate and self-loop arcs in the input are dropped by the builder
    s = [0, 1, 2, 3, 4, 5, 6, 1, 3, 3]
    d = [1, 2, 3, 4, 0, 6, 7, 2, 3, 3]
    c["twocomp10"] = G.from_arcs(10, s + d[:7], d + s[:7])
    # empty graph rows at both ends: vertex 0 and the last vertex have no arcs
    c["emptyends8"] = G.from_arcs(8, [1, 2, 3, 4, 5, 6], [2, 3, 1, 6, 4, 5])
    c["rmat8"] = G.rmat(8, 16, seed=11)
    c["rmat8sym"] = G.rmat(8, 16, seed=12, symmetrise=True)
    c["rmat10"] = G.rmat(10, 16, seed=1)
    c["rmat10sym"] = G.rmat(10, 16, seed=2, symmetrise=True)
    c["grid8x8"] = G.grid(8, 8)
    return c


def run_ref(binname: str, root: Path, out: Path) -> list[int]:
    args = [str(REF_BIN / binname), str(root), str(out), "2", "static", "0"]
    if binname.startswith("sssp"):
        args.append("0")
    p = subprocess.run(args, capture_output=True, text=True, check=True)
    return [int(line.split(":")[1]) for line in p.stdout.splitlines() if "ActiveVertexCount" in line]


# cases also written with 64-bit ids (<case>_w64.*: the .adj holds uint64
# entries, as the _64 builds read it) and run by the _64 builds
WIDE_CASES = ["chain10", "star65", "twocomp10", "rmat10"]
# reference binaries built in the _32 width only (oracle/ref.mk)
NARROW_ONLY = {"pagerank_directed_single_broadcast_32"}


def write_wide(g: G.Graph, root: Path) -> None:
    g.write(root)
    adj = np.fromfile(f"{root}.adj", dtype=np.uint32)
    adj.astype(np.uint64).tofile(f"{root}.adj")


def main() -> None:
    if not REF_BIN.exists():
        raise SystemExit("oracle/_ref/bin is missing: build the reference first (oracle/build.py)")
    for name, g in cases().items():
        root = HERE / name
        g.write(root)
        for app, bins in BINS.items():
            for b in bins:
                out = HERE / f"{name}__{b}.out"
                trace = run_ref(b, root, out)
                (HERE / f"{name}__{b}.trace").write_text("".join(f"{t}\n" for t in trace))
        if name in WIDE_CASES:
            wroot = HERE / f"{name}_w64"
            write_wide(g, wroot)
            for app, bins in BINS.items():
                for b in bins:
                    if b in NARROW_ONLY:
                        continue
                    b64 = b.replace("_32", "_64")
                    out = HERE / f"{name}_w64__{b64}.out"
                    trace = run_ref(b64, wroot, out)
                    (HERE / f"{name}_w64__{b64}.trace").write_text("".join(f"{t}\n" for t in trace))
        print(name, g.V, g.E)


if __name__ == "__main__":
    main()
