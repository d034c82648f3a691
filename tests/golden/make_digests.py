"""Writes tests/golden/digests.json: for the benchmark configurations at full
size, the SHA-256 of the ip_dump text and the per-superstep ActiveVertexCount
trace of the reference's own binaries (compiled from the reference tree into
oracle/_ref/bin by oracle/ref.mk), run on the graphs this project's generator
writes:

  pr16           pagerank_single_broadcast_32, R-MAT scale 16 ef 16 seed 1
  pr16_directed  pagerank_directed_single_broadcast_32 (directed=true), same graph
  cc22           cc_spread_32, symmetrised R-MAT scale 22 ef 16 seed 2
  ss24a          sssp_spread_32, R-MAT scale 24 ef 16 seed 3 (source 0, which
                 is the highest out-degree vertex of this graph)
  cc24           cc_spread_32, symmetrised R-MAT scale 24 ef 16 seed 2

The texts are too large to commit (up to 300 MB); the digests are data that
the GPU tests compare the engine's own dump against.

Run in the build container only: python tests/golden/make_digests.py
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from ipregel_amd import graph as G  # noqa: E402

REF_BIN = ROOT / "oracle" / "_ref" / "bin"

CONFIGS = {
    "pr16": ("pagerank_single_broadcast_32", lambda: G.rmat(16, 16, seed=1)),
    "pr16_directed": ("pagerank_directed_single_broadcast_32", lambda: G.rmat(16, 16, seed=1)),
    "cc22": ("cc_spread_32", lambda: G.rmat(22, 16, seed=2, symmetrise=True)),
    "ss24a": ("sssp_spread_32", lambda: G.rmat(24, 16, seed=3)),
    "cc24": ("cc_spread_32", lambda: G.rmat(24, 16, seed=2, symmetrise=True)),
}


def sha256_file(p: Path) -> str:
    h = hashlib.sha256()
    with open(p, "rb") as f:
        for block in iter(lambda: f.read(1 << 24), b""):
            h.update(block)
    return h.hexdigest()


def main() -> None:
    if not REF_BIN.exists():
        raise SystemExit("oracle/_ref/bin is missing: build the reference first (oracle/build.py)")
    out = {}
    threads = str(os.cpu_count() or 1)
    for key, (binname, make) in CONFIGS.items():
        g = make()
        if binname.startswith("sssp") and g.max_degree_vertex() != 0:
            raise SystemExit(f"{key}: the highest out-degree vertex is not 0; the reference runs from 0")
        with tempfile.TemporaryDirectory(dir="/tmp") as td:
            root = Path(td) / "g"
            g.write(root)
            dump = Path(td) / "dump.txt"
            args = [str(REF_BIN / binname), str(root), str(dump), threads, "static", "0"]
            if binname.startswith("sssp"):
                args.append("0")
            p = subprocess.run(args, capture_output=True, text=True, check=True)
            trace = [int(x.split(":")[1]) for x in p.stdout.splitlines() if "ActiveVertexCount" in x]
            out[key] = {"binary": binname, "V": g.V, "E": g.E, "dump_sha256": sha256_file(dump),
                        "dump_bytes": dump.stat().st_size, "trace": trace}
        print(key, out[key]["dump_sha256"][:16], len(trace), flush=True)
    (HERE / "digests.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
