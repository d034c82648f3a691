"""Build recipe for the oracle (test infrastructure only).

  oracle/liboracle.so   the C restatement (ipr_oracle.c) plus the graph
                        generator it reads graphs through
  oracle/_ref/bin/*     the reference's own applications, compiled from
                        /root/reference with its makefile flags, when that
                        tree is present (see oracle/ref.mk); never shipped
                        in git, never used by the product
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
REF = Path("/root/reference")


def build() -> Path:
    out = HERE / "liboracle.so"
    cmd = ["gcc", "-std=c11", "-O2", "-fPIC", "-shared", "-fopenmp", "-Wall", "-Wextra",
           f"-I{ROOT / 'include'}", str(HERE / "ipr_oracle.c"),
           str(ROOT / "ipregel_amd/csrc/graph/ipr_graph.c"), "-o", str(out)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}")
    build_ref()
    return out


def build_ref() -> bool:
    """Compile the reference applications into oracle/_ref when the reference
    tree is present; returns False when it is absent."""
    if not (REF / "src" / "iPregel.h").exists():
        return False
    r = subprocess.run(["make", "-s", "-f", str(HERE / "ref.mk"), f"REF={REF}",
                        f"OUT={HERE / '_ref'}", "-j8"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode:
        raise RuntimeError(f"reference build failed:\n{r.stdout}")
    return True


if __name__ == "__main__":
    build()
