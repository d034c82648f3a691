"""Build recipe for the MI355X iPregel engine (gfx950).

Produces, in-tree (git-ignored, shipped with the tree):
  ipregel_amd/lib/libipregel_graph.so   graph generators and file triple (C, OpenMP)
  ipregel_amd/lib/libipregel_hip.so     thin HIP C-ABI + type-independent kernels
  ipregel_amd/lib/libipregel_batch.so   resident graphs, batched queries, serving loop
  ipregel_amd/lib/ip_driver.o           C host driver (linked into every app)
  ipregel_amd/lib/libipr_<app>_32.so    app TU + driver, loadable over ctypes
  ipregel_amd/bin/<app>_32              the same app as a command-line program
(and _64 builds, IP_VERTEX_ID_TYPE uint64_t, of the API-coverage apps).

Every HIP translation unit is compiled for gfx950 only, with packed-FP32
instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
Application TUs use -ffp-contract=off so that FP64 arithmetic in ip_compute
rounds like the reference's x86-64 build (no fused multiply-add).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INC = ROOT / "include"
LIB = PKG / "lib"
BIN = PKG / "bin"
OBJ = PKG / "lib" / "obj"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
C_FLAGS = ["-std=c11", "-O3", "-fPIC", "-fopenmp", "-Wall", "-Wextra"]

# (binary name, source, defines); the name follows the reference makefile's
# <app>[_single_broadcast][_spread]_32 scheme (makefile:108-205).
APPS = [
    ("pagerank_single_broadcast", "pagerank.c", ["-DIP_USE_SINGLE_BROADCAST", "-DIP_COMBINER_SUM"]),
    ("pagerank", "pagerank.c", ["-DIP_COMBINER_SUM"]),
    # the same with directed=true: PageRank over the mirrored in-lists
    ("pagerank_directed_single_broadcast", "pagerank.c",
     ["-DIP_USE_SINGLE_BROADCAST", "-DIP_COMBINER_SUM", "-DPAGERANK_DIRECTED"]),
    ("cc_spread", "cc.c", ["-DIP_USE_SPREAD", "-DIP_COMBINER_MIN"]),
    ("cc", "cc.c", ["-DIP_COMBINER_MIN"]),
    ("cc_single_broadcast", "cc.c", ["-DIP_USE_SINGLE_BROADCAST", "-DIP_COMBINER_MIN"]),
    ("cc_single_broadcast_spread", "cc.c",
     ["-DIP_USE_SINGLE_BROADCAST", "-DIP_USE_SPREAD", "-DIP_COMBINER_MIN"]),
    ("cc_generic_spread", "cc.c", ["-DIP_USE_SPREAD"]),
    ("sssp_spread", "sssp.c", ["-DIP_USE_SPREAD", "-DIP_COMBINER_MIN"]),
    ("sssp", "sssp.c", ["-DIP_COMBINER_MIN"]),
    ("sssp_single_broadcast", "sssp.c", ["-DIP_USE_SINGLE_BROADCAST", "-DIP_COMBINER_MIN"]),
    ("sssp_single_broadcast_spread", "sssp.c",
     ["-DIP_USE_SINGLE_BROADCAST", "-DIP_USE_SPREAD", "-DIP_COMBINER_MIN"]),
    ("sssp_weighted_spread", "sssp_weighted.c", ["-DIP_USE_SPREAD", "-DIP_COMBINER_MIN"]),
    ("sssp_weighted", "sssp_weighted.c", ["-DIP_COMBINER_MIN"]),
    # API-coverage applications, in both id widths: ip_send_message to each
    # out-neighbour (and its abort in a single-broadcast build), 4-byte .idx
    # offsets; neighbour-id lists and the host-side vertex lookup
    ("cc_send", "cc_send.c", ["-DIP_COMBINER_MIN"], (32, 64)),
    ("cc_send_spread", "cc_send.c", ["-DIP_USE_SPREAD", "-DIP_COMBINER_MIN"], (32, 64)),
    ("cc_send_single_broadcast", "cc_send.c", ["-DIP_USE_SINGLE_BROADCAST", "-DIP_COMBINER_MIN"], (32,)),
    ("nbr_view", "nbr_view.c", ["-DIP_COMBINER_MIN"], (32, 64)),
    ("nbr_view_single_broadcast", "nbr_view.c", ["-DIP_USE_SINGLE_BROADCAST", "-DIP_COMBINER_MIN"], (32, 64)),
    ("nbr_view_single_broadcast_spread", "nbr_view.c",
     ["-DIP_USE_SINGLE_BROADCAST", "-DIP_USE_SPREAD", "-DIP_COMBINER_MIN"], (32, 64)),
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")


def build_graph_lib() -> Path:
    out = LIB / "libipregel_graph.so"
    _run(["gcc", *C_FLAGS, "-shared", f"-I{INC}", str(PKG / "csrc/graph/ipr_graph.c"), "-o", str(out)])
    return out


def build_hip_lib() -> Path:
    out = LIB / "libipregel_hip.so"
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INC}", str(PKG / "csrc/hip/ipr_hip.hip"), "-o", str(out)])
    return out


def build_batch_lib() -> Path:
    """libipregel_batch.so: resident graphs, scalar and batched query kernels,
    the serving loop. -ffp-contract=off keeps the compensated FP32 sums and the
    FP64 reductions free of fused multiply-adds, so the scalar and batched
    kernels round identically."""
    out = LIB / "libipregel_batch.so"
    src = PKG / "csrc" / "batch"
    _run([HIPCC, *HIP_FLAGS, "-ffp-contract=off", "-shared", f"-I{INC}", f"-I{src}",
          str(src / "ipr_batch.hip"), str(src / "ipr_serve.cpp"),
          f"-L{LIB}", "-lipregel_hip", "-Wl,-rpath,$ORIGIN", "-o", str(out)])
    return out


def build_driver() -> Path:
    out = OBJ / "ip_driver.o"
    _run(["gcc", *C_FLAGS, f"-I{INC}", "-c", str(PKG / "csrc/driver/ip_driver.c"), "-o", str(out)])
    return out


def build_app(name: str, src: str, defines: list[str], driver_obj: Path, width: int = 32) -> tuple[Path, Path]:
    """One application variant: libipr_<name>_<width>.so and bin/<name>_<width>
    (IP_VERTEX_ID_TYPE uint32_t or uint64_t, as the reference makefile's _32
    and _64 builds, makefile:16-17)."""
    obj = OBJ / f"{name}_{width}.o"
    app_label = name
    id_type = "uint32_t" if width == 32 else "uint64_t"
    _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-ffp-contract=off", f"-I{INC}",
          f"-DIP_VERTEX_ID_TYPE={id_type}", f'-DIP_APPLICATION="{app_label}"',
          '-DIP_MACHINE="MI355X"', '-DVERSION="1.0.0-mi355x"',
          *defines, "-c", str(PKG / "apps" / src), "-o", str(obj)])
    so = LIB / f"libipr_{name}_{width}.so"
    exe = BIN / f"{name}_{width}"
    link = [str(obj), str(driver_obj), f"-L{LIB}", "-lipregel_hip", "-fopenmp"]
    _run([HIPCC, "-shared", "-fPIC", "-Wl,-Bsymbolic", "-Wl,-rpath,$ORIGIN", *link, "-o", str(so)])
    _run([HIPCC, "-Wl,-rpath,$ORIGIN/../lib", *link, "-o", str(exe)])
    return so, exe


def build_all(jobs: int | None = None, verbose: bool = True) -> None:
    for d in (LIB, BIN, OBJ):
        d.mkdir(parents=True, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        f_graph = ex.submit(build_graph_lib)
        f_hip = ex.submit(build_hip_lib)
        f_drv = ex.submit(build_driver)
        drv = f_drv.result()
        f_hip.result()
        f_graph.result()
        f_batch = ex.submit(build_batch_lib)
        futs = {}
        for entry in APPS:
            n, s, d = entry[:3]
            for w in (entry[3] if len(entry) > 3 else (32,)):
                futs[ex.submit(build_app, n, s, d, drv, w)] = f"{n}_{w}"
        for f in cf.as_completed(futs):
            f.result()
            if verbose:
                print(f"built {futs[f]}", flush=True)
        f_batch.result()
    # the reference's own applications against this engine (needs the
    # reference tree; a host without it uses the binaries built here)
    from . import dropin
    dropin.build()


if __name__ == "__main__":
    build_all()
