"""Build the reference's own applications (benchmarks/cc.c, pagerank.c,
sssp.c of an iPregel tree) against this engine's iPregel.h.

The edits are the ones the boundary asks for, applied mechanically to a copy
made under /tmp (the copy is never written into this tree):
  - `__device__` on ip_compute and ip_combine;
  - `__device__` on the application globals that ip_compute reads and main()
    assigns (pagerank.c: ratio, initial_value; sssp.c: start_vertex), and one
    `IP_PUBLISH(name);` per such global just before `ip_run();`.
Everything else (types, hooks, main, printf formats) is compiled as written.
Outputs: ipregel_amd/dropin/bin/<app><variant>_32 and _64 (IP_VERTEX_ID_TYPE
uint32_t / uint64_t, as the reference makefile builds them; git-ignored,
shipped with the tree like the engine's own binaries).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import tempfile
from pathlib import Path

from . import build as B

REF_DEFAULT = Path("/root/reference")
OUT = B.PKG / "dropin" / "bin"

# application -> (globals read on the device, [(variant suffix, defines)])
APPS = {
    "cc": ([], [("", ["-DIP_COMBINER_MIN"]),
                ("_spread", ["-DIP_USE_SPREAD", "-DIP_COMBINER_MIN"]),
                ("_single_broadcast", ["-DIP_USE_SINGLE_BROADCAST", "-DIP_COMBINER_MIN"]),
                ("_single_broadcast_spread", ["-DIP_USE_SINGLE_BROADCAST", "-DIP_USE_SPREAD",
                                              "-DIP_COMBINER_MIN"])]),
    "pagerank": (["ratio", "initial_value"],
                 [("_single_broadcast", ["-DIP_USE_SINGLE_BROADCAST", "-DIP_COMBINER_SUM"])]),
    "sssp": (["start_vertex"], [("", ["-DIP_COMBINER_MIN"]),
                                ("_spread", ["-DIP_USE_SPREAD", "-DIP_COMBINER_MIN"]),
                                ("_single_broadcast", ["-DIP_USE_SINGLE_BROADCAST", "-DIP_COMBINER_MIN"])]),
}


# the reference makefile's id widths (makefile:16-17): <name>_32 and <name>_64
WIDTHS = [(32, "uint32_t"), (64, "uint64_t")]


def qualify(text: str, device_globals: list[str]) -> str:
    """The device-qualifier edits (and the publish lines) on an app's source."""
    out = re.sub(r"^(\s*)void\s+ip_compute\s*\(", r"\1__device__ void ip_compute(", text, flags=re.M)
    out = re.sub(r"^(\s*)void\s+ip_combine\s*\(", r"\1__device__ void ip_combine(", out, flags=re.M)
    for g in device_globals:
        out, n = re.subn(rf"^(\s*)((?:[A-Za-z_][\w ]*?\s))({g})\s*;", r"\1__device__ \2\3;", out,
                         count=1, flags=re.M)
        if n != 1:
            raise RuntimeError(f"global {g} not found")
    if device_globals:
        pub = "".join(f"IP_PUBLISH({g}); " for g in device_globals)
        out, n = re.subn(r"^(\s*)ip_run\(\);", rf"\1{pub}ip_run();", out, count=1, flags=re.M)
        if n != 1:
            raise RuntimeError("ip_run(); not found")
    return out


def build(ref: Path = REF_DEFAULT) -> list[Path]:
    """Returns the binaries built; an empty list when the reference tree is absent."""
    src_dir = ref / "benchmarks"
    if not (src_dir / "cc.c").exists():
        return []
    OUT.mkdir(parents=True, exist_ok=True)
    drv = B.OBJ / "ip_driver.o"
    if not drv.exists():
        B.build_driver()
    with tempfile.TemporaryDirectory(prefix="ipr_dropin_", dir="/tmp") as td:
        jobs = []
        for app, (globs, variants) in APPS.items():
            qsrc = Path(td) / f"{app}.c"
            qsrc.write_text(qualify((src_dir / f"{app}.c").read_text(), globs))
            for suffix, defines in variants:
                for width, id_type in WIDTHS:
                    jobs.append((app, suffix, defines, width, id_type, qsrc))

        def one(job) -> Path:
            app, suffix, defines, width, id_type, qsrc = job
            name = f"{app}{suffix}_{width}"
            obj = Path(td) / f"{name}.o"
            B._run([B.HIPCC, *B.HIP_FLAGS, "-x", "hip", "-ffp-contract=off", "-w",
                    f"-I{B.INC}", f"-DIP_VERTEX_ID_TYPE={id_type}",
                    f'-DIP_APPLICATION="{app}{suffix}"', '-DIP_MACHINE="MI355X"',
                    '-DVERSION="1.0.0"', '-DCOMMITS="none"', *defines, "-c", str(qsrc),
                    "-o", str(obj)])
            exe = OUT / name
            B._run([B.HIPCC, f"-Wl,-rpath,{B.LIB}", "-Wl,-rpath,$ORIGIN/../../lib", str(obj), str(drv),
                    f"-L{B.LIB}", "-lipregel_hip", "-fopenmp", "-o", str(exe)])
            return exe

        with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
            return list(ex.map(one, jobs))


if __name__ == "__main__":
    for p in build():
        print(p)
