"""Build the reference's own applications (benchmarks/cc.c, pagerank.c,
sssp.c of an iPregel tree) against this engine's iPregel.h.

The edits are the ones the boundary asks for, applied mechanically to a copy
made under /tmp (the copy is never written into this tree):
  - `__device__` on ip_compute and ip_combine;
  - `__device__` on the application globals that ip_compute reads and main()
    assigns (pagerank.c: ratio, initial_value; sssp.c: start_vertex), and one
    `IP_PUBLISH(name);` per such global just before `ip_run();`.
Everything else (types, hooks, main, printf formats) is compiled as written.
Outputs: ipregel_amd/dropin/bin/<app><variant>_32 (git-ignored, shipped with
the tree like the engine's own binaries).
"""
from __future__ import annotations

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
    built = []
    with tempfile.TemporaryDirectory(prefix="ipr_dropin_", dir="/tmp") as td:
        for app, (globs, variants) in APPS.items():
            qsrc = Path(td) / f"{app}.c"
            qsrc.write_text(qualify((src_dir / f"{app}.c").read_text(), globs))
            for suffix, defines in variants:
                name = f"{app}{suffix}_32"
                obj = Path(td) / f"{name}.o"
                B._run([B.HIPCC, *B.HIP_FLAGS, "-x", "hip", "-ffp-contract=off", "-w",
                        f"-I{B.INC}", "-DIP_VERTEX_ID_TYPE=uint32_t",
                        f'-DIP_APPLICATION="{app}{suffix}"', '-DIP_MACHINE="MI355X"',
                        '-DVERSION="1.0.0"', '-DCOMMITS="none"', *defines, "-c", str(qsrc),
                        "-o", str(obj)])
                exe = OUT / name
                B._run([B.HIPCC, f"-Wl,-rpath,{B.LIB}", "-Wl,-rpath,$ORIGIN/../../lib", str(obj), str(drv),
                        f"-L{B.LIB}", "-lipregel_hip", "-fopenmp", "-o", str(exe)])
                built.append(exe)
    return built


if __name__ == "__main__":
    for p in build():
        print(p)
