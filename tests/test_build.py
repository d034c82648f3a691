"""CPU checks of the built libraries: they load, and every function that
include/*.h declares is exported (no compute calls; no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from ipregel_amd import engine as EN

ROOT = Path(__file__).resolve().parents[1]
INC = ROOT / "include"
LIB = ROOT / "ipregel_amd" / "lib"


def declared(header):
    text = (INC / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(ipr?_\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if not n.startswith("ipr_mix64") and n != "ipr_hash3"))


def test_hip_layer_exports():
    L = C.CDLL(str(LIB / "libipregel_hip.so"))
    names = declared("ipregel_hip.h")
    assert len(names) > 20
    for n in names:
        assert hasattr(L, n), n


def test_graph_lib_exports():
    L = C.CDLL(str(LIB / "libipregel_graph.so"))
    for n in declared("ipregel_graph.h"):
        assert hasattr(L, n), n


@pytest.mark.parametrize("lib", sorted(p.name for p in LIB.glob("libipr_*_32.so")))
def test_app_libs_export_driver_and_ipregel_api(lib):
    L = C.CDLL(str(LIB / lib))
    for n in declared("ipregel_driver.h"):
        assert hasattr(L, n), n
    for n in ("ip_init", "ip_run", "ip_dump", "main"):
        assert hasattr(L, n), n


def test_all_apps_built():
    from ipregel_amd.build import APPS
    for name, _, _ in APPS:
        assert (LIB / f"libipr_{name}_32.so").exists(), name
        assert (ROOT / "ipregel_amd" / "bin" / f"{name}_32").exists(), name


def test_batch_lib_exports():
    L = C.CDLL(str(LIB / "libipregel_batch.so"))
    names = declared("ipregel_batch.h")
    assert "ipr_serve" in names and "ipr_batch_ppr" in names
    for n in names:
        assert hasattr(L, n), n
    from ipregel_amd import batch as BA
    assert set(BA.EXPORTS) == set(names)


def test_batch_kernels_call_and_scratch_free(tmp_path):
    """Every kernel of the batch library is a single straight function with no
    private (scratch) segment: an accidental self-call once recursed through
    the scratch stack until the GPU faulted."""
    import subprocess
    src = ROOT / "ipregel_amd" / "csrc" / "batch" / "ipr_batch.hip"
    out = tmp_path / "batch.s"
    r = subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=off",
                        "-Wno-unused-command-line-argument", f"-I{INC}", "--cuda-device-only", "-S",
                        str(src), "-o", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    asm = out.read_text()
    assert "s_swappc_b64" not in asm
    sizes = re.findall(r"\.private_segment_fixed_size:\s+(\d+)", asm)
    assert sizes and all(int(x) == 0 for x in sizes)
