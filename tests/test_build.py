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
    for entry in APPS:
        name = entry[0]
        for width in (entry[3] if len(entry) > 3 else (32,)):
            assert (LIB / f"libipr_{name}_{width}.so").exists(), name
            assert (ROOT / "ipregel_amd" / "bin" / f"{name}_{width}").exists(), name


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


def _resource_usage(src, defines, tmp_path):
    """Per-kernel (scratch bytes per lane, VGPR spills, waves per SIMD by
    registers) of a HIP source compiled for gfx950 as the build compiles it."""
    import subprocess
    from ipregel_amd import build as B
    r = subprocess.run([B.HIPCC, *B.HIP_FLAGS, "-x", "hip", "-ffp-contract=off", f"-I{INC}",
                        "-DIP_VERTEX_ID_TYPE=uint32_t", '-DIP_APPLICATION="check"', '-DIP_MACHINE="check"',
                        '-DVERSION="check"', *defines, "--cuda-device-only", "-c",
                        "-Rpass-analysis=kernel-resource-usage", str(src), "-o", str(tmp_path / "k.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    usage = {}
    for block in r.stderr.split("Function Name: ")[1:]:
        name = block.split()[0]
        get = lambda key: int(re.search(re.escape(key) + r":\s*(\d+)", block).group(1))  # noqa: E731
        usage[name] = (get("ScratchSize [bytes/lane]"), get("VGPRs Spill"), get("Occupancy [waves/SIMD]"))
    assert usage
    return usage


@pytest.mark.parametrize("app,defines,gather_waves", [
    ("cc.c", ["-DIP_USE_SPREAD", "-DIP_COMBINER_MIN"], 5),
    ("sssp.c", ["-DIP_USE_SPREAD", "-DIP_COMBINER_MIN"], 5),
    ("cc.c", ["-DIP_COMBINER_MIN"], 5),
    ("sssp_weighted.c", ["-DIP_USE_SPREAD", "-DIP_COMBINER_MIN"], 4),
    ("cc.c", ["-DIP_USE_SPREAD", "-DIP_COMBINER_MIN", "-UIP_VERTEX_ID_TYPE", "-DIP_VERTEX_ID_TYPE=uint64_t"], 4),
    ("pagerank.c", ["-DIP_USE_SINGLE_BROADCAST", "-DIP_COMBINER_SUM"], 4)])
def test_engine_kernels_spill_free(app, defines, gather_waves, tmp_path):
    """The engine's kernels compile without scratch or spills, and the
    dense-superstep gather keeps the 5 wavefronts per SIMD that its 32 KB LDS
    copy allows (its launch is sized for 5 workgroups per CU); with arc
    weights or 8-byte messages it keeps 4 rather than spill."""
    usage = _resource_usage(ROOT / "ipregel_amd" / "apps" / app, defines, tmp_path)
    for name, (scratch, spill, _) in usage.items():
        assert scratch == 0 and spill == 0, (name, scratch, spill)
    gather = [u for n, u in usage.items() if "ipr_k_gather7" in n]
    assert len(gather) == 1 and gather[0][2] >= gather_waves, gather


def test_hip_layer_kernels_spill_free(tmp_path):
    """The HIP layer's own kernels (not rocprim's, used once at load) compile
    without scratch or spills."""
    usage = _resource_usage(ROOT / "ipregel_amd" / "csrc" / "hip" / "ipr_hip.hip", [], tmp_path)
    own = {n: u for n, u in usage.items() if "rocprim" not in n}
    assert any("k_compact" in n for n in own) and any("k_finalize" in n for n in own)
    for name, (scratch, spill, _) in own.items():
        assert scratch == 0 and spill == 0, (name, scratch, spill)


def test_bench_refuses_more_than_one_gpu():
    import json
    import subprocess
    import sys
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_every_engine_variable_is_documented():
    """Each IPR_* environment variable the native code reads has a line in
    INTEGRATION.md's tables."""
    read = set()
    for f in list((ROOT / "ipregel_amd" / "csrc").rglob("*")) + list(INC.glob("*.h")):
        if f.suffix in (".c", ".cpp", ".hip", ".h"):
            read |= set(re.findall(r'(?:env_\w+|getenv)\("(IPR_[A-Z0-9_]+)"', f.read_text()))
    documented = set(re.findall(r"IPR_[A-Z0-9_]+", (ROOT / "INTEGRATION.md").read_text()))
    assert read and not read - documented, sorted(read - documented)
