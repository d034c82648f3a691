import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def golden_cases():
    return sorted(p.stem for p in GOLDEN.glob("*.config"))


def golden_runs(wide: bool = False):
    """(case, reference binary) pairs that have a golden output: the _32
    builds on the cases, or (wide) the _64 builds on the <case>_w64 copies
    written with 64-bit ids."""
    out = []
    for p in sorted(GOLDEN.glob("*__*.out")):
        case, binname = p.stem.split("__")
        if case.endswith("_w64") == wide:
            out.append((case, binname))
    return out


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN
