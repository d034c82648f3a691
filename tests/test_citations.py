"""Every reference citation in the tree (`src/<file>:<lines>`,
`benchmarks/<file>:<lines>`, `makefile:<lines>`, the same without the
directory, and bare `:<lines>` continuations after one of them on the same
line) names a file of the reference and lines that exist in it. Runs only where
the reference tree is present (/root/reference); it checks the numbers, not
what the cited lines say."""
from __future__ import annotations

import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference")
SUFFIXES = {".c", ".h", ".hip", ".cpp", ".py", ".md", ".mk"}
SKIP_DIRS = {"_work", "gpurun_out", ".git", "__pycache__", "_ref", "golden"}
# notes about the build written by others, quoting the reference's own numbers
SKIP_FILES = {"VERDICT.md", "ADVICE.md", "SURVEY.md", "BASELINE.md"}
FILE_RE = r"(?:(?:src|benchmarks)/)?(?:[\w-]+/)*(?:combiner[a-z_]*|iPregel[a-z_]*|[a-z_]+)\.(?:h|c)|makefile"
CITE = re.compile(r"(?<![\w/.])(" + FILE_RE + r")((?::\s?\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*)+)")
BARE = re.compile(r"(?<=[\s(`,])(:\d+(?:-\d+)?(?:,\d+(?:-\d+)?)*)(?![\w:])")


def ref_file(name: str) -> Path | None:
    p = REF / name
    if p.is_file():
        return p
    base = name.split("/")[-1]
    hits = [q for q in (REF / "src").rglob(base)] + [q for q in (REF / "benchmarks").glob(base)]
    return hits[0] if len(hits) == 1 else None


def ranges(spec: str):
    for part in spec.replace(" ", "").lstrip(":").split(","):
        part = part.lstrip(":")
        a, _, b = part.partition("-")
        yield int(a), int(b) if b else int(a)


def sources():
    for p in sorted(ROOT.rglob("*")):
        if p.is_file() and p.suffix in SUFFIXES and p.name not in SKIP_FILES \
                and not SKIP_DIRS.intersection(p.relative_to(ROOT).parts):
            yield p


@pytest.mark.skipif(not REF.is_dir(), reason="the reference tree is not present")
def test_every_cited_reference_range_exists():
    bad, seen = [], 0
    lengths: dict[Path, int] = {}
    for p in sources():
        for ln, line in enumerate(p.read_text(errors="ignore").splitlines(), 1):
            last = None
            spans = []
            for m in CITE.finditer(line):
                name = m.group(1)
                rf = ref_file(name)
                if rf is None:
                    # a file of this project named like a C source (e.g. ip_driver.c:12)
                    if name.startswith(("src/", "benchmarks/")) or name == "makefile":
                        bad.append(f"{p.relative_to(ROOT)}:{ln}: no reference file {name}")
                    continue
                last = rf
                spans.append((m.start(), m.end()))
                seen += 1
                n = lengths.setdefault(rf, len(rf.read_text(errors="ignore").splitlines()))
                for a, b in ranges(m.group(2)):
                    if not 1 <= a <= b <= n:
                        bad.append(f"{p.relative_to(ROOT)}:{ln}: {name}:{a}-{b} (file has {n} lines)")
            if last is None:
                continue
            n = lengths[last]
            for m in BARE.finditer(line):
                if any(s <= m.start() < e for s, e in spans) or m.start() < spans[0][0]:
                    continue
                seen += 1
                for a, b in ranges(m.group(1)):
                    if not 1 <= a <= b <= n:
                        bad.append(f"{p.relative_to(ROOT)}:{ln}: {last.name}{m.group(1)} (file has {n} lines)")
    assert seen > 60, seen
    assert not bad, "\n".join(bad)
