"""Every `file:line` citation of a reference file lies inside that file.

The line counts come from tests/golden/reference_line_counts.json (written
from the reference checkout by tests/golden/make_reference_line_counts.py),
so the check runs without the reference.  A citation is `Name.ext:N`,
`Name.ext:N-M` or a comma list of those, optionally with a path prefix; a
name shared by several reference files must fit the longest of them unless
the prefix picks one.
"""
from __future__ import annotations

import json
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
COUNTS = json.loads((ROOT / "tests" / "golden" / "reference_line_counts.json").read_text())

CITE = re.compile(r"((?:[A-Za-z0-9_]+/)*[A-Za-z0-9_]+\.(?:cpp|hpp|cl|h|c|txt|bat|vcxproj))"
                  r"(:\d+(?:-\d+)?(?:,\d+(?:-\d+)?)*)")
SCANNED = ["include", "oracle", "hulktracer_amd", "tests", "DESIGN.md", "INTEGRATION.md", "README.md",
           "bench.py", "__graft_entry__.py"]
SUFFIXES = {".py", ".h", ".hpp", ".hip", ".cpp", ".c", ".md", ".txt", ""}


def _by_name():
    names = {}
    for rel, n in COUNTS.items():
        names.setdefault(rel.rsplit("/", 1)[-1], []).append((rel, n))
    return names


def _files():
    for entry in SCANNED:
        p = ROOT / entry
        if p.is_file():
            yield p
            continue
        for f in sorted(p.rglob("*")):
            if (f.is_file() and f.suffix in SUFFIXES and "golden" not in f.parts and "_build" not in f.parts
                    and "__pycache__" not in f.parts and f.name not in ("Makefile", "test_citations.py")):
                yield f


def citations():
    names = _by_name()
    for f in _files():
        for ln, line in enumerate(f.read_text(errors="replace").splitlines(), 1):
            for m in CITE.finditer(line):
                path = m.group(1)
                cands = names.get(path.rsplit("/", 1)[-1])
                if not cands:
                    continue  # not a reference file (e.g. this build's own sources)
                if "/" in path:
                    picked = [c for c in cands if c[0].endswith(path)]
                    cands = picked or cands
                limit = max(n for _, n in cands)
                for part in m.group(2)[1:].split(","):
                    lo, _, hi = part.partition("-")
                    yield f"{f.relative_to(ROOT)}:{ln}", m.group(0), int(lo), int(hi or lo), limit


def test_fixture_covers_the_hot_path_files():
    names = _by_name()
    for n in ("GPUTrace.cl", "GPUTrace_stackless.cl", "shadePicture.cl", "GPU_common.cl", "SharedFunctions.hpp",
              "SharedStructs.hpp", "KDTreeSAH.cpp", "KDTreeCPU.cpp", "GLVisitorGPUTrace.cpp"):
        assert n in names, n
    assert dict(names["GPUTrace.cl"])["HulkEngine/src/GPUTrace.cl"] == 161
    assert dict(names["shadePicture.cl"])["HulkEngine/src/shadePicture.cl"] == 55


def test_every_citation_lies_inside_its_file():
    found = list(citations())
    assert len(found) > 200, "the scan found too few citations: the pattern or the scanned files changed"
    bad = [f"{where}: {text} (file has {limit} lines)" for where, text, lo, hi, limit in found
           if not (1 <= lo <= hi <= limit)]
    assert not bad, "citations outside their file:\n" + "\n".join(bad)


@pytest.mark.parametrize("text,ok", [("GPUTrace.cl:122-160", True), ("GPUTrace.cl:193-228", False),
                                     ("shadePicture.cl:7-55", True), ("shadePicture.cl:73-121", False),
                                     ("HulkEngine/src/GPUTrace_stackless.cl:163-165", True)])
def test_checker_catches_an_out_of_range_line(text, ok):
    names = _by_name()
    m = CITE.search(text)
    cands = names[m.group(1).rsplit("/", 1)[-1]]
    limit = max(n for _, n in cands)
    lo, _, hi = m.group(2)[1:].partition("-")
    assert (int(hi or lo) <= limit) == ok
