"""Writes tests/golden/reference_line_counts.json: the line count of every
text source file of the reference repository (path relative to its root).

Run in a container that holds the reference checkout (default
/root/reference); the output is a data fixture (file names and counts, no
source text) that tests/test_citations.py checks citations against.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

EXTS = {".cpp", ".hpp", ".h", ".c", ".cl", ".txt", ".bat", ".aff", ".obj", ".log", ".vcxproj", ".sln"}


def main(root: str = "/root/reference") -> None:
    base = Path(root)
    counts = {}
    for p in sorted(base.rglob("*")):
        if p.is_file() and p.suffix.lower() in EXTS and ".git" not in p.parts:
            with open(p, "rb") as f:
                counts[str(p.relative_to(base))] = sum(1 for _ in f)
    out = Path(__file__).resolve().parent / "reference_line_counts.json"
    out.write_text(json.dumps(counts, indent=0, sort_keys=True) + "\n")
    print(f"{len(counts)} files -> {out}")


if __name__ == "__main__":
    main(*sys.argv[1:])
