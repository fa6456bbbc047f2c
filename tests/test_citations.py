"""Every file:line citation of a reference file in this repository's sources
points inside the cited file.

Sources cite the reference (TreESaP) as `path/to/File.hpp:12-34` relative to
its tree, where path is any trailing part of the file's path (`stencil.hpp`,
`util/stencil.hpp`, `src/treesap/util/stencil.hpp`). The check needs the
reference tree and is skipped where it is absent.
"""
from __future__ import annotations

import os
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path(os.environ.get("TSP_REFERENCE_TREE", "/root/reference"))

# files this project writes (the survey, verdict and advice files are not)
SOURCES = ["include", "treesap_amd", "oracle", "tests", "bench.py", "__graft_entry__.py", "DESIGN.md",
           "INTEGRATION.md", "README.md"]
SUFFIXES = {".hpp", ".h", ".hip", ".c", ".cpp", ".py", ".md"}
CITE = re.compile(r"((?:[A-Za-z0-9_.-]+/)*[A-Za-z0-9_.-]+\.(?:hpp|cpp|txt|cmake|md|gif|png)):"
                  r"(\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*)")


def _ref_files():
    out = {}
    for p in REF.rglob("*"):
        if p.is_file() and ".git" not in p.parts:
            out[p.relative_to(REF).as_posix()] = p
    return out


def _source_files():
    for s in SOURCES:
        p = ROOT / s
        if p.is_file():
            yield p
        elif p.is_dir():
            for q in sorted(p.rglob("*")):
                if q.is_file() and q.suffix in SUFFIXES and "bin" not in q.parts:
                    yield q


def _citations():
    for f in _source_files():
        if f.resolve() == Path(__file__).resolve():
            continue
        text = f.read_text(errors="replace")
        for lineno, line in enumerate(text.splitlines(), 1):
            for m in CITE.finditer(line):
                yield f.relative_to(ROOT).as_posix(), lineno, m.group(1), m.group(2)


@pytest.mark.skipif(not REF.is_dir(), reason="reference tree not present")
def test_reference_citations_are_in_range():
    ref = _ref_files()
    lengths = {}
    bad, seen = [], 0
    for src, lineno, name, ranges in _citations():
        cands = [k for k in ref if k == name or k.endswith("/" + name)]
        if not cands:
            # a file of this repository (or a generic name such as README.md of ours)
            if (ROOT / name).exists() or list(ROOT.glob(f"**/{name}")):
                continue
            bad.append(f"{src}:{lineno}: {name} is not a reference file")
            continue
        top = max(int(x) for r in ranges.split(",") for x in r.strip().split("-"))
        seen += 1
        ok = False
        for c in cands:
            if c not in lengths:
                lengths[c] = len(ref[c].read_text(errors="replace").splitlines())
            ok = ok or top <= lengths[c]
        if not ok:
            bad.append(f"{src}:{lineno}: {name}:{ranges} past the end ({[lengths[c] for c in cands]} lines)")
    assert seen > 50, seen
    assert not bad, "\n".join(bad)
