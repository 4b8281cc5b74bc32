"""Every `src/<file>:N` citation of the reference in the boundary headers, the
library, the oracle, the tests and the design notes points at an existing line
of that file (runs where the reference tree is present)."""
from __future__ import annotations

import pathlib
import re

import pytest

from conftest import ROOT

REF = pathlib.Path("/root/reference")
CITE = re.compile(r"((?:src|tests|third_party|scenes)/[\w/.\-]+\.(?:h|cpp|hpp|cmake))((?::\d+(?:-\d+)?)(?:,\s*:\d+(?:-\d+)?)*)")
WHERE = ["include", "cpu-renderer_amd", "oracle", "tests", "DESIGN.md", "INTEGRATION.md", "README.md", "bench.py",
         "__graft_entry__.py"]
SUFFIXES = {".h", ".hpp", ".cpp", ".hip", ".c", ".py", ".md", ""}


def _files():
    for w in WHERE:
        p = ROOT / w
        if p.is_file():
            yield p
        elif p.is_dir():
            for f in sorted(p.rglob("*")):
                if f.is_file() and f.suffix in SUFFIXES and "_build" not in f.parts and "_ref" not in f.parts \
                        and "__pycache__" not in f.parts and f.name != "Makefile":
                    yield f


@pytest.mark.skipif(not (REF / "src" / "render.cpp").exists(), reason="reference tree absent")
def test_reference_citations_are_in_range():
    lengths, bad, seen = {}, [], 0
    for f in _files():
        for ln, line in enumerate(f.read_text(errors="replace").splitlines(), 1):
            for m in CITE.finditer(line):
                target = REF / m.group(1)
                if not target.exists():
                    if (ROOT / m.group(1)).exists():  # a path of this repository (tests/csrc/...)
                        continue
                    bad.append(f"{f.relative_to(ROOT)}:{ln}: {m.group(0)} (no such file)")
                    continue
                if target not in lengths:
                    lengths[target] = len(target.read_text(errors="replace").splitlines())
                seen += 1
                for num in re.findall(r"\d+", m.group(2)):
                    if not 1 <= int(num) <= lengths[target]:
                        bad.append(f"{f.relative_to(ROOT)}:{ln}: {m.group(0)} ({lengths[target]} lines)")
    assert seen > 100
    assert not bad, "\n".join(bad)
