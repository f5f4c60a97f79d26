"""Every reference citation of the form `path.cr:A` or `path.cr:A-B` in the
boundary header, the oracle, the product sources, the tests and the design
documents names a file of the reference and lines that exist in it.  Runs
only where the reference checkout is present (it is read as text, never
run); elsewhere it skips."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
CITE = re.compile(r"([A-Za-z0-9_/.]*[A-Za-z0-9_]\.cr):(\d+)(?:-(\d+))?")
# a bare ":A-B" after a citation on the same line continues that file
CONT = re.compile(r"(?<![\w/.):])[ (,]:(\d+)(?:-(\d+))?")
SCAN = ["include", "oracle", "ph-core_amd/csrc", "ph-core_amd/host", "ph-core_amd/phase", "tests",
        "DESIGN.md", "INTEGRATION.md", "README.md", "bench.py", "__graft_entry__.py"]


def _files():
    for p in SCAN:
        full = os.path.join(ROOT, p)
        if os.path.isfile(full):
            yield full
            continue
        for dp, _, fs in os.walk(full):
            if "_build" in dp or "__pycache__" in dp or "golden" in dp:
                continue
            for f in fs:
                if f.endswith((".h", ".hpp", ".cpp", ".hip", ".py", ".md")):
                    yield os.path.join(dp, f)


def _by_name():
    names = {}
    for dp, _, fs in os.walk(REF):
        for f in fs:
            if f.endswith(".cr"):
                names.setdefault(f, []).append(os.path.join(dp, f))
    return names


def _resolve(rel, names):
    """A citation path relative to the reference root; a bare or partial
    path resolves by its file name when that name is unique."""
    full = os.path.join(REF, rel)
    if os.path.isfile(full):
        return full
    cands = [p for p in names.get(os.path.basename(rel), []) if p.endswith("/" + rel)]
    return cands[0] if len(cands) == 1 else None


def _lines(path, cache={}):
    if path not in cache:
        with open(path, errors="replace") as fh:
            cache[path] = sum(1 for _ in fh)
    return cache[path]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout not present")
def test_reference_citations_exist():
    names = _by_name()
    bad = []
    n = 0

    def check(where, path, a, b, text):
        last = int(b) if b else a
        if a < 1 or last < a or last > _lines(path):
            bad.append(f"{where}: {text} past the end of {os.path.relpath(path, REF)} "
                       f"({_lines(path)} lines)")

    for f in _files():
        for ln, text in enumerate(open(f, errors="replace"), 1):
            where = f"{os.path.relpath(f, ROOT)}:{ln}"
            cites = list(CITE.finditer(text))
            for i, m in enumerate(cites):
                rel, a, b = m.group(1), int(m.group(2)), m.group(3)
                path = _resolve(rel, names)
                n += 1
                if path is None:
                    bad.append(f"{where}: {rel} does not name one reference file")
                    continue
                check(where, path, a, b, m.group(0))
                end = cites[i + 1].start() if i + 1 < len(cites) else len(text)
                for c in CONT.finditer(text, m.end(), end):
                    check(where, path, int(c.group(1)), c.group(2), c.group(0).strip())
    assert n > 100
    assert not bad, "\n".join(bad)
