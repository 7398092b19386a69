"""Every line citation of the reference's files (``name.jl:a-b``, ``name.ipynb:a``, ...)
in this repository's code and documents points inside the file it names.

A citation is a file name (with or without its directory) followed by ``:a`` or
``:a-b``, and optionally more ``, c-d`` ranges of the same file. A bare name is looked
up among the reference's files by base name. Runs where the reference sources are
mounted (it reads them as text only); skipped elsewhere."""

import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"

SCANNED = ["cloud.jl_amd", "include", "oracle", "tests", "bench.py", "__graft_entry__.py", "DESIGN.md",
           "INTEGRATION.md", "README.md"]
SUFFIXES = (".py", ".hip", ".h", ".cpp", ".md", ".jl")
CITE = re.compile(r"([A-Za-z0-9_./-]+\.(?:jl|ipynb|toml))((?::\d+(?:-\d+)?)(?:,\s*\d+(?:-\d+)?)*)")


def _reference_files():
    by_path, by_name = {}, {}
    for dirpath, dirnames, filenames in os.walk(REFERENCE):
        dirnames[:] = [d for d in dirnames if d != ".git"]
        for f in filenames:
            if f.endswith((".jl", ".ipynb", ".toml")):
                full = os.path.join(dirpath, f)
                rel = os.path.relpath(full, REFERENCE)
                with open(full, errors="replace") as fh:
                    n = sum(1 for _ in fh)
                by_path[rel] = n
                by_name.setdefault(f, []).append(rel)
    return by_path, by_name


def _citations():
    for entry in SCANNED:
        base = os.path.join(ROOT, entry)
        paths = [base] if os.path.isfile(base) else [
            os.path.join(d, f) for d, _, fs in os.walk(base) for f in fs if f.endswith(SUFFIXES)]
        for path in paths:
            if "__pycache__" in path or os.sep + "_build" + os.sep in path:
                continue
            with open(path, errors="replace") as fh:
                for lineno, line in enumerate(fh, 1):
                    for m in CITE.finditer(line):
                        ranges = re.findall(r"(\d+)(?:-(\d+))?", m.group(2))
                        yield os.path.relpath(path, ROOT), lineno, m.group(1), [
                            (int(a), int(b) if b else int(a)) for a, b in ranges]


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference sources not mounted")
def test_every_reference_citation_resolves_inside_its_file():
    by_path, by_name = _reference_files()
    bad, count = [], 0
    for where, lineno, name, ranges in _citations():
        name = name.lstrip("./")
        if name.startswith("root/reference/"):
            name = name[len("root/reference/"):]
        if name in by_path:
            candidates = [name]
        else:
            candidates = [p for p in by_name.get(os.path.basename(name), []) if p.endswith(name)]
        if not candidates:
            bad.append(f"{where}:{lineno}: {name} is not a file of the reference")
            continue
        count += 1
        n = max(by_path[c] for c in candidates)
        for a, b in ranges:
            if not (1 <= a <= b <= n):
                bad.append(f"{where}:{lineno}: {name}:{a}-{b} outside 1-{n}")
    assert count > 100
    assert not bad, "\n".join(bad)
