"""Checker of the reference citations in this repository's sources.

A citation is `path/file.jl:a-b` (or `:a`, or `:a-b, c, d-e` continuing with commas)
naming a file of the reference tree (src/, test/, docs/, bench/, README.md).  A
citation without a directory (`cubeki.jl:37-48`) names the reference file of that
base name when exactly one exists.  Every cited line range must lie inside the file.

    python tests/citations.py [REFERENCE_ROOT]    # prints every broken citation
"""
from __future__ import annotations

import os
import re
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCAN_DIRS = ("include", "oracle", "flexiblefmm.jl_amd", "tests")
SCAN_FILES = ("bench.py", "__graft_entry__.py", "DESIGN.md", "INTEGRATION.md")
SCAN_EXT = (".h", ".hip", ".c", ".py", ".md", ".cpp")

CITE = re.compile(r"(?<![\w./-])((?:(?:src|test|docs|bench)/)?(?:[\w]+/)*[\w.]+\.(?:jl|md|toml))"
                  r":(\d+(?:-\d+)?(?:,\s*\d+(?:-\d+)?(?![\w.]))*)")


def reference_files(root):
    out = {}
    for d, _, fs in os.walk(root):
        if "/.git" in d:
            continue
        for f in fs:
            if f.endswith((".jl", ".md", ".toml")):
                rel = os.path.relpath(os.path.join(d, f), root)
                out[rel] = os.path.join(d, f)
    return out


def line_counts(files):
    n = {}
    for rel, path in files.items():
        with open(path, "rb") as fh:
            n[rel] = fh.read().count(b"\n") + 1
    return n


def scanned_files():
    for d in SCAN_DIRS:
        for base, _, fs in os.walk(os.path.join(REPO, d)):
            if "_build" in base or "__pycache__" in base:
                continue
            for f in fs:
                if f.endswith(SCAN_EXT):
                    yield os.path.join(base, f)
    for f in SCAN_FILES:
        p = os.path.join(REPO, f)
        if os.path.exists(p):
            yield p


def resolve(name, files):
    if name in files:
        return name
    for pre in ("src/", "src/operators/", "src/tree/", "test/", "docs/", "bench/"):
        if pre + name in files:
            return pre + name
    base = os.path.basename(name)
    hits = [r for r in files if r.endswith("/" + name) or os.path.basename(r) == base and "/" not in name]
    return hits[0] if len(hits) == 1 else None


def check(root):
    """[(repo file, line, citation, reason)] of every citation that does not resolve."""
    files = reference_files(root)
    counts = line_counts(files)
    bad = []
    for path in scanned_files():
        rel_repo = os.path.relpath(path, REPO)
        with open(path, encoding="utf-8", errors="replace") as fh:
            for ln, line in enumerate(fh, 1):
                for m in CITE.finditer(line):
                    name, spans = m.group(1), m.group(2)
                    if name in ("README.md",) and "reference" not in line and rel_repo.endswith(".md"):
                        pass
                    ref = resolve(name, files)
                    if ref is None:
                        if name.endswith(".jl"):
                            bad.append((rel_repo, ln, m.group(0), "no such reference file"))
                        continue
                    for span in re.split(r",\s*", spans):
                        a, _, b = span.partition("-")
                        a = int(a)
                        b = int(b) if b else a
                        if a < 1 or b < a or b > counts[ref]:
                            bad.append((rel_repo, ln, m.group(0),
                                        f"{ref} has {counts[ref]} lines"))
    return bad


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    bad = check(root)
    for b in bad:
        print(f"{b[0]}:{b[1]}: {b[2]}  ({b[3]})")
    print(f"{len(bad)} broken citations")
