"""The drop-in boundary as the reference would meet it.

* Every reference citation in the sources (file:line, file:a-b, lists)
  points inside the cited file of the reference checkout.
* The GPU binding (fastfib_amd/integration/mul_gpu.c) compiles inside the
  reference's own unity build next to its Big (the header defers to the
  caller's Big under FF_CALLER_BIG), and a program that does so
  (oracle/ref_binding_main.c, built into oracle/_ref by oracle/Makefile)
  gets the same F(n), products and decimal digits from the reference's code
  and from the GPU through the binding.

The first two need the reference checkout and are skipped without it.
"""
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
BINDING_EXE = os.path.join(ROOT, "oracle", "_ref", "ref_gpu_binding")

need_ref = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "fastfib.c")), reason="no reference checkout")

CITE = re.compile(r"\b((?:src/|tests/|tools/)?[A-Za-z_][A-Za-z0-9_]*\.(?:c|h|py|ps1))((?::\d+(?:-\d+)?)(?:,\d+(?:-\d+)?)*)")


def _ref_line_counts():
    counts = {}
    for p in glob.glob(os.path.join(REF, "**", "*"), recursive=True):
        if os.path.isfile(p) and p.endswith((".c", ".h", ".py", ".ps1")):
            rel = os.path.relpath(p, REF)
            with open(p, encoding="utf-8", errors="replace") as f:
                counts[rel] = sum(1 for _ in f)
    return counts


def _sources():
    paths = [os.path.join(ROOT, "bench.py"), os.path.join(ROOT, "__graft_entry__.py")]
    for d in ("oracle", "include", "fastfib_amd", "tests"):
        for p in glob.glob(os.path.join(ROOT, d, "**", "*"), recursive=True):
            if p.endswith((".c", ".h", ".hip", ".py")) and "_ref" not in p:
                paths.append(p)
    return paths


@need_ref
def test_citations_point_inside_the_cited_files():
    counts = _ref_line_counts()
    by_name = {}
    for rel, n in counts.items():
        by_name.setdefault(os.path.basename(rel), []).append((rel, n))
    bad, seen = [], 0
    for path in _sources():
        with open(path, encoding="utf-8") as f:
            for ln, line in enumerate(f, 1):
                for m in CITE.finditer(line):
                    name, spans = m.group(1), m.group(2)
                    cands = by_name.get(os.path.basename(name))
                    if not cands:
                        continue  # not a reference file (this repo's own names)
                    if "/" in name:
                        cands = [c for c in cands if c[0] == name] or cands
                    limit = max(n for _, n in cands)
                    seen += 1
                    if max(int(x) for x in re.findall(r"\d+", spans)) > limit:
                        bad.append(f"{os.path.relpath(path, ROOT)}:{ln}: {name}{spans} (file has {limit} lines)")
    assert seen > 100
    assert not bad, "\n".join(bad)


@need_ref
def test_binding_compiles_in_the_reference_unity_build():
    r = subprocess.run(
        ["gcc", "-fsyntax-only", "-std=gnu11", "-march=x86-64-v3", "-D_GNU_SOURCE", "-DFASTFIB_GPU", "-Wall",
         "-Wno-unused-function", "-I", os.path.join(ROOT, "oracle", "ref_port"), "-I-",
         "-I", os.path.join(ROOT, "fastfib_amd", "integration"), "-I", os.path.join(ROOT, "include"),
         "-I", os.path.join(REF, "src"), "-I", REF, os.path.join(ROOT, "oracle", "ref_binding_main.c")],
        capture_output=True, text=True)
    errors = [l for l in r.stderr.splitlines() if "obsolete option" not in l and l.strip()]
    assert r.returncode == 0 and not any("warning" in l or "error" in l for l in errors), r.stderr


def test_header_compiles_without_the_caller_big():
    """Alone, the header supplies its own layout-identical Big."""
    src = '#include "fastfib_amd.h"\nint main(void) { Big b = {0, 1}; return (int)b.n - 1; }\n'
    r = subprocess.run(["gcc", "-std=c99", "-pedantic", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-x", "c", "-fsyntax-only", "-"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_reference_with_gpu_binding():
    if not os.path.exists(BINDING_EXE):
        pytest.skip("oracle/_ref/ref_gpu_binding is not built (needs the reference checkout at build time)")
    ns = ["0", "1", "2", "93", "94", "1000", "100000", "1000001", "10000000"]
    r = subprocess.run([BINDING_EXE, *ns], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    assert r.stdout.split("\n")[:-1] == [f"n {n} ok" for n in ns]
