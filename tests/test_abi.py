"""CPU checks of the C-ABI boundary (no GPU calls).

* libfastfib_amd.so loads and exports exactly the functions include/*.h
  declares with FF_API;
* the product package never reaches the oracle (no CPU fallback);
* the oracle libraries load.
"""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        for m in re.finditer(r"FF_API\s+[\w\s\*]+?\b(ff_\w+)\s*\(", open(h).read()):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    import subprocess

    import fastfib_amd
    lib = ctypes.CDLL(fastfib_amd.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 20
    for name in sorted(declared):
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", fastfib_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert exported == declared, (exported - declared, declared - exported)


def test_product_code_has_no_cpu_fallback():
    """Nothing in the product package loads the oracle or the reference."""
    for path in glob.glob(os.path.join(ROOT, "fastfib_amd", "**", "*.*"), recursive=True):
        if path.endswith((".py", ".hip", ".h", ".cpp")):
            text = open(path).read()
            assert "oracle" not in text.lower().replace("oracle/", ""), path
            assert "libffref" not in text, path


def test_oracle_libraries_load():
    from oracle.oracle import ORACLE_SO, REF_SO, Oracle
    assert os.path.exists(ORACLE_SO)
    Oracle()
    if os.path.exists(REF_SO):
        ctypes.CDLL(REF_SO)
