"""Builds libfastfib_amd.so (gfx950) in place, next to this file.

    python -m fastfib_amd.build          (or: python fastfib_amd/build.py)

hipcc compiles each HIP source for gfx950 only, with packed FP32 ops
disabled, then links one shared library whose exported symbols are the
C-ABI declared in include/fastfib_amd.h; gcc then builds the fastfib
command line (cli/fastfib.c) against that C-ABI.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfastfib_amd.so")
SOURCES = ["ff_ntt.hip", "ff_ntt_list.hip", "ff_queue.hip", "ff_crt.hip", "ff_limbs.hip", "ff_runtime.hip", "ff_engine.hip", "ff_batch.hip", "ff_tune.hip", "ff_mlp.hip", "ff_bigops.hip",
           "ff_decimal.hip", "ff_split.hip", "ff_early.hip", "ff_noalloc.hip", "ff_probe.hip", "ff_school.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _compile(src: str, objdir: str) -> str:
    obj = os.path.join(objdir, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    err = "\n".join(l for l in (r.stderr or "").splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{err}")
    if err.strip():
        print(err, file=sys.stderr)
    return obj


def build(verbose: bool = True) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, objdir), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    build_cli()
    if verbose:
        print(f"built {OUT} and {CLI_OUT}")
    return OUT


CLI_SRC = os.path.join(HERE, "cli", "fastfib.c")
CLI_OUT = os.path.join(HERE, "fastfib")


def build_cli() -> str:
    """The fastfib command line (cli/fastfib.c), a C program linked to the
    library beside it (found at run time through its rpath)."""
    cc = os.environ.get("CC", "gcc")
    include = os.path.join(os.path.dirname(HERE), "include")
    tmp = CLI_OUT + ".tmp"
    cmd = [cc, "-O2", "-std=c11", "-D_POSIX_C_SOURCE=200809L", "-Wall", "-Wextra", "-Werror", f"-I{include}",
           "-o", tmp, CLI_SRC, f"-L{HERE}", "-lfastfib_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"cli build failed:\n{r.stderr}")
    os.replace(tmp, CLI_OUT)
    return CLI_OUT


if __name__ == "__main__":
    build()
