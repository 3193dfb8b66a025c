"""The fastfib command line and a plain C caller of the C-ABI.

The command line re-expresses the 22 cases of the reference's
tests/cli.ps1:9-32 (argument order, --print, --help/-h, the n prompt on
stdin with blanks, the 64-bit edge F(93)/F(94), and exit status 1 with a
message for bad input), byte for byte in what each case expects or rejects.
Cases that stop at argument checking run anywhere; the ones that compute
F(n) need the GPU.

tests/c/abi_hexfib.c is compiled by gcc against include/fastfib_amd.h and
linked to libfastfib_amd.so (no Python in between); it prints F(n) in the
format of tests/hexfib.c:9-10, which must equal Python's format(x, "x").
"""
import os
import subprocess

import pytest

from oracle.oracle import FINGERPRINT_PRIMES, decimal_mod, fib_int, fib_mod

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fastfib_amd", "fastfib")

# (arguments, stdin or None, exit status, expected text, rejected text or None)
CASES = [
    ("10 --print", None, 0, "fibonacci(10) = 55", None),
    ("--print 10", None, 0, "fibonacci(10) = 55", None),
    ("0 --print", None, 0, "fibonacci(0) = 0", None),
    ("1 --print", None, 0, "fibonacci(1) = 1", None),
    ("2 --print", None, 0, "fibonacci(2) = 1", None),
    ("93 --print", None, 0, "fibonacci(93) = 12200160415121876738", None),
    ("94 --print", None, 0, "fibonacci(94) = 19740274219868223167", None),
    ("007 --print", None, 0, "fibonacci(7) = 13", None),
    ("1000", None, 0, "Time:", "fibonacci("),
    ("--help", None, 0, "usage: fastfib", None),
    ("-h", None, 0, "usage: fastfib", None),
    ("--print", "12", 0, "fibonacci(12) = 144", None),
    ("--print", "  12  ", 0, "fibonacci(12) = 144", None),
    ("--print", "abc", 1, "invalid n: abc", None),
    ("--print", "", 1, "invalid n", None),
    ("-5", None, 1, "invalid argument: -5", None),
    ("+5", None, 1, "invalid argument: +5", None),
    ("1e9", None, 1, "invalid argument: 1e9", None),
    ("0x10", None, 1, "invalid argument: 0x10", None),
    ("18446744073709551616", None, 1, "invalid argument", None),
    ("10 20", None, 1, "invalid argument: 20", None),
    ("--bogus", None, 1, "usage: fastfib", None),
]


def _needs_gpu(case) -> bool:
    args, stdin, status, expect, _ = case
    return status == 0 and "usage" not in expect


def _run(args: str, stdin):
    # stdin as PowerShell pipes a string: the text and a newline
    r = subprocess.run([EXE, *args.split(" ")], input=None if stdin is None else stdin + "\n",
                       capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout + r.stderr


def _ids(cases):
    return [f"{a} <{s}>" if s is not None else a for a, s, *_ in cases]


CPU_CASES = [c for c in CASES if not _needs_gpu(c)]
GPU_CASES = [c for c in CASES if _needs_gpu(c)]


def test_case_split():
    assert len(CASES) == 22 and len(CPU_CASES) == 11 and len(GPU_CASES) == 11


@pytest.mark.parametrize("case", CPU_CASES, ids=_ids(CPU_CASES))
def test_cli_checks_arguments(case):
    args, stdin, status, expect, reject = case
    assert os.path.exists(EXE), "build the library and the command line first (python -m fastfib_amd.build)"
    code, text = _run(args, stdin)
    assert code == status, (code, text)
    assert expect in text, text
    assert reject is None or reject not in text


@pytest.mark.gpu
@pytest.mark.parametrize("case", GPU_CASES, ids=_ids(GPU_CASES))
def test_cli_computes(case):
    args, stdin, status, expect, reject = case
    code, text = _run(args, stdin)
    assert code == status, (code, text)
    assert expect in text, text
    assert reject is None or reject not in text
    lines = [l for l in text.splitlines() if l.strip()]
    # the last lines are the timings, as main.c:98-102 prints them
    if "--print" in args:
        assert lines[-2].startswith("Time: ") and lines[-2].endswith(" seconds")
        assert lines[-1].startswith("Decimal conversion: ") and lines[-1].endswith(" seconds")
    else:
        assert lines[-1].startswith("Time: ") and lines[-1].endswith(" seconds")


@pytest.mark.gpu
def test_cli_large_n_digits():
    n = 100000
    code, text = _run(f"{n} --print", None)
    assert code == 0
    line = next(l for l in text.splitlines() if l.startswith(f"fibonacci({n}) = "))
    assert line.split(" = ", 1)[1] == str(fib_int(n))


@pytest.fixture(scope="module")
def abi_hexfib(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("abi") / "abi_hexfib")
    lib_dir = os.path.join(ROOT, "fastfib_amd")
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-o", exe, os.path.join(ROOT, "tests", "c", "abi_hexfib.c"), f"-L{lib_dir}",
                        "-lfastfib_amd", f"-Wl,-rpath,{lib_dir}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_caller_builds(abi_hexfib):
    """The header compiles as C11 with warnings as errors and the program links."""
    assert os.path.exists(abi_hexfib)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [0, 1, 2, 93, 94, 1000, 12345, 100000, 1000003])
def test_c_caller_hexfib(abi_hexfib, n):
    r = subprocess.run([abi_hexfib, str(n)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    hexline, sq, dec = r.stdout.splitlines()
    x = fib_int(n)
    assert hexline == format(x, "x")
    assert sq == "square ok"
    if n <= 100000:
        assert dec == str(x)
    else:
        assert [decimal_mod(dec, p) for p in FINGERPRINT_PRIMES] == [fib_mod(n, p) for p in FINGERPRINT_PRIMES]
