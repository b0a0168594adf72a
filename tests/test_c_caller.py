"""The drop-in boundary from a C caller: tests/c_abi/bigint_caller.c is compiled
against include/ alone and linked to libntt_amd.so, as a caller of the
reference's ntt/lib/arith/multiply.h would be.

CPU: the program and every public header compile (C and C++); a managed BigInt
that is too short for its product aborts in bigint_resize, as in the reference
(ntt/lib/arith/bigint.c:46-60).
GPU: the program's checks against GMP: the reference's own differential test
shape (ntt/app/compute_e/test_gmp_oracle.c:55-76: xorshift64 seed
0x123456789abcdef, operands of N/2 - 16 limbs) at log_n 20, 22, 24 and 27, and
the API battery (threshold, dispatch count, aliasing, resize, zero operands).
"""
import ctypes
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "c_abi", "bigint_caller")
LIB = os.path.join(ROOT, "cryptography_amd", "libntt_amd.so")


def _need_prog():
    if not os.path.exists(PROG):
        pytest.fail(f"{PROG} is not built: run `make -C tests/c_abi` (or __graft_entry__.build())")


def test_caller_program_built():
    _need_prog()
    out = subprocess.run([PROG], capture_output=True, text=True, timeout=60)
    assert out.returncode == 2 and "usage" in out.stderr


@pytest.mark.parametrize("lang,flags", [("c", ["-x", "c", "-std=c11"]), ("c++", ["-x", "c++", "-std=c++17"])])
def test_public_headers_compile(tmp_path, lang, flags):
    """Every header in include/ compiles on its own and together, warnings as errors."""
    headers = sorted(os.path.basename(h) for h in glob.glob(os.path.join(ROOT, "include", "*.h")))
    src = tmp_path / ("all." + ("c" if lang == "c" else "cc"))
    src.write_text("".join(f'#include "{h}"\n' for h in headers) + "int main(void) { return 0; }\n")
    cc = "gcc" if lang == "c" else "g++"
    r = subprocess.run([cc, *flags, "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


_ABORT_SCRIPT = r"""
import ctypes, sys
class BigInt(ctypes.Structure):
    _fields_ = [("limbs", ctypes.POINTER(ctypes.c_uint64)), ("n_limbs", ctypes.c_int),
                ("cap", ctypes.c_int), ("managed", ctypes.c_int)]
L = ctypes.CDLL(sys.argv[1])
buf = (ctypes.c_uint64 * 4)()
b = BigInt(ctypes.cast(buf, ctypes.POINTER(ctypes.c_uint64)), 0, 4, 1)
L.bigint_resize(ctypes.byref(b), 4)      # within the capacity: no-op
print("within-cap ok", flush=True)
L.bigint_resize(ctypes.byref(b), 100)    # a managed BigInt cannot grow
print("not reached", flush=True)
"""


def test_managed_bigint_too_short_aborts():
    """A managed BigInt that would have to grow aborts (host-side check: the
    limbs here are ordinary memory marked managed, no GPU is touched)."""
    r = subprocess.run([sys.executable, "-c", _ABORT_SCRIPT, LIB], capture_output=True, text=True, timeout=120)
    assert "within-cap ok" in r.stdout and "not reached" not in r.stdout
    assert r.returncode == -6, (r.returncode, r.stderr)
    assert "managed BigInt" in r.stderr


def _run(args, timeout):
    _need_prog()
    print("bigint_caller", *args, flush=True)
    r = subprocess.run([PROG, *args], capture_output=True, text=True, timeout=timeout)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    return r.stdout


@pytest.mark.gpu
def test_c_caller_api_battery():
    out = _run(["api"], 120)
    assert out.count("PASS") >= 50


@pytest.mark.gpu
@pytest.mark.parametrize("log_n", [20, 22, 24])
def test_c_caller_reference_gmp_oracle(log_n):
    out = _run(["gmp", str(log_n)], 120)
    assert out.count("PASS") == 4


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_c_caller_reference_gmp_oracle_log_n_27():
    """log_n 27: operands of 2^26 - 16 limbs (about 2^32 bits each), a
    transform of 2^29 coefficients; GMP takes about a minute on one core."""
    out = _run(["gmp", "27"], 560)
    assert out.count("PASS") == 4
