"""Modular primitives of the transform kernels (cryptography_amd/csrc/modarith.h,
lazy_group.h), compiled for the host and checked against 128-bit integer
arithmetic: values and the output ranges the kernels rely on (Montgomery
product (0, 3p) / (0, 2p), Shoup [0, 2p), the one-step reduction of any 64-bit
value), every lazy butterfly group plan the transform rounds use, with inputs
at their bounds, the compile-time roots and the CRT checks.  The reference's
counterparts are ntt/lib/primes.h:193-279 and ntt/lib/shoup.h:45-112."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_modarith_primitives(tmp_path):
    exe = str(tmp_path / "modarith_check")
    src = os.path.join(HERE, "host_stub", "modarith_check.cpp")
    subprocess.run([HIPCC, "-O2", "-o", exe, src], check=True, capture_output=True)
    out = subprocess.run([exe, "2000000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("bad 0")
