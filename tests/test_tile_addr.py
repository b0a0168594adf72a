"""Address identities of the LDS-tiled transform passes
(cryptography_amd/csrc/tile_geom.h), compiled for the host: the split tile
transfer address (a wave-uniform chunk part plus a lane offset) equals the
whole one, a thread's stored words lie at base + k*step, and the smallest
round's register groups cover the tile once and store at gb + i*gs, for every
pass geometry the launcher makes, tiles at both ends and the row pads in use.  The
kernels' memory accesses are built on these identities; the GPU tests check
the products they lead to."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_tile_address_identities(tmp_path):
    exe = str(tmp_path / "tile_addr_check")
    src = os.path.join(HERE, "host_stub", "tile_addr_check.cpp")
    subprocess.run([HIPCC, "-O2", "-o", exe, src], check=True, capture_output=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("bad 0")
