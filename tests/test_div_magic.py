"""The hot-cell box test divides a voxel index by dimxy and nx with the
multiply-high constants of div_magic (mcx_amd/csrc/mcx_kernel.h); this
builds a host checker of that formula with g++ and compares it with
integer division over all divisors up to 20000, random large divisors and
the edge numerators (0, d - 1, d, 2^31, 2^32 - 1, ...)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_div_magic_matches_integer_division(tmp_path):
    exe = str(tmp_path / "div_magic_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "mcx_amd", "csrc"), "-I", "/opt/rocm/include",
                    "-D__HIP_PLATFORM_AMD__", os.path.join(HERE, "div_magic_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    checked, bad = map(int, out.stdout.split())
    assert out.returncode == 0 and bad == 0 and checked > 4_000_000, out.stdout
