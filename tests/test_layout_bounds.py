"""Bounds audit of the skewed layout used by the lexicographic sweep (CPU).

Compiles tests/native/skew_bounds.cpp against the device layout header and
runs it: every index the sweep and the assembly kernels' skewed writers form
is checked against its allocation, and each fetch against the cell the
wavefront step needs, for mesh shapes from 1 x 1 to 1024 rows.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_skewed_layout_indices_stay_in_bounds(tmp_path):
    exe = tmp_path / "skew_bounds"
    subprocess.run(["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    os.path.join(HERE, "native", "skew_bounds.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=False, capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
