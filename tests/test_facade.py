"""The reference-style C++ API (include/cfdgpu_facade.hpp) compiled against
libcfdgpu.so: on the CPU the host-side parts (meshes, face conditions, fields,
sweepAngles, create's argument check); on the GPU a 200-iteration laminar
C-grid solve with surface extraction and a warm-start write through field()."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "cfd-simulator_amd")


@pytest.fixture(scope="module")
def facade_exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    if not os.path.exists(os.path.join(LIBDIR, "libcfdgpu.so")):
        pytest.fail("libcfdgpu.so is not built")
    exe = tmp_path_factory.mktemp("facade") / "facade_smoke"
    subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "native", "facade_smoke.cpp"), "-L", LIBDIR, "-lcfdgpu",
                    "-Wl,-rpath," + LIBDIR, "-o", str(exe)], check=True, capture_output=True)
    return exe


def test_facade_host_side(facade_exe):
    out = subprocess.run([str(facade_exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stdout + out.stderr


@pytest.mark.gpu
def test_facade_solve_on_the_gpu(facade_exe, gpu_available):
    print("case: C++ facade, NACA0012 118x27 laminar, 200 iterations", flush=True)
    out = subprocess.run([str(facade_exe), "solve"], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stdout + out.stderr
