"""The HIP sources built for the CPU emulation (tests/emu: lanes as threads,
wave shuffles, DPP shifts and LDS emulated) under AddressSanitizer, run on
cases whose indexing differs from the small parity meshes: a sweep that spans
two waves, and batches of members of different widths on the plain and on the
fast path. Any out-of-bounds access aborts the child process. Runs on the CPU
only; it checks the kernels' memory safety and their results before they go
to the GPU."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EMU = os.path.join(HERE, "emu")


def _asan_runtime():
    out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def emu_library():
    runtime = _asan_runtime()
    if runtime is None:
        pytest.skip("no AddressSanitizer runtime for gcc")
    subprocess.run(["make", "-C", EMU, "-j", "4"], check=True, capture_output=True)
    return os.path.join(EMU, "libcfdgpu_emu.so"), runtime


@pytest.mark.parametrize("case", ["parity100", "batch_plain", "batch_fast"])
def test_emulated_kernels_are_memory_safe(emu_library, case):
    lib, runtime = emu_library
    env = dict(os.environ, CFDGPU_LIBRARY=lib, LD_PRELOAD=runtime, ASAN_OPTIONS="detect_leaks=0",
               PYTHONPATH=os.pathsep.join([os.path.join(ROOT, "cfd-simulator_amd"), os.path.join(ROOT, "oracle"),
                                           HERE]))
    proc = subprocess.run([sys.executable, os.path.join(HERE, "emu_cases.py"), case], env=env, capture_output=True,
                          text=True, timeout=900)
    assert proc.returncode == 0, proc.stderr[-4000:]
    assert proc.stdout.strip().startswith("ok"), proc.stdout
