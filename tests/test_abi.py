"""The C ABI library loads and exports every symbol include/cfdgpu.h declares
(no compute call: runs without a GPU)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "cfdgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cfdgpu_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_exports(G):
    assert sorted(G.EXPORTS) == declared_symbols()


def test_every_declared_symbol_is_exported(G):
    lib = G.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_error_strings(G):
    assert G.lib().cfdgpu_version().startswith(b"cfdgpu")
    assert isinstance(G.lib().cfdgpu_last_error(), bytes)


def test_host_side_validation_needs_no_gpu(G):
    """Argument checks happen before any device call."""
    lib = G.lib()
    assert lib.cfdgpu_server_create(None, None) == 1  # InvalidArgument
    opts = G.server_options()
    assert opts.max_slots == 64 and opts.chunk_iterations == 20 and opts.warm_start == 1
    assert opts.hbm_budget_bytes == 200e9 and opts.model == G.MODEL_SST
    import ctypes as C
    h = C.c_void_p()
    solver = G.server_options(settings=G.default_settings(pressure_solver=9))
    assert lib.cfdgpu_server_create(C.byref(solver), C.byref(h)) == 1  # InvalidArgument
    passes = G.server_options(settings=G.default_settings(refinement_passes=0))
    assert lib.cfdgpu_server_create(C.byref(passes), C.byref(h)) == 1
    bad = G.server_options(settings=G.default_settings(precision=7))
    assert lib.cfdgpu_server_create(C.byref(bad), C.byref(h)) == 1  # InvalidArgument


def test_launch_tuning_is_validated_and_kept(G):
    before = G.launch_tuning()
    assert before in {(c, r) for c in (0, 1, 2, 4) for r in (16, 32)}
    try:
        for c, r in ((1, 16), (0, 16), (4, 32), (2, 32)):
            G.set_launch_tuning(c, r)
            assert G.launch_tuning() == (c, r)
        for c, r in ((3, 32), (-1, 16), (8, 32), (2, 8), (2, 64)):
            with pytest.raises(G.CfdError):
                G.set_launch_tuning(c, r)
            assert G.launch_tuning() == (2, 32)  # a rejected call changes nothing
    finally:
        G.set_launch_tuning(*before)


def test_server_create_fails_loudly_without_a_device(G):
    if G.device_count() > 0:
        pytest.skip("a GPU is visible")
    import ctypes as C
    h = C.c_void_p()
    opts = G.server_options()
    assert G.lib().cfdgpu_server_create(C.byref(opts), C.byref(h)) == 6  # InitializationFailure
    assert b"GPU" in G.lib().cfdgpu_last_error()
