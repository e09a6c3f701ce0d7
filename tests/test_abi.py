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
