"""The C-ABI library loads and exports every symbol declared in include/geowarp_amd.h
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, 'include', 'geowarp_amd.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(gm_[a-z0-9_]+)\s*\(', txt)))


def test_header_declares_the_python_bindings():
    from geowarp_amd import _lib
    assert set(_declared()) == set(_lib.exported_symbols())


def test_library_exports_every_declared_symbol():
    from geowarp_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('library not built')
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing
    assert lib.gm_version() >= 1


def test_product_path_refuses_to_run_without_a_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip('GPU present')
    from geowarp_amd.solvers._engine import Engine
    from geowarp_amd import GMError
    with pytest.raises(GMError):
        Engine(2, False, (4, 4), (0.1, 0.1, 0.0))
