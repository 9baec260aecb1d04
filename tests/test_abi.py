"""The C-ABI library loads and exports every entry point include/pqgpu.h declares (no GPU needed)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pqgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pqg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("pqg_ci_create", "pqg_ci_sigma", "pqg_ci_davidson", "pqg_ao2mo", "pqg_serve_submit"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    import ctypes

    from polyquant_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpqgpu.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if getattr(L, s, None) is None]
    assert missing == []
    # the ctypes binding covers exactly the declared entry points
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == declared_symbols()


def test_no_gpu_means_loud_failure():
    """Without a visible GPU the product raises instead of falling back to the CPU."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from polyquant_amd import _lib
    from polyquant_amd.ci import Context

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpqgpu.so not built")
    with pytest.raises(_lib.PqgError):
        Context()
