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


LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "pqgpu.h"
#define F(T, m) printf(#T " " #m " %zu %zu\n", offsetof(T, m), sizeof(((T*)0)->m))
#define S(T) printf(#T " sizeof %zu %zu\n", sizeof(T), _Alignof(T))
int main(void) {
  S(pqg_ci_desc); F(pqg_ci_desc, nspecies); F(pqg_ci_desc, n_orb); F(pqg_ci_desc, n_alpha);
  F(pqg_ci_desc, n_beta); F(pqg_ci_desc, charge); F(pqg_ci_desc, nspin_blocks); F(pqg_ci_desc, ex_level);
  F(pqg_ci_desc, max_collective); F(pqg_ci_desc, h); F(pqg_ci_desc, eri);
  S(pqg_dav_params); F(pqg_dav_params, nev); F(pqg_dav_params, nvec_init); F(pqg_dav_params, nvec_max);
  F(pqg_dav_params, maxit); F(pqg_dav_params, tol); F(pqg_dav_params, precision);
  F(pqg_dav_params, guess_perturb);
  S(pqg_dav_report); F(pqg_dav_report, iterations); F(pqg_dav_report, converged);
  F(pqg_dav_report, sigma_calls); F(pqg_dav_report, sigma_calls_fp64); F(pqg_dav_report, restarts);
  F(pqg_dav_report, residual); F(pqg_dav_report, seconds);
  S(pqg_result); F(pqg_result, request_id); F(pqg_result, status); F(pqg_result, nev);
  F(pqg_result, energies); F(pqg_result, t_submit); F(pqg_result, t_admit); F(pqg_result, t_done);
  F(pqg_result, iterations); F(pqg_result, message);
  return 0;
}
"""


def test_struct_layouts_match_the_ctypes_mirror(tmp_path):
    """sizeof/offsetof of every struct in include/pqgpu.h, from a C program compiled against the
    header, equal the ctypes mirror in polyquant_amd/_lib.py field by field."""
    import ctypes
    import shutil
    import subprocess

    from polyquant_amd import _lib

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    mirror = {"pqg_ci_desc": _lib.CiDesc, "pqg_dav_params": _lib.DavParams, "pqg_dav_report": _lib.DavReport,
              "pqg_result": _lib.Result}
    seen = set()
    for line in filter(None, out):
        name, field, off, size = line.split()
        cls = mirror[name]
        if field == "sizeof":
            assert ctypes.sizeof(cls) == int(off), name
            continue
        desc = getattr(cls, field)
        assert (desc.offset, desc.size) == (int(off), int(size)), (name, field)
        seen.add((name, field))
    # every ctypes field was checked against the header
    for name, cls in mirror.items():
        for f in cls._fields_:
            assert (name, f[0]) in seen, (name, f[0])


def test_header_constants_match_the_binding():
    """Every #define PQG_<name> <integer> of include/pqgpu.h has the same value in _lib.py."""
    from polyquant_amd import _lib

    text = open(os.path.join(ROOT, "include", "pqgpu.h")).read()
    consts = dict(re.findall(r"#define\s+(PQG_[A-Z0-9_]+)\s+(\d+)\b", text))
    assert "PQG_SIGMA_LINK_GEMM" in consts and "PQG_MIXED" in consts
    for name, value in consts.items():
        if name.startswith("PQG_NOALLOC"):
            continue
        assert getattr(_lib, name) == int(value), name
