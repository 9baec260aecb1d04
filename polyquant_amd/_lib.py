"""ctypes binding of libpqgpu.so (the C ABI in include/pqgpu.h).

The product has no CPU fallback: if the HIP library is missing or no GPU is visible, every
call raises.
"""
from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpqgpu.so")

PQG_OK, PQG_ERR_ARG, PQG_ERR_HIP, PQG_ERR_OOM, PQG_ERR_NOCONV, PQG_ERR_UNSUPPORTED = 0, 1, 2, 3, 4, 5
PQG_FP64, PQG_MIXED = 0, 1
PQG_AO2MO_FP64, PQG_AO2MO_BF16X3 = 0, 1
PQG_SIGMA_ENUMERATION, PQG_SIGMA_LINK_GEMM, PQG_SIGMA_PAIR_GEMM, PQG_SIGMA_STORED_H = 0, 1, 2, 3

_dp = ct.POINTER(ct.c_double)
_vp = ct.c_void_p


class CiDesc(ct.Structure):
    _fields_ = [
        ("nspecies", ct.c_int),
        ("n_orb", ct.c_int * 2),
        ("n_alpha", ct.c_int * 2),
        ("n_beta", ct.c_int * 2),
        ("charge", ct.c_double * 2),
        ("nspin_blocks", ct.c_int * 2),
        ("ex_level", (ct.c_int * 3) * 2),
        ("max_collective", ct.c_int),
        ("h", (_dp * 2) * 2),
        ("eri", (((_dp * 2) * 2) * 2) * 2),
    ]


class DavParams(ct.Structure):
    _fields_ = [
        ("nev", ct.c_int),
        ("nvec_init", ct.c_int),
        ("nvec_max", ct.c_int),
        ("maxit", ct.c_int),
        ("tol", ct.c_double),
        ("precision", ct.c_int),
        ("guess_perturb", ct.c_double),
    ]


class DavReport(ct.Structure):
    _fields_ = [
        ("iterations", ct.c_int),
        ("converged", ct.c_int),
        ("sigma_calls", ct.c_int),
        ("sigma_calls_fp64", ct.c_int),
        ("restarts", ct.c_int),
        ("residual", ct.c_double * 64),
        ("seconds", ct.c_double),
    ]


class Result(ct.Structure):
    _fields_ = [
        ("request_id", ct.c_int64),
        ("status", ct.c_int),
        ("nev", ct.c_int),
        ("energies", ct.c_double * 64),
        ("t_submit", ct.c_double),
        ("t_admit", ct.c_double),
        ("t_done", ct.c_double),
        ("iterations", ct.c_int),
        ("message", ct.c_char * 128),
    ]


# (name, restype, argtypes) for every entry point of include/pqgpu.h
SIGNATURES = [
    ("pqg_ctx_create", ct.c_int, [ct.c_int, ct.c_uint64, ct.POINTER(_vp)]),
    ("pqg_ctx_destroy", ct.c_int, [_vp]),
    ("pqg_ctx_memory", ct.c_int, [_vp, ct.POINTER(ct.c_uint64), ct.POINTER(ct.c_uint64), ct.POINTER(ct.c_uint64)]),
    ("pqg_ctx_sync", ct.c_int, [_vp]),
    ("pqg_last_error", ct.c_char_p, []),
    ("pqg_ctx_profile", ct.c_int, [_vp, ct.c_int]),
    ("pqg_ctx_profile_read", ct.c_int, [_vp, ct.c_char_p, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_double),
                                        ct.POINTER(ct.c_int64)]),
    ("pqg_ci_create", ct.c_int, [_vp, ct.POINTER(CiDesc), ct.POINTER(_vp)]),
    ("pqg_ci_destroy", ct.c_int, [_vp]),
    ("pqg_ci_ndets", ct.c_int64, [_vp]),
    ("pqg_ci_nstrings", ct.c_int64, [_vp, ct.c_int, ct.c_int]),
    ("pqg_ci_sigma_path", ct.c_int, [_vp]),
    ("pqg_ci_diag", ct.c_int, [_vp, _vp]),
    ("pqg_ci_elements", ct.c_int, [_vp, _vp, _vp, ct.c_int64, _vp]),
    ("pqg_ci_dense_h", ct.c_int, [_vp, _vp, ct.c_int64, ct.c_int]),
    ("pqg_ci_dense_s2", ct.c_int, [_vp, ct.c_int, ct.c_double, ct.c_int, _vp, ct.c_int64, ct.c_int]),
    ("pqg_ao2mo_pair", ct.c_int, [_vp, ct.c_int, ct.c_int, _vp, _vp, _vp, _vp, ct.c_int, ct.c_int,
                                  ct.POINTER(ct.c_double)]),
    ("pqg_ao2mo_mixed", ct.c_int, [_vp, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _vp, _vp, _vp, _vp, ct.c_int,
                                   ct.c_int, ct.POINTER(ct.c_double)]),
    ("pqg_fc_ints", ct.c_int, [_vp, ct.c_int, ct.c_int, _vp, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_double,
                               _vp, _vp]),
    ("pqg_ci_csr", ct.c_int, [_vp, ct.c_double, _vp, _vp, _vp, ct.c_int64, ct.POINTER(ct.c_int64)]),
    ("pqg_ci_set_explicit", ct.c_int, [_vp, ct.c_int, ct.c_double, ct.POINTER(ct.c_int64)]),
    ("pqg_ci_add_spin_penalty", ct.c_int, [_vp, ct.c_int, ct.c_double, ct.c_int, ct.c_double]),
    ("pqg_ci_sigma", ct.c_int, [_vp, _vp, ct.c_int64, ct.c_int, _vp, ct.c_int64, ct.c_int, ct.c_int]),
    ("pqg_ci_sigma_noalloc", ct.c_int, [_vp, _vp, ct.c_int64, ct.c_int, _vp, ct.c_int64, ct.c_int]),
    ("pqg_ci_s2", ct.c_int, [_vp, _vp, ct.c_int64, ct.c_int, _vp, ct.c_int]),
    ("pqg_ci_rdm1", ct.c_int, [_vp, _vp, ct.c_int64, ct.c_int, ct.c_int, ct.c_int, _vp, ct.c_int]),
    ("pqg_ci_davidson", ct.c_int, [_vp, ct.POINTER(DavParams), _vp, _vp, ct.c_int64, ct.POINTER(DavReport)]),
    ("pqg_dav_create", ct.c_int, [_vp, ct.POINTER(DavParams), ct.POINTER(_vp)]),
    ("pqg_dav_step", ct.c_int, [_vp, ct.POINTER(ct.c_int)]),
    ("pqg_dav_result", ct.c_int, [_vp, _vp, _vp, ct.c_int64, ct.POINTER(DavReport)]),
    ("pqg_dav_destroy", ct.c_int, [_vp]),
    ("pqg_ci_dump_strings", ct.c_int, [_vp, ct.c_int, ct.c_int, _vp]),
    ("pqg_ci_dump_unfold", ct.c_int, [_vp, _vp]),
    ("pqg_ci_dump_excitations", ct.c_int, [_vp, ct.c_int, ct.c_int, ct.c_int, _vp, _vp, _vp]),
    ("pqg_ao2mo", ct.c_int, [_vp, ct.c_int, ct.c_int, _vp, _vp, _vp, ct.c_int, ct.c_int, ct.POINTER(ct.c_double)]),
    ("pqg_serve_create", ct.c_int, [_vp, ct.c_int, ct.POINTER(_vp)]),
    ("pqg_serve_destroy", ct.c_int, [_vp]),
    ("pqg_serve_submit", ct.c_int, [_vp, ct.POINTER(CiDesc), ct.POINTER(DavParams), ct.POINTER(ct.c_int64)]),
    ("pqg_serve_step", ct.c_int, [_vp, ct.POINTER(Result), ct.c_int, ct.POINTER(ct.c_int)]),
    ("pqg_serve_stats", ct.c_int, [_vp, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.POINTER(ct.c_uint64)]),
    ("pqg_serve_clock", ct.c_double, [_vp]),
]

_lib = None


class PqgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pqgpu error {code}: {msg}")
        self.code = code


def load(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `python -m polyquant_amd.build`")
        L = ct.CDLL(path)
        for name, res, args in SIGNATURES:
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def missing_symbols(path: str = LIB_PATH):
    """Entry points of include/pqgpu.h that the built library does not export."""
    L = ct.CDLL(path)
    return [name for name, _, _ in SIGNATURES if getattr(L, name, None) is None]


def check(status: int):
    if status != PQG_OK:
        raise PqgError(status, load().pqg_last_error().decode(errors="replace"))
