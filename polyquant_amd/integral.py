"""Host-side mirror of POLYQUANT_INTEGRAL's MO two-body transform over the HIP library.

``transform_mo_2_body_integrals`` stands in for the reference function of that name
(src/integral/integral.cpp:207-477) for one coefficient matrix: it returns the 4-fold
packed ``n_pair x n_pair`` matrix ``(ij|kl)`` addressed by ``idx2(i,j), idx2(k,l)``
(src/integral/integral.hpp:14-20).  The AO integrals come in 8-fold packed form
(``idx8``, integral.hpp:76-88) instead of being recomputed by libint.
"""
from __future__ import annotations

import ctypes as ct
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, load
from .ci import Context, default_context

FP64 = _lib.PQG_AO2MO_FP64
BF16X3 = _lib.PQG_AO2MO_BF16X3


def pack8(ao_full: np.ndarray) -> np.ndarray:
    """8-fold packed vector of a (n,n,n,n) AO tensor: eri8[idx2(idx2(p,q), idx2(r,s))]."""
    n = ao_full.shape[0]
    ii, jj = np.tril_indices(n)  # pairs in idx2 order
    mat = ao_full[ii[:, None], jj[:, None], ii[None, :], jj[None, :]]
    r, c = np.tril_indices(len(ii))
    return np.ascontiguousarray(mat[r, c])


def transform_mo_2_body_integrals(eri8: np.ndarray, C: np.ndarray, mode: int = FP64,
                                  ctx: Optional[Context] = None):
    """(ij|kl) on the GPU from 8-fold packed AO integrals and C (nao x nmo, nmo == nao).
    Returns (out4, gpu_seconds)."""
    ctx = ctx or default_context()
    n = C.shape[0]
    npair = n * (n + 1) // 2
    eri8 = np.ascontiguousarray(eri8, dtype=np.float64)
    assert eri8.size == npair * (npair + 1) // 2
    Cf = np.asfortranarray(C, dtype=np.float64)
    out = np.zeros((npair, npair), order="F")
    sec = ct.c_double()
    check(load().pqg_ao2mo(ctx.h, n, C.shape[1], eri8.ctypes.data_as(ct.c_void_p), Cf.ctypes.data_as(ct.c_void_p),
                           out.ctypes.data_as(ct.c_void_p), 0, mode, ct.byref(sec)))
    return out, sec.value


def transform_device(ctx: Context, n: int, eri8_ptr: int, C_ptr: int, out_ptr: int, mode: int = FP64) -> float:
    """Same on device buffers (C column-major n x n, out column-major npair x npair)."""
    sec = ct.c_double()
    check(load().pqg_ao2mo(ctx.h, n, n, ct.c_void_p(eri8_ptr), ct.c_void_p(C_ptr), ct.c_void_p(out_ptr), 1, mode,
                           ct.byref(sec)))
    return sec.value
