"""Data of the QMCPACK multideterminant dump.

POLYQUANT_CALCULATION::dump_post_mf_for_qmcpack (src/calculation/calculation.cpp:622-657) hands
every determinant's strings, padded with the frozen-core orbitals by
POLYQUANT_DETSET::get_det_withfcorbs (src/ci/detset/determinant_creation.cpp:391-454), to
POLYQUANT_HDF5::dump_post_mf_to_hdf5_for_QMCPACK (src/io/hdf5_utilities.cpp:399-468), which
writes the group /MultiDet:

  CI_{2 part + spin}   uint64 (N_dets, N_int): row i = the padded string of determinant i
                        (address order), its 64-bit words least significant first
  Coeff, Coeff_1, ...   float64 (N_dets,): the CI vector of each state
  NbDet                 N_dets (int32)
  Nbits                 N_int, words per padded string (uint64)
  nexcitedstate         the number of states (int32)
  nstate                the largest active + frozen-core orbital count over the species (int32)

``multidet_arrays`` builds these datasets from a determinant set (the strings and the address
unfold come from the GPU determinant set through the C-ABI); ``save_multidet`` stores them as
an ``.npz`` archive under the same names ("MultiDet/CI_0", ...).  The HDF5 container itself is
not written here: this image has no HDF5 library, and a caller with h5py writes each array of
the dict to the path of the same name.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

_M64 = (1 << 64) - 1


def padded_words(string: int, nfc: int, num_int: int) -> list:
    """The occupation string with its nfc frozen-core orbitals: active orbital o becomes orbital
    o + nfc and orbitals 0 .. nfc - 1 are occupied (the result of get_det_withfcorbs,
    determinant_creation.cpp:391-454), as num_int 64-bit words, most significant first (orbital
    o in word num_int - 1 - (o >> 6), determinant_creation.cpp:63-68)."""
    if not 0 <= nfc <= 64:
        raise ValueError("frozen core must hold 0 to 64 orbitals")
    x = (int(string) << nfc) | ((1 << nfc) - 1)
    if x >> (64 * num_int):
        raise ValueError("the padded string does not fit num_int words")
    return [(x >> (64 * (num_int - 1 - w))) & _M64 for w in range(num_int)]


def multidet_arrays(dets, C: np.ndarray, n_orb: Sequence[int], frozen_core: Optional[Sequence[int]] = None,
                    n_states: Optional[int] = None) -> dict:
    """The /MultiDet datasets for the determinant set ``dets`` (an object with ``strings(part,
    spin)`` in reference order and ``unfold_all()``, e.g. ``ci.DetSet``) and CI vectors C
    (N_dets x states, reference address order).  n_orb: active orbitals per species (the
    reference's max_orb); frozen_core: frozen orbitals per species (default none); n_states:
    the columns of C written (default all)."""
    nsp = len(n_orb)
    fc = [0] * nsp if frozen_core is None else [int(x) for x in frozen_core]
    if len(fc) != nsp:
        raise ValueError("one frozen-core count per species")
    C = np.asarray(C, dtype=np.float64)
    if C.ndim == 1:
        C = C[:, None]
    n_states = C.shape[1] if n_states is None else int(n_states)
    unfold = np.asarray(dets.unfold_all())
    N = unfold.shape[0]
    if C.shape[0] != N or not 1 <= n_states <= C.shape[1]:
        raise ValueError("C must hold N_dets rows and at least n_states columns")
    # words per padded string: from the largest active + frozen orbital count over the species
    # (determinant_creation.cpp:396-403); every species is written with that many
    num_int = (max(n + f for n, f in zip(n_orb, fc)) >> 6) + 1
    out = {}
    for p in range(nsp):
        for s in range(2):
            strs = np.asarray(dets.strings(p, s), dtype=np.uint64)
            table = np.array([padded_words(int(x), fc[p], num_int)[::-1] for x in strs], dtype=np.uint64)
            table = table.reshape(len(strs), num_int)
            out[f"MultiDet/CI_{2 * p + s}"] = np.ascontiguousarray(table[unfold[:, 2 * p + s]])
    for i in range(n_states):
        out["MultiDet/Coeff" if i == 0 else f"MultiDet/Coeff_{i}"] = np.ascontiguousarray(C[:, i])
    # scalar types as the reference passes them (int N_dets, N_states, N_mo; size_t words)
    out["MultiDet/NbDet"] = np.int32(N)
    out["MultiDet/Nbits"] = np.uint64(num_int)
    out["MultiDet/nexcitedstate"] = np.int32(n_states)
    out["MultiDet/nstate"] = np.int32(max(n + f for n, f in zip(n_orb, fc)))
    return out


def save_multidet(path: str, arrays: dict) -> None:
    """The datasets of ``multidet_arrays`` as an uncompressed .npz archive (names kept)."""
    np.savez(path, **arrays)
