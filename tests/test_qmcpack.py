"""QMCPACK multideterminant dump data: the oracle's restatement of get_det_withfcorbs
(determinant_creation.cpp:391-454) pinned by what the padding must mean (the occupied orbitals of
the padded string are the frozen core plus the active orbitals shifted by it), and the product's
assembly (polyquant_amd/qmcpack.py) against the oracle's dump_post_mf_for_qmcpack restatement."""
import random

import numpy as np
import pytest

from oracle import refcpu
from polyquant_amd import qmcpack
from polyquant_amd import workloads as wl


def _orbitals(words):
    """Occupied orbitals of a string given as 64-bit words, most significant first."""
    n = len(words)
    return {64 * (n - 1 - w) + b for w, x in enumerate(words) for b in range(64) if (x >> b) & 1}


def _cases():
    rng = random.Random(20261204)
    out = []
    # (active orbitals of this species, its frozen core, max_orb and frozen_core of all species)
    layouts = [(5, 0, [5], [0]), (7, 1, [7], [1]), (13, 5, [13], [5]), (40, 24, [40], [24]),
               (63, 0, [63], [0]), (63, 1, [63], [1]), (60, 10, [60], [10]), (10, 60, [10], [60]),
               (30, 63, [30], [63]), (10, 0, [10, 8], [0, 60]), (8, 60, [10, 8], [0, 60]),
               (63, 64, [63], [64])]
    for n, nfc, mo, fc in layouts:
        for _ in range(12):
            k = rng.randint(0, n)
            occ = rng.sample(range(n), k)
            out.append((sum(1 << o for o in occ), set(occ), n, nfc, mo, fc))
    return out


def test_oracle_padding_is_core_plus_shifted_active():
    for x, occ, n, nfc, mo, fc in _cases():
        words = refcpu.get_det_withfcorbs([x], nfc, mo, fc)
        num_int = (max(m + f for m, f in zip(mo, fc)) >> 6) + 1
        assert len(words) == num_int, (n, nfc, mo, fc)
        assert all(0 <= w < 1 << 64 for w in words)
        assert _orbitals(words) == set(range(nfc)) | {o + nfc for o in occ}, (x, nfc, mo, fc)
        if nfc == 0 and num_int == 1:
            assert words == [x]  # returned as is (determinant_creation.cpp:405-408)
        if nfc == 0 and num_int == 2:
            assert words == [0, x]  # a zero word in front (:409-413)


def test_oracle_padding_known_words():
    # 0b1011 with two core orbitals: active 0, 1, 3 -> 2, 3, 5, core 0, 1
    assert refcpu.get_det_withfcorbs([0b1011], 2, [10], [2]) == [0b101111]
    # top four bits overflow into a new most significant word
    assert refcpu.get_det_withfcorbs([0xF000000000000001], 4, [63], [4]) == [0xF, 0x1F]


def test_padded_words_matches_oracle():
    for x, occ, n, nfc, mo, fc in _cases():
        num_int = (max(m + f for m, f in zip(mo, fc)) >> 6) + 1
        assert qmcpack.padded_words(x, nfc, num_int) == refcpu.get_det_withfcorbs([x], nfc, mo, fc)
    with pytest.raises(ValueError):
        qmcpack.padded_words(1, 65, 3)
    with pytest.raises(ValueError):
        qmcpack.padded_words(1 << 63, 1, 1)


@pytest.mark.parametrize("fc", [(0, 0), (2, 0), (0, 3), (60, 0)])
def test_multidet_assembly_matches_oracle(fc, tmp_path):
    """Host logic on the oracle's determinant set (strings and unfold in reference order): every
    dataset equal to the restated reference dump, for one and two words per string."""
    prob = wl.small_problem(106, 6, 2, 1, positron=(5, 1, 0))
    oc = refcpu.OracleCI(prob)
    C = np.random.default_rng(5).standard_normal((oc.N, 3))
    n_orb = [sp.n_orb for sp in prob.species]
    got = qmcpack.multidet_arrays(oc, C, n_orb, fc, n_states=2)
    ref = refcpu.qmcpack_multidet(oc, C, n_orb, list(fc), 2)
    assert set(got) == {"MultiDet/" + k for k in ref}
    for k, v in ref.items():
        g = got["MultiDet/" + k]
        if isinstance(v, np.ndarray):
            assert g.dtype == v.dtype and np.array_equal(g, v), k
        else:
            assert int(g) == v, k
    assert got["MultiDet/CI_0"].shape == (oc.N, 2 if fc[0] == 60 else 1)
    # dataset types of the reference's arguments (hdf5_utilities.cpp:399-400, 454-467)
    assert got["MultiDet/NbDet"].dtype == np.int32 and got["MultiDet/nstate"].dtype == np.int32
    assert got["MultiDet/nexcitedstate"].dtype == np.int32 and got["MultiDet/Nbits"].dtype == np.uint64
    path = tmp_path / "multidet.npz"
    qmcpack.save_multidet(str(path), got)
    back = np.load(path)
    for k, v in got.items():
        assert np.array_equal(back[k], v), k
