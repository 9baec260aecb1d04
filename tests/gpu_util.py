"""Shared helpers for the GPU parity tests."""
import numpy as np


def compare_strings_and_excitations(ds, oc, prob):
    """Bit-exact: strings, address unfold, excitation lists and phases (GPU vs oracle)."""
    from oracle import refcpu

    for p in range(prob.nspecies):
        for s in range(2):
            gs = ds.strings(p, s)
            os_ = oc.strings(p, s)
            assert np.array_equal(gs, os_), (p, s)
            for which in (1, 2):
                off, J, ph = ds.excitations(which, p, s)
                for i in range(len(os_)):
                    ref = oc.excitations(which, p, s, i)
                    got = J[off[i]:off[i + 1]]
                    assert np.array_equal(got, ref), (p, s, which, i)
                    for q, j in enumerate(ref):
                        di, dj = [int(os_[i])], [int(os_[j])]
                        exp = refcpu.get_phase(di, refcpu.get_holes(di, dj), refcpu.get_parts(di, dj))
                        assert ph[off[i] + q] == exp, (p, s, which, i, j)
    assert np.array_equal(ds.unfold_all(), oc.unfold_all())
