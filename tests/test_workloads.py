"""Host-side workload generation: determinism, sizes, integral symmetries."""
import math

import numpy as np

from polyquant_amd import workloads as wl


def test_splitmix_known_values():
    # SplitMix64 with seed 0: first outputs (Steele, Lea & Flood 2014 reference stream)
    r = wl.SplitMix64(0).raw(3)
    assert [int(x) for x in r] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_generator_is_deterministic():
    a = wl.config2()
    b = wl.config2()
    for k in a.eri:
        assert np.array_equal(a.eri[k], b.eri[k])
    assert np.array_equal(a.h[0][0], b.h[0][0])


def test_config_sizes():
    assert wl.config2().n_dets_fci() == 48 * 48 * 48
    assert wl.config3().n_dets_fci() == 12870 ** 2 == 165_636_900
    assert math.comb(13, 5) ** 2 == 1_656_369


def test_integral_symmetries():
    p = wl.small_problem(5, 6, 2, 2, positron=(5, 1, 0))
    v = p.eri[(0, 0, 0, 0)]
    assert np.allclose(v, v.T)
    h = p.h[0][0]
    assert np.allclose(h, h.T)
    # Cholesky-built two-body matrices are positive semidefinite
    assert np.linalg.eigvalsh(v).min() > -1e-10
    assert p.eri[(0, 0, 1, 0)].shape == (21, 15)


def test_serving_mix_draw_rules():
    reqs = wl.serving_requests(64)
    ts = [t for t, _ in reqs]
    assert all(b > a for a, b in zip(ts, ts[1:]))
    for _, pr in reqs:
        sp = pr.species[0]
        assert 6 <= sp.n_orb <= 14 and 2 <= sp.n_alpha <= sp.n_orb // 2 and 2 <= sp.n_beta <= sp.n_orb // 2
        assert 1 <= pr.nroots <= 4
        if pr.nspecies == 2:
            assert 6 <= pr.species[1].n_orb <= 12


def test_spin_penalty_host_logic_follows_reference_loop():
    """create_S_sq_penalty (s2.cpp:126-153) host loop: first order adds a S^2 (expected 0),
    second order a (S^2 - expected)^2, and a zero penalty stops every later species (the species
    counter only advances past a nonzero penalty)."""
    import numpy as np

    from polyquant_amd.ci import EPCI

    calls = []

    class FakeDetSet:
        def create_s2_penalty_dense(self, part, expected, power):
            calls.append((part, expected, power))
            return np.full((2, 2), 10.0 ** part)

    class Prob:
        species = [object(), object()]

    def make(kind, pen, exp):
        ci = object.__new__(EPCI)
        ci.prob, ci.detset = Prob(), FakeDetSet()
        ci.first_order_spin_penalty = kind == "first_order"
        ci.second_order_spin_penalty = kind == "second_order"
        ci.spin_penalty, ci.expected_S2 = pen, exp
        return ci

    H = make("first_order", [2.0, 3.0], [5.0, 6.0]).add_spin_penalty(np.zeros((2, 2)))
    assert calls == [(0, 0.0, 1), (1, 0.0, 1)] and np.all(H == 2.0 + 30.0)
    calls.clear()
    H = make("second_order", [2.0, 3.0], [5.0, 6.0]).add_spin_penalty(np.zeros((2, 2)))
    assert calls == [(0, 5.0, 2), (1, 6.0, 2)]
    calls.clear()
    H = make("second_order", [0.0, 3.0], [5.0, 6.0]).add_spin_penalty(np.zeros((2, 2)))
    assert calls == [] and np.all(H == 0.0)


def test_exact_diag_screening_prunes_before_the_penalty():
    """create_ham prunes |H_ij| <= screening_threshold (explicit_ham.cpp:78-80) before
    create_S_sq_penalty adds the penalty (epci.cpp:623-631); the stored explicit H of the
    Davidson path is FP64 only."""
    import numpy as np
    import pytest

    from polyquant_amd.ci import EPCI

    H0 = np.array([[-1.0, 0.01, 0.3], [0.01, -0.5, 0.02], [0.3, 0.02, 0.2]])

    class FakeDetSet:
        def create_ham_dense(self):
            return H0.copy()

        def create_s2_penalty_dense(self, part, expected, power):
            return np.full((3, 3), 0.001)

    class Prob:
        species = [object()]

    ci = object.__new__(EPCI)
    ci.prob, ci.detset = Prob(), FakeDetSet()
    ci.exact_diag, ci.num_states, ci.constant_shift = True, 2, 0.0
    ci.screening_threshold = 0.05
    ci.first_order_spin_penalty, ci.second_order_spin_penalty = True, False
    ci.spin_penalty, ci.expected_S2 = [2.0], [0.0]
    E = ci.run()
    Hp = np.where(np.abs(H0) <= 0.05, 0.0, H0) + 2.0 * 0.001
    assert np.allclose(E, np.linalg.eigvalsh(Hp)[:2], atol=1e-14)
    # the argument checks run before any GPU object is made
    with pytest.raises(ValueError, match="FP64 only"):
        EPCI(Prob(), screening_threshold=0.1, precision=1)
    with pytest.raises(ValueError, match="FP64 only"):
        EPCI(Prob(), spin_penalty_type="first_order", spin_penalty=[1.0], precision=1)
    with pytest.raises(ValueError, match="first_order or second_order"):
        EPCI(Prob(), exact_diag=True, spin_penalty_type="third_order")


def test_frozen_core_ints_host_loop_follows_reference():
    """calculate_frozen_core_ints (integral.cpp:166-205): for each species a and spin block of
    fc_dm[a], one compute_frozen_core_ints call per species b with a frozen core and per spin
    block of fc_dm[b] (so an unrestricted core species contributes once per block also to another
    species), then kinetic - charge * nuclear; a > b blocks use the transposed (b, a) integrals."""
    from types import SimpleNamespace

    import numpy as np

    from polyquant_amd import integral

    calls = []

    def fake(out, eri, dm_b, same, sa, sb, rb, cp, ctx=None):
        calls.append((out.shape[0], eri.shape, len(dm_b), same, sa, sb, rb, cp))
        out += 1.0
        return out

    saved = integral.fc_ints_block
    integral.fc_ints_block = fake
    try:
        sp = [SimpleNamespace(charge=-1.0, restricted=False), SimpleNamespace(charge=1.0, restricted=True)]
        fc_dm = [[np.zeros((3, 3)), np.zeros((3, 3))], [np.zeros((2, 2))]]
        eri = {(0, 0): np.zeros(21), (0, 1): np.zeros((6, 3)), (1, 1): np.zeros(6)}
        kin = [np.eye(3), 2 * np.eye(2)]
        nuc = [np.ones((3, 3)), np.ones((2, 2))]
        out = integral.calculate_frozen_core_ints(sp, fc_dm, [1, 0], eri, kin, nuc)
    finally:
        integral.fc_ints_block = saved
    # species 1 has no frozen core: only species 0 contributes, twice (two spin blocks)
    assert calls == [(3, (21,), 2, True, 0, 0, False, 1.0), (3, (21,), 2, True, 0, 1, False, 1.0),
                     (3, (21,), 2, True, 1, 0, False, 1.0), (3, (21,), 2, True, 1, 1, False, 1.0),
                     (2, (3, 6), 2, False, 0, 0, False, -1.0), (2, (3, 6), 2, False, 0, 1, False, -1.0)]
    assert np.allclose(out[0][0], 2.0 + np.eye(3) + np.ones((3, 3)))
    assert np.allclose(out[1][0], 2.0 + 2 * np.eye(2) - np.ones((2, 2)))


def test_spin_penalty_defaults_follow_epci_setup():
    """POLYQUANT_EPCI::setup (epci.cpp:15-26): with a penalty type and no list, spin_penalty is
    default_spin_penalty = 0.1 per species (epci.hpp:130) and expected_S2 the SCF <S^2> per
    species; without an SCF value the second order needs expected_S2 from the caller.  The checks
    run before any GPU object is made."""
    import pytest

    from polyquant_amd.ci import EPCI, spin_penalty_defaults

    class Prob:
        species = [object(), object()]

    class ProbScf(Prob):
        scf_S2 = [0.75, 2.0]

    assert spin_penalty_defaults(Prob(), None) == ([0.0, 0.0], [0.0, 0.0])
    assert spin_penalty_defaults(Prob(), "first_order") == ([0.1, 0.1], [0.0, 0.0])
    assert spin_penalty_defaults(ProbScf(), "first_order") == ([0.1, 0.1], [0.75, 2.0])
    assert spin_penalty_defaults(ProbScf(), "second_order") == ([0.1, 0.1], [0.75, 2.0])
    assert spin_penalty_defaults(Prob(), "second_order", [0.5, 0.0], [1.0, 2.0]) == ([0.5, 0.0], [1.0, 2.0])
    with pytest.raises(ValueError, match="expected_S2"):
        spin_penalty_defaults(Prob(), "second_order")
    with pytest.raises(ValueError, match="one value per species"):
        spin_penalty_defaults(Prob(), "first_order", [0.3])
    with pytest.raises(ValueError, match="expected_S2"):
        EPCI(Prob(), exact_diag=True, spin_penalty_type="second_order")
