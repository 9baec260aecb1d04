"""GPU parity of the one-particle density matrix (create_1rdm, src/ci/detset/rdm1.cpp:6-79).

Bars: FP64, within 1e-12 of the oracle (normalised vectors, so the scale is 1) on small and
water-sized spaces; at 16 orbitals 8a8b (165,636,900 determinants), where the oracle is too
slow, the identity <C|H1|C> = tr(h_a D_a) + tr(h_b D_b) for a one-body Hamiltonian (GPU sigma,
itself checked against the oracle) within 1e-11, symmetry, and the trace (particle count).
"""
import numpy as np
import pytest

from polyquant_amd import workloads as wl

pytestmark = pytest.mark.gpu

CASES = {
    "fci_restricted_7o_3a3b": lambda: wl.small_problem(201, 7, 3, 3),
    "fci_open_shell_8o_4a2b": lambda: wl.small_problem(202, 8, 4, 2),
    "fci_unrestricted_6o_3a2b": lambda: wl.small_problem(203, 6, 3, 2, nspin=2),
    "cisd_h2o_layout_7o_5a5b": lambda: wl.small_problem(204, 7, 5, 5, ex_level=(2, 2, 2)),
    "cis_h2o_layout_7o_5a5b": lambda: wl.small_problem(205, 7, 5, 5, ex_level=(1, 1, 1)),
    "ep_fci_6e_2a1b_5p_1": lambda: wl.small_problem(206, 6, 2, 1, positron=(5, 1, 0)),
    "ep_fci_5e_2a2b_4p_2a1b": lambda: wl.small_problem(207, 5, 2, 2, positron=(4, 2, 1)),
    "ep_truncated_psh_layout": lambda: wl.small_problem(208, 6, 1, 1, positron=(6, 1, 0), ex_level=(2, 2, 2),
                                                       pos_ex_level=(1, 0, 1), max_collective=2),
    "fci_single_electron_9o_1a0b": lambda: wl.small_problem(209, 9, 1, 0),
    "fci_full_shell_5o_5a2b": lambda: wl.small_problem(210, 5, 5, 2),
    "ep_fci_1p_strings_9e_1a1b_8p_1": lambda: wl.small_problem(211, 9, 1, 1, positron=(8, 1, 0)),
    "fci_12o_6a5b": lambda: wl.small_problem(212, 12, 6, 5),
}


@pytest.fixture(scope="module")
def ctx():
    from polyquant_amd.ci import Context

    c = Context()
    yield c
    c.close()


def _slots(prob):
    return [(p, s) for p in range(prob.nspecies) for s in range(2)]


@pytest.mark.parametrize("name", list(CASES))
def test_rdm1_matches_oracle(name, ctx):
    from oracle import refcpu
    from polyquant_amd.ci import DetSet

    prob = CASES[name]()
    print(f"[gpu] case rdm1 {name}", flush=True)
    ds = DetSet(prob, ctx)
    oc = refcpu.OracleCI(prob)
    assert ds.N_dets == oc.N
    C = np.random.default_rng(5).standard_normal((oc.N, 3))
    C /= np.linalg.norm(C, axis=0)
    for p, s in _slots(prob):
        got = ds.create_1rdm(C, p, s)
        ref = oc.rdm1(C, p, s)
        assert got.shape == ref.shape
        assert np.max(np.abs(got - ref)) < 1e-12, (name, p, s)
        one = ds.create_1rdm(C[:, 1], p, s)
        assert np.max(np.abs(one - ref[1])) < 1e-12


@pytest.mark.parametrize("name", ["fci_open_shell_8o_4a2b", "cisd_h2o_layout_7o_5a5b", "ep_fci_5e_2a2b_4p_2a1b",
                                  "ep_truncated_psh_layout", "fci_12o_6a5b"])
def test_rdm1_grouped_form_equals_per_string_form(name, ctx, monkeypatch):
    """The grouped Gram form (strings sharing nocc - 1 orbitals) and the per-string link form
    give the same matrices."""
    from polyquant_amd.ci import DetSet

    prob = CASES[name]()
    print(f"[gpu] case rdm1 grouped vs links {name}", flush=True)
    grouped = DetSet(prob, ctx)
    C = np.random.default_rng(12).standard_normal((grouped.N_dets, 2))
    C /= np.linalg.norm(C, axis=0)
    # the form is chosen on a handle's first call for each string set
    ga = {ps: grouped.create_1rdm(C, *ps) for ps in _slots(prob)}
    monkeypatch.setenv("PQG_RDM_LINKS", "1")
    links = DetSet(prob, ctx)
    for ps in _slots(prob):
        assert np.max(np.abs(ga[ps] - links.create_1rdm(C, *ps))) < 1e-13, (name, ps)


def test_rdm1_water_fci_matches_oracle(ctx):
    from oracle import refcpu
    from polyquant_amd.ci import DetSet

    prob = wl.water_problem("6-31g")
    print("[gpu] case rdm1 water 6-31g fci", flush=True)
    ds = DetSet(prob, ctx)
    oc = refcpu.OracleCI(prob)
    C = np.random.default_rng(6).standard_normal((oc.N, 2))
    C /= np.linalg.norm(C, axis=0)
    for s in range(2):
        got = ds.create_1rdm(C, 0, s)
        ref = oc.rdm1(C, 0, s)
        assert np.max(np.abs(got - ref)) < 1e-12, s


def test_rdm1_device_pointers_equal_host_call(ctx):
    """on_device = 1: C and D in device memory (allocated with the HIP runtime the library uses)."""
    import ctypes as ct

    from polyquant_amd.ci import DetSet

    hip = ct.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    hip.hipFree.argtypes = [ct.c_void_p]
    H2D, D2H = 1, 2

    def dmalloc(nbytes):
        p = ct.c_void_p()
        assert hip.hipMalloc(ct.byref(p), nbytes) == 0
        return p

    prob = CASES["ep_fci_5e_2a2b_4p_2a1b"]()
    print("[gpu] case rdm1 device pointers", flush=True)
    ds = DetSet(prob, ctx)
    C = np.asfortranarray(np.random.default_rng(8).standard_normal((ds.N_dets, 2)))
    dC = dmalloc(C.nbytes)
    assert hip.hipMemcpy(dC, C.ctypes.data, C.nbytes, H2D) == 0
    try:
        for p, s in _slots(prob):
            n = prob.species[p].n_orb
            out = np.zeros((2, n, n))
            dD = dmalloc(out.nbytes)
            try:
                ds.create_1rdm_device(dC.value, ds.N_dets, 2, p, s, dD.value)
                assert hip.hipMemcpy(out.ctypes.data, dD, out.nbytes, D2H) == 0
            finally:
                hip.hipFree(dD)
            got = out.transpose(0, 2, 1)
            assert np.max(np.abs(got - ds.create_1rdm(C, p, s))) < 1e-13
    finally:
        hip.hipFree(dC)


def test_rdm1_rejects_bad_indices(ctx):
    from polyquant_amd._lib import PqgError
    from polyquant_amd.ci import DetSet

    ds = DetSet(CASES["fci_restricted_7o_3a3b"](), ctx)
    C = np.ones(ds.N_dets)
    with pytest.raises(PqgError):
        ds.create_1rdm(C, 1, 0)  # one species only
    with pytest.raises(PqgError):
        ds.create_1rdm(C, 0, 2)


def test_natural_orbitals_of_ground_state(ctx):
    """calculate_NOs flow: occupations descending in [0, 1], summing to the particle count;
    D V = V diag(occ)."""
    from polyquant_amd.ci import EPCI

    prob = wl.small_problem(213, 7, 3, 3, positron=(5, 1, 0))
    print("[gpu] case natural orbitals", flush=True)
    ci = EPCI(prob, num_states=2, convergence_E=1e-8)
    ci.run()
    nos = ci.calculate_NOs()
    for st in (0, 1):
        for (p, s), (D, occ, vec) in nos[st].items():
            sp = prob.species[p]
            nocc = sp.n_alpha if s == 0 else sp.n_beta
            assert abs(occ.sum() - nocc) < 1e-10
            assert np.all(np.diff(occ) <= 1e-14)
            assert occ.min() > -1e-10 and occ.max() < 1 + 1e-10
            assert np.allclose(D @ vec, vec * occ, atol=1e-10)


def _one_body_problem(n, na, nb, seed):
    """Unrestricted one-body Hamiltonian (all two-body blocks zero)."""
    rng = wl.SplitMix64(seed)
    ha = wl.sym_gaussian(rng, n, 0.5)
    hb = wl.sym_gaussian(rng, n, 0.5)
    npr = wl.npair(n)
    z = np.zeros((npr, npr))
    eri = {(0, 0, 0, 0): z, (0, 1, 0, 1): z, (0, 0, 0, 1): z, (0, 1, 0, 0): z}
    sp = [wl.Species(n, na, nb, -1.0, 2)]
    return wl.CIProblem(species=sp, h=[[ha, hb]], eri=eri, nroots=1, name="one_body"), ha, hb


def test_rdm1_full_size_16o_8a8b_one_body_expectation(ctx):
    """165,636,900 determinants: <C|H1 C> (GPU sigma) = tr(h_a D_a) + tr(h_b D_b)."""
    from polyquant_amd.ci import DetSet

    prob, ha, hb = _one_body_problem(16, 8, 8, wl.SEED_BASE + 77)
    print("[gpu] case rdm1 full size 16o 8a8b", flush=True)
    ds = DetSet(prob, ctx)
    N = ds.N_dets
    assert N == 12870 * 12870
    C = np.random.default_rng(9).standard_normal(N)
    C /= np.linalg.norm(C)
    S = ds.sigma(C)[:, 0]
    e_sigma = float(C @ S)
    del S
    Da = ds.create_1rdm(C, 0, 0)
    Db = ds.create_1rdm(C, 0, 1)
    for D in (Da, Db):
        assert np.max(np.abs(D - D.T)) < 1e-14
        assert abs(np.trace(D) - 8.0) < 1e-11
    e_rdm = float(np.sum(ha * Da) + np.sum(hb * Db))
    assert abs(e_sigma - e_rdm) < 1e-11 * max(1.0, abs(e_sigma)), (e_sigma, e_rdm)
    # linearity in C C^T: D(2C) = 4 D(C)
    assert np.max(np.abs(ds.create_1rdm(2.0 * C, 0, 0) - 4.0 * Da)) < 1e-12
