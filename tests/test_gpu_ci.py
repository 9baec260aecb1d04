"""GPU parity: the HIP CI path against the oracle on the same seeded inputs.

Bars: strings, addresses, excitation lists and phases bit-exact; FP64 matrix elements,
diagonal and sigma within 1e-12 (relative to the vector scale); FP64 Davidson energies
within 1e-8 Eh; FP32 (mixed) sigma within 2e-5 relative.
"""
import numpy as np
import pytest

from polyquant_amd import workloads as wl

pytestmark = pytest.mark.gpu

PROBLEMS = {
    "fci_restricted_7o_3a3b": lambda: wl.small_problem(101, 7, 3, 3),
    "fci_open_shell_8o_4a2b": lambda: wl.small_problem(102, 8, 4, 2),
    "fci_unrestricted_6o_3a2b": lambda: wl.small_problem(103, 6, 3, 2, nspin=2),
    "cisd_h2o_layout_7o_5a5b": lambda: wl.small_problem(104, 7, 5, 5, ex_level=(2, 2, 2)),
    "cis_h2o_layout_7o_5a5b": lambda: wl.small_problem(105, 7, 5, 5, ex_level=(1, 1, 1)),
    "ep_fci_6e_2a1b_5p_1": lambda: wl.small_problem(106, 6, 2, 1, positron=(5, 1, 0)),
    "ep_fci_5e_2a2b_4p_2a1b": lambda: wl.small_problem(107, 5, 2, 2, positron=(4, 2, 1)),
    "ep_truncated_psh_layout": lambda: wl.small_problem(108, 6, 1, 1, positron=(6, 1, 0), ex_level=(2, 2, 2),
                                                       pos_ex_level=(1, 0, 1), max_collective=2),
    "fci_single_electron_9o_1a0b": lambda: wl.small_problem(109, 9, 1, 0),
    # one particle per string: the dense pair-GEMM sigma path
    "ep_fci_1p_strings_9e_1a1b_8p_1": lambda: wl.small_problem(110, 9, 1, 1, positron=(8, 1, 0)),
    "fci_1p_strings_11o_1a1b": lambda: wl.small_problem(111, 11, 1, 1),
    "fci_1p_strings_unrestricted_7o_1a1b": lambda: wl.small_problem(112, 7, 1, 1, nspin=2),
}


@pytest.fixture(scope="module")
def ctx():
    from polyquant_amd.ci import Context

    c = Context()
    yield c
    c.close()


def _pair(name, ctx):
    from oracle import refcpu
    from polyquant_amd.ci import DetSet

    prob = PROBLEMS[name]()
    print(f"[gpu] case {name}", flush=True)
    return prob, DetSet(prob, ctx), refcpu.OracleCI(prob)


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_strings_excitations_phases_bit_exact(name, ctx):
    from tests.gpu_util import compare_strings_and_excitations

    prob, ds, oc = _pair(name, ctx)
    assert ds.N_dets == oc.N
    compare_strings_and_excitations(ds, oc, prob)


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_elements_and_diagonal(name, ctx):
    prob, ds, oc = _pair(name, ctx)
    N = oc.N
    Hor = oc.dense()
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    Hg = ds.elements(ii.ravel(), jj.ravel()).reshape(N, N)
    scale = max(1.0, np.max(np.abs(Hor)))
    assert np.max(np.abs(Hg - Hor)) < 1e-12 * scale
    assert np.max(np.abs(ds.diagonal_Hii - oc.diag())) < 1e-12 * scale


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_sigma_fp64_and_mixed(name, ctx):
    from polyquant_amd._lib import PQG_MIXED

    prob, ds, oc = _pair(name, ctx)
    N = oc.N
    rng = np.random.default_rng(7)
    C = rng.standard_normal((N, 3))
    ref = oc.sigma(C)
    got = ds.sigma(C)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) < 1e-12 * scale
    # reference unit-vector cases (ci_test.cpp:692-853)
    for k in (0, min(5, N - 1)):
        e = np.zeros(N)
        e[k] = 1.0
        assert np.max(np.abs(ds.sigma(e) - oc.sigma(e))) < 1e-12 * max(1.0, scale)
    g32 = ds.sigma(C, precision=PQG_MIXED)
    assert np.max(np.abs(g32 - ref)) < 2e-5 * scale


@pytest.mark.parametrize("name,path", [("fci_restricted_7o_3a3b", "link_gemm"), ("cisd_h2o_layout_7o_5a5b", "enumeration"),
                                       ("fci_1p_strings_11o_1a1b", "pair_gemm"), ("fci_beyond_link_gemm_17o_2a2b", "enumeration")])
def test_sigma_path_reported(name, path, ctx):
    """pqg_ci_sigma_path names the kernel family each space runs; full CI past the link GEMM's
    limits (17 orbitals: 136 column pairs) takes the enumeration kernel and still equals the
    oracle's sigma; the stored explicit H is reported while it is set."""
    problems = dict(PROBLEMS)
    problems["fci_beyond_link_gemm_17o_2a2b"] = lambda: wl.small_problem(113, 17, 2, 2)
    prob = problems[name]()
    print(f"[gpu] case sigma path {name}", flush=True)
    from oracle import refcpu
    from polyquant_amd.ci import DetSet

    ds = DetSet(prob, ctx)
    assert ds.sigma_path == path
    if name.startswith("fci_beyond"):
        oc = refcpu.OracleCI(prob)
        C = np.random.default_rng(9).standard_normal((oc.N, 2))
        ref = oc.sigma(C)
        assert np.max(np.abs(ds.sigma(C) - ref)) < 1e-12 * np.max(np.abs(ref))
    if ds.N_dets <= 2000:
        ds.set_explicit(True)
        assert ds.sigma_path == "stored_h"
        ds.set_explicit(False)
        assert ds.sigma_path == path


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_davidson_fp64_matches_dense(name, ctx):
    prob, ds, oc = _pair(name, ctx)
    nev = min(3, oc.N)
    w = np.linalg.eigvalsh(oc.dense())
    E, X, rep = ds.davidson(nev, nvec_init=min(5, oc.N), tol=1e-8)
    assert rep.converged == 1
    assert np.max(np.abs(E - w[:nev])) < 1e-8
    # Ritz vectors are eigenvectors of the oracle Hamiltonian
    R = oc.sigma(X) - X * E[None, :]
    assert np.max(np.linalg.norm(R, axis=0)) < 1e-7


@pytest.mark.parametrize("name", ["fci_restricted_7o_3a3b", "ep_fci_6e_2a1b_5p_1"])
def test_davidson_mixed_reaches_fp64_energy(name, ctx):
    from polyquant_amd._lib import PQG_MIXED

    prob, ds, oc = _pair(name, ctx)
    w = np.linalg.eigvalsh(oc.dense())
    E, X, rep = ds.davidson(2, nvec_init=5, tol=1e-7, precision=PQG_MIXED)
    assert rep.converged == 1
    assert np.max(np.abs(E - w[:2])) < 1e-8
    assert rep.sigma_calls_fp64 >= 2  # FP64 residual checks happened


@pytest.mark.parametrize("name", ["ep_fci_1p_strings_9e_1a1b_8p_1", "fci_1p_strings_11o_1a1b"])
def test_tensor_path_equals_enumeration_path(name, ctx, monkeypatch):
    """The pair-GEMM sigma for one-particle strings equals the general enumeration kernel."""
    from polyquant_amd._lib import PQG_MIXED
    from polyquant_amd.ci import DetSet

    prob = PROBLEMS[name]()
    C = np.random.default_rng(3).standard_normal((wl.CIProblem.n_dets_fci(prob), 2))
    fast = DetSet(prob, ctx)
    monkeypatch.setenv("PQG_NO_TENSOR_PATH", "1")
    slow = DetSet(prob, ctx)
    a, b = fast.sigma(C), slow.sigma(C)
    assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(b))
    a32 = fast.sigma(C, precision=PQG_MIXED)
    assert np.max(np.abs(a32 - b)) < 2e-5 * np.max(np.abs(b))


def test_water_sto3g_cisd_on_gpu_matches_reference(ctx):
    """H2O/STO-3G CISD through the HIP path: the reference's energies
    (calculation_test.cpp:259-260) and its element fixture (absolute values, MO-sign free)."""
    import os

    from polyquant_amd.ci import DetSet

    prob = wl.water_problem("sto-3g", nroots=2, ex_level=(2, 2, 2))
    print("[gpu] case water_sto3g_cisd", flush=True)
    ds = DetSet(prob, ctx)
    assert ds.N_dets == 141
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "h2o_sto3g_cisd_ham_elements.npz"))
    ii, jj = np.meshgrid(np.arange(141), np.arange(141), indexing="ij")
    Hg = ds.elements(ii.ravel(), jj.ravel()).reshape(141, 141)
    assert np.max(np.abs(np.abs(Hg) - np.abs(fx["value"]))) < 2e-6
    E, X, rep = ds.davidson(2, nvec_init=20, tol=1e-9)
    assert rep.converged == 1
    E = E + prob.e_shift
    assert abs(E[0] - float(fx["e0"])) < 1e-8
    assert abs(E[1] - float(fx["e1"])) < 1e-7  # 6e-8: from the input elements (test_oracle)


def test_water_631g_fci_sigma_matches_oracle_rows(ctx):
    """H2O/6-31G full CI (13 orbitals, 5a5b, 1,656,369 determinants): GPU sigma against the
    oracle on every 97th row, FP64 within 1e-12 relative."""
    from oracle import refcpu
    from polyquant_amd.ci import DetSet

    prob = wl.water_problem("6-31g")
    print("[gpu] case water_631g_fci", flush=True)
    ds = DetSet(prob, ctx)
    assert ds.N_dets == 1656369
    oc = refcpu.OracleCI(prob)
    C = np.random.default_rng(5).standard_normal((ds.N_dets, 2))
    got = ds.sigma(C)
    rows = np.arange(0, ds.N_dets, 97)
    ref = oc.sigma_rows(C, rows)
    assert np.max(np.abs(got[rows] - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("case", ["fci_15o_6a6b_rows_in_place", "ep_12e_4a4b_12p_1_rows_in_place"])
def test_sigma_large_rows_read_in_place(case, ctx):
    """Rows too long to stage in LDS are read in place (L2): FP64 and FP32 sigma against the oracle
    on every 331st row."""
    from oracle import refcpu
    from polyquant_amd._lib import PQG_MIXED
    from polyquant_amd.ci import DetSet

    if case.startswith("fci"):
        prob = wl.small_problem(121, 15, 6, 6)
    else:
        prob = wl.small_problem(122, 12, 4, 4, positron=(12, 1, 0))
    print(f"[gpu] case {case}", flush=True)
    ds = DetSet(prob, ctx)
    oc = refcpu.OracleCI(prob)
    C = np.random.default_rng(9).standard_normal((ds.N_dets, 2))
    rows = np.arange(0, ds.N_dets, 331)
    ref = oc.sigma_rows(C, rows)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(ds.sigma(C)[rows] - ref)) < 1e-12 * scale
    assert np.max(np.abs(ds.sigma(C, precision=PQG_MIXED)[rows] - ref)) < 2e-5 * scale


LINK_GEMM_CASES = {
    "fci_restricted_7o_3a3b": lambda: wl.small_problem(101, 7, 3, 3),
    "fci_open_shell_8o_4a2b": lambda: wl.small_problem(102, 8, 4, 2),
    "fci_unrestricted_6o_3a2b": lambda: wl.small_problem(103, 6, 3, 2, nspin=2),
    "fci_unrestricted_9o_2a4b": lambda: wl.small_problem(114, 9, 2, 4, nspin=2),
    "fci_12o_6a5b": lambda: wl.small_problem(115, 12, 6, 5),
    # with positrons: rows are (electron alpha, positron) strings, positron links in the GEMM
    "ep_fci_6e_2a1b_5p_1": lambda: wl.small_problem(106, 6, 2, 1, positron=(5, 1, 0)),
    "ep_fci_5e_2a2b_4p_2a1b": lambda: wl.small_problem(107, 5, 2, 2, positron=(4, 2, 1)),
    "ep_fci_9e_3a3b_7p_1": lambda: wl.small_problem(116, 9, 3, 3, positron=(7, 1, 0)),
    "ep_fci_unrestricted_7e_3a2b_6p_2a0b": lambda: wl.small_problem(117, 7, 3, 2, nspin=2, positron=(6, 2, 0)),
}


@pytest.mark.parametrize("name", list(LINK_GEMM_CASES))
def test_link_gemm_path_equals_enumeration_path(name, ctx, monkeypatch):
    """One-species full CI: sigma with the alpha-beta couplings and same-spin doubles on MFMA
    (link GEMM) equals the all-enumeration kernel, FP64 to 1e-12 and FP32 to 2e-5 relative,
    for several vectors at once."""
    from polyquant_amd._lib import PQG_MIXED
    from polyquant_amd.ci import DetSet

    prob = LINK_GEMM_CASES[name]()
    print(f"[gpu] case link_gemm {name}", flush=True)
    C = np.random.default_rng(11).standard_normal((prob.n_dets_fci(), 3))
    fast = DetSet(prob, ctx)
    monkeypatch.setenv("PQG_NO_LINK_GEMM", "1")
    slow = DetSet(prob, ctx)
    a, b = fast.sigma(C), slow.sigma(C)
    scale = np.max(np.abs(b))
    assert np.max(np.abs(a - b)) < 1e-12 * scale
    a32 = fast.sigma(C, precision=PQG_MIXED)
    assert np.max(np.abs(a32 - b)) < 2e-5 * scale


def test_davidson_perturbed_start_reaches_ground_state(ctx):
    """Serving request 281 (8 orbitals, 3a3b, 3136 determinants): the reference's unit-vector
    start converges to the second eigenvalue of the dense Hamiltonian (the ground state has no
    weight on the lowest-diagonal determinants); the perturbed start finds the lowest one."""
    from oracle import refcpu
    from polyquant_amd.ci import DetSet

    prob = wl.serving_requests(282, only={281})[281][1]
    print("[gpu] case davidson_perturbed_start req281", flush=True)
    oc = refcpu.OracleCI(prob)
    w = np.linalg.eigvalsh(oc.sigma(np.eye(oc.N)))
    ds = DetSet(prob, ctx)
    E0, _, r0 = ds.davidson(1, nvec_init=8, tol=1e-7, want_vectors=False)
    E1, _, r1 = ds.davidson(1, nvec_init=8, tol=1e-7, want_vectors=False, perturb=1e-2)
    assert r0.converged == 1 and r1.converged == 1
    assert abs(E0[0] - w[1]) < 1e-8  # the unit-vector start lands on the excited root
    assert abs(E1[0] - w[0]) < 1e-8


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_dense_hamiltonian_matches_oracle(name, ctx):
    """create_ham + exact_diag's mirrored copy (explicit_ham.cpp:51-80, epci.cpp:633-643): the
    dense H equals the oracle's Slater_Condon matrix and is exactly symmetric."""
    prob, ds, oc = _pair(name, ctx)
    H = ds.create_ham_dense()
    Hor = oc.dense()
    scale = max(1.0, np.max(np.abs(Hor)))
    assert np.max(np.abs(H - Hor)) < 1e-12 * scale
    assert np.array_equal(H, H.T)
    iu, ju = np.triu_indices(ds.N_dets)
    assert np.array_equal(H[iu, ju], ds.elements(iu, ju))


def test_dense_hamiltonian_device_pointer_padded(ctx):
    """pqg_ci_dense_h into a device buffer with ldh > N equals the host call."""
    import ctypes as ct

    from polyquant_amd import _lib
    from polyquant_amd.ci import DetSet

    hip = ct.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    hip.hipFree.argtypes = [ct.c_void_p]
    D2H = 2
    prob = PROBLEMS["ep_fci_5e_2a2b_4p_2a1b"]()
    print("[gpu] case dense_h device ep_fci_5e_2a2b_4p_2a1b", flush=True)
    ds = DetSet(prob, ctx)
    n = ds.N_dets
    ldh = n + 5
    p = ct.c_void_p()
    assert hip.hipMalloc(ct.byref(p), ldh * n * 8) == 0
    try:
        _lib.check(_lib.load().pqg_ci_dense_h(ds.h, p, ldh, 1))
        out = np.zeros((ldh, n), order="F")
        assert hip.hipMemcpy(out.ctypes.data, p, ldh * n * 8, D2H) == 0
    finally:
        hip.hipFree(p)
    assert np.array_equal(out[:n], ds.create_ham_dense())
    with pytest.raises(_lib.PqgError):
        _lib.check(_lib.load().pqg_ci_dense_h(ds.h, p, n - 1, 1))


def test_exact_diag_water_sto3g_cisd(ctx):
    """EPCI with exact_diag (epci.cpp:633-662) on H2O/STO-3G CISD: the reference's two lowest
    energies (calculation_test.cpp:259-260) and the Davidson roots of the same space."""
    import os

    from polyquant_amd.ci import EPCI

    prob = wl.water_problem("sto-3g", nroots=2, ex_level=(2, 2, 2))
    print("[gpu] case exact_diag water_sto3g_cisd", flush=True)
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "h2o_sto3g_cisd_ham_elements.npz"))
    ci = EPCI(prob, ctx, num_states=2, exact_diag=True)
    E = ci.run()
    assert abs(E[0] - float(fx["e0"])) < 1e-8
    assert abs(E[1] - float(fx["e1"])) < 1e-7  # 6e-8: from the input elements (test_oracle)
    assert ci.C_ci.shape == (141, 2)
    dv = EPCI(prob, ctx, num_states=2, num_subspace_vec=20, convergence_E=1e-10)
    Ed = dv.run()
    assert np.max(np.abs(E - Ed)) < 1e-8
    # the eigenvectors feed the same post-processing as the Davidson ones
    s2 = ci.calculate_S_squared()
    assert np.max(np.abs(s2[:, 0] - np.round(s2[:, 0]))) < 1e-8


def _csr_dense(rowptr, cols, vals, n):
    U = np.zeros((n, n))
    for r in range(n):
        for q in range(rowptr[r], rowptr[r + 1]):
            U[r, cols[q]] += vals[q]
    return U


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_csr_hamiltonian_matches_oracle(name, ctx):
    """create_ham (explicit_ham.cpp:51-84): compressed upper-triangle rows in reference order,
    diagonal first, columns strictly ascending, no explicit zeros off the diagonal, values within
    1e-12 of the oracle's Slater_Condon matrix, every oracle entry above 1e-12 present; the
    symmetric product with random vectors equals the GPU sigma."""
    prob, ds, oc = _pair(name, ctx)
    n = ds.N_dets
    rowptr, cols, vals = ds.create_ham_csr()
    assert rowptr[0] == 0 and rowptr[-1] == len(cols) == len(vals)
    for r in range(n):
        c = cols[rowptr[r]:rowptr[r + 1]]
        assert c[0] == r and np.all(np.diff(c) > 0)
        assert np.all(vals[rowptr[r] + 1:rowptr[r + 1]] != 0.0)
    Hor = oc.dense()
    scale = max(1.0, np.max(np.abs(Hor)))
    U = _csr_dense(rowptr, cols, vals, n)
    assert np.max(np.abs(U - np.triu(Hor))) < 1e-12 * scale
    assert np.all(U[np.abs(np.triu(Hor)) > 1e-12 * scale] != 0.0)
    C = np.random.default_rng(3).standard_normal((n, 2))
    Hs = U + np.triu(U, 1).T
    assert np.max(np.abs(Hs @ C - ds.sigma(C))) < 1e-11 * scale * np.max(np.abs(C)) * max(1, n) ** 0.5


def test_csr_hamiltonian_screening(ctx):
    """ham.prune(screening, 1.0) (explicit_ham.cpp:78-80): every kept entry exceeds the threshold
    and every unscreened entry above it is kept with the same value."""
    prob, ds, oc = _pair("cisd_h2o_layout_7o_5a5b", ctx)
    n = ds.N_dets
    full = _csr_dense(*ds.create_ham_csr(), n)
    s = 0.05
    rowptr, cols, vals = ds.create_ham_csr(s)
    assert np.all(np.abs(vals) > s)
    kept = _csr_dense(rowptr, cols, vals, n)
    assert np.array_equal(kept, np.where(np.abs(full) > s, full, 0.0))


@pytest.mark.parametrize("name", ["fci_open_shell_8o_4a2b", "fci_unrestricted_6o_3a2b", "cisd_h2o_layout_7o_5a5b",
                                  "ep_fci_5e_2a2b_4p_2a1b", "ep_truncated_psh_layout", "fci_1p_strings_11o_1a1b"])
def test_stored_explicit_h_sigma_and_davidson(name, ctx):
    """build_matrix = true: the stored explicit H (pqg_ci_set_explicit) holds both triangles of
    the CSR entries; its sigma equals the matrix-free sigma, its Davidson roots the matrix-free
    ones; mixed precision is refused while it is stored, and switching it off restores the
    matrix-free sigma."""
    from polyquant_amd import _lib

    prob, ds, oc = _pair(name, ctx)
    n = ds.N_dets
    C = np.random.default_rng(9).standard_normal((n, 3))
    S_free = ds.sigma(C)
    E_free, _, _ = ds.davidson(2, tol=1e-9)
    rowptr, cols, vals = ds.create_ham_csr()
    nnz = ds.set_explicit(True)
    assert nnz == 2 * len(cols) - n
    scale = max(1.0, np.max(np.abs(S_free)))
    assert np.max(np.abs(ds.sigma(C) - S_free)) < 1e-12 * scale
    E, _, rep = ds.davidson(2, tol=1e-9)
    assert rep.converged == 1 and np.max(np.abs(E - E_free)) < 1e-8
    with pytest.raises(_lib.PqgError):
        ds.sigma(C, precision=_lib.PQG_MIXED)
    ds.set_explicit(False)
    assert np.max(np.abs(ds.sigma(C, precision=_lib.PQG_MIXED) - S_free)) < 2e-5 * scale


def test_screened_explicit_h_davidson_matches_pruned_oracle(ctx):
    """screening_threshold with build_matrix (explicit_ham.cpp:78-80, epci.cpp:664-682): the
    Davidson roots on the stored, screened H equal the lowest eigenvalues of the oracle's H with
    |H_ij| <= threshold removed, and the exact_diag path with the same threshold."""
    from polyquant_amd.ci import EPCI

    prob, ds, oc = _pair("cisd_h2o_layout_7o_5a5b", ctx)
    s = 0.05
    Hor = oc.dense()
    ref = np.linalg.eigvalsh(np.where(np.abs(Hor) <= s, 0.0, Hor))[:3] + prob.e_shift
    E = EPCI(prob, ctx, num_states=3, num_subspace_vec=10, convergence_E=1e-10, screening_threshold=s).run()
    assert np.max(np.abs(E - ref)) < 1e-8
    Ex = EPCI(prob, ctx, num_states=3, exact_diag=True, screening_threshold=s).run()
    assert np.max(np.abs(Ex - ref)) < 1e-9
    Hfull = np.linalg.eigvalsh(Hor)[:3] + prob.e_shift
    assert np.max(np.abs(Hfull - ref)) > 1e-6  # the threshold changes the answer


@pytest.mark.parametrize("name", ["fci_restricted_7o_3a3b", "ep_fci_6e_2a1b_5p_1"])
def test_stepwise_davidson_equals_whole_solve(name, ctx):
    """pqg_dav_create / pqg_dav_step / pqg_dav_result run the same solve as pqg_ci_davidson: same
    iterations, bit-identical energies and vectors, in FP64 and mixed precision; the result is
    refused before the solve has finished."""
    from polyquant_amd import _lib
    from polyquant_amd._lib import PQG_FP64, PQG_MIXED

    prob, ds, oc = _pair(name, ctx)
    for prec in (PQG_FP64, PQG_MIXED):
        E, X, rep = ds.davidson(2, nvec_init=5, tol=1e-7, precision=prec)
        sv = ds.davidson_solver(2, nvec_init=5, tol=1e-7, precision=prec)
        with pytest.raises(_lib.PqgError):
            sv.result()
        steps = 0
        while not sv.step():
            steps += 1
        E2, X2, rep2 = sv.result()
        sv.close()
        assert rep2.iterations == rep.iterations and steps >= 1
        assert np.array_equal(E, E2) and np.array_equal(X, X2)
