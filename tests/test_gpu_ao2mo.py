"""GPU parity of the AO->MO transform (FP64 MFMA and BF16x3 MFMA) against the oracle's
quarter transform and the Cholesky closed form."""
import numpy as np
import pytest

from polyquant_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _case(n, K, seed):
    rng = np.random.default_rng(seed)
    L = rng.normal(0.0, 0.05, size=(K, n, n))
    L = L + L.transpose(0, 2, 1)
    ao = np.einsum("kpq,krs->pqrs", L, L)
    C = np.linalg.qr(rng.standard_normal((n, n)))[0]
    return L, ao, C


@pytest.mark.parametrize("n", [5, 16, 37, 70])
def test_ao2mo_fp64_matches_oracle(n):
    from oracle import refcpu
    from polyquant_amd.integral import FP64, pack8, transform_mo_2_body_integrals

    print(f"[gpu] ao2mo fp64 n={n}", flush=True)
    L, ao, C = _case(n, 2 * n, n)
    out, sec = transform_mo_2_body_integrals(pack8(ao), C, FP64)
    ref = refcpu.ao2mo(ao, C)
    assert np.max(np.abs(out - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    # closed form: sum_k (C^T L_k C)_ij (C^T L_k C)_kl
    M = np.einsum("pi,kpq,qj->kij", C, L, C)
    ii, jj = np.tril_indices(n)
    Mp = M[:, ii, jj]
    assert np.max(np.abs(out - Mp.T @ Mp)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n", [16, 37, 70])
def test_ao2mo_bf16x3_error_bound(n):
    """BF16x3 split products with FP32 accumulation against the oracle's FP64 quarter transform:
    within 5e-5 of the largest element (and nowhere near exact, so the FP64 path is not taken)."""
    from oracle import refcpu
    from polyquant_amd.integral import BF16X3, pack8, transform_mo_2_body_integrals

    print(f"[gpu] ao2mo bf16x3 n={n}", flush=True)
    L, ao, C = _case(n, 2 * n, 100 + n)
    ref = refcpu.ao2mo(ao, C)
    out, _ = transform_mo_2_body_integrals(pack8(ao), C, BF16X3)
    err = np.max(np.abs(out - ref))
    print(f"bf16x3 n={n} max abs err {err:.3e} (max |value| {np.max(np.abs(ref)):.3e})")
    assert err < 5e-5 * np.max(np.abs(ref))
    assert err > 1e-12 * np.max(np.abs(ref))


def _pair_block(X):
    """(na, na, nb, nb) -> npair_a x npair_b in idx2 order."""
    na, nb = X.shape[0], X.shape[2]
    ia, ja = np.array([(i, j) for i in range(na) for j in range(i + 1)]).T
    ib, jb = np.array([(i, j) for i in range(nb) for j in range(i + 1)]).T
    return np.asfortranarray(X[ia[:, None], ja[:, None], ib[None, :], jb[None, :]])


@pytest.mark.parametrize("na,nma,nb,nmb", [(9, 9, 5, 5), (9, 7, 5, 3), (5, 4, 12, 12), (40, 33, 23, 17)])
def test_ao2mo_two_bases_matches_oracle(na, nma, nb, nmb):
    """pqg_ao2mo_mixed (the electron-positron block, two bases, integral.cpp:207-477 with
    num_ao_a != num_ao_b), rectangular coefficients included: FP64 within 1e-12 of the oracle's
    quarter transform (relative to the largest element), BF16x3 within 5e-5."""
    from oracle import refcpu
    from polyquant_amd.integral import BF16X3, FP64, transform_mo_2_body_integrals_mixed

    print(f"[gpu] ao2mo two bases {na}x{nma} / {nb}x{nmb}", flush=True)
    rng = np.random.default_rng(na * 100 + nb)
    K = 2 * max(na, nb)
    A = rng.normal(0.0, 0.05, size=(K, na, na))
    A = A + A.transpose(0, 2, 1)
    B = rng.normal(0.0, 0.05, size=(K, nb, nb))
    B = B + B.transpose(0, 2, 1)
    ao = np.einsum("kpq,krs->pqrs", A, B)
    Ca = np.linalg.qr(rng.standard_normal((na, na)))[0][:, :nma]
    Cb = np.linalg.qr(rng.standard_normal((nb, nb)))[0][:, :nmb]
    ref = refcpu.ao2mo(ao, Ca, Cb)
    amax = np.max(np.abs(ref))
    got, _ = transform_mo_2_body_integrals_mixed(_pair_block(ao), Ca, Cb, FP64)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) < 1e-12 * amax
    got3, _ = transform_mo_2_body_integrals_mixed(_pair_block(ao), Ca, Cb, BF16X3)
    err3 = np.max(np.abs(got3 - ref))
    assert 1e-12 * amax < err3 < 5e-5 * amax


@pytest.mark.parametrize("n,first,nmo", [(9, 2, 6), (37, 3, 30)])
def test_ao2mo_active_columns_match_oracle(n, first, nmo):
    """pqg_ao2mo with fewer MO columns than basis functions (frozen core and deleted virtuals,
    integral.cpp:214-217): equal to the oracle's transform of the same columns (1e-12) and to the
    active block of the full transform."""
    from oracle import refcpu
    from polyquant_amd.integral import FP64, active_two_body, pack8, transform_mo_2_body_integrals

    print(f"[gpu] ao2mo active columns n={n} first={first} nmo={nmo}", flush=True)
    L, ao, C = _case(n, 2 * n, 300 + n)
    sub = np.asfortranarray(C[:, first:first + nmo])
    got, _ = transform_mo_2_body_integrals(pack8(ao), sub, FP64)
    ref = refcpu.ao2mo(ao, sub)
    amax = np.max(np.abs(ref))
    assert got.shape == (nmo * (nmo + 1) // 2,) * 2
    assert np.max(np.abs(got - ref)) < 1e-12 * amax
    full, _ = transform_mo_2_body_integrals(pack8(ao), C, FP64)
    assert np.max(np.abs(active_two_body(full, n, first, nmo) - got)) < 1e-12 * amax
