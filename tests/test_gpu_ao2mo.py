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
