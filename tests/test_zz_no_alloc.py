"""The allocation-free sigma (pqg_ci_sigma_noalloc) equals pqg_ci_sigma through the C-ABI: two
calls in a row on the link-GEMM path (the second with a smaller input), FP64 and FP32, then the
one-particle-string and enumeration paths; sizes past the fixed scratch are refused."""
import ctypes as ct

import numpy as np
import pytest

from polyquant_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _sigma_pair(ds, C, k, precision):
    """(pqg_ci_sigma on device, pqg_ci_sigma_noalloc) for the device vectors C (N x k)."""
    import torch

    from polyquant_amd._lib import check, load

    N = ds.N_dets
    S_ref = torch.empty_like(C)
    S_new = torch.full_like(C, float("nan"))
    check(load().pqg_ci_sigma(ds.h, ct.c_void_p(C.data_ptr()), N, k, ct.c_void_p(S_ref.data_ptr()), N, 1, precision))
    check(load().pqg_ci_sigma_noalloc(ds.h, ct.c_void_p(C.data_ptr()), N, k, ct.c_void_p(S_new.data_ptr()), N,
                                      precision))
    torch.cuda.synchronize()
    return S_ref.cpu().numpy(), S_new.cpu().numpy()


def _check(ds, k, seed, precision, tol):
    import torch

    g = torch.Generator(device="cuda").manual_seed(seed)
    C = torch.randn(k * ds.N_dets, dtype=torch.float64, device="cuda", generator=g)
    ref, got = _sigma_pair(ds, C, k, precision)
    scale = np.max(np.abs(ref))
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - ref)) <= tol * scale
    return np.max(np.abs(got - ref)) / scale


def test_noalloc_sigma_equals_sigma_two_calls_link_gemm():
    import torch

    from polyquant_amd._lib import PQG_FP64, PQG_MIXED
    from polyquant_amd.ci import Context, DetSet

    torch.zeros(1, device="cuda")
    ctx = Context()
    big = DetSet(wl.small_problem(901, 12, 6, 5, nroots=1), ctx)  # 924 x 792 = 731,808 determinants
    small = DetSet(wl.small_problem(902, 8, 3, 3, nroots=1), ctx)  # 56 x 56 = 3,136
    print(f"[gpu] case noalloc link gemm N={big.N_dets} k=3 then N={small.N_dets} k=1", flush=True)
    assert _check(big, 3, 1, PQG_FP64, 1e-12) == 0.0  # same kernels, same summation order
    assert _check(small, 1, 2, PQG_FP64, 1e-12) == 0.0
    # mixed precision once the handle has its FP32 tables
    assert _check(big, 2, 3, PQG_MIXED, 2e-5) <= 2e-5
    assert _check(small, 1, 4, PQG_MIXED, 2e-5) <= 2e-5
    big.close()
    small.close()
    ctx.close()


def test_noalloc_sigma_other_paths_and_limits():
    import torch

    from polyquant_amd import _lib
    from polyquant_amd._lib import PQG_FP64, PQG_MIXED, check, load
    from polyquant_amd.ci import Context, DetSet

    torch.zeros(1, device="cuda")
    ctx = Context()
    # one-particle strings (pair-GEMM path) and a truncated space (enumeration path)
    for prob in (wl.small_problem(903, 9, 1, 1, positron=(8, 1, 0)),
                 wl.small_problem(904, 7, 5, 5, ex_level=(2, 2, 2))):
        ds = DetSet(prob, ctx)
        print(f"[gpu] case noalloc N={ds.N_dets}", flush=True)
        assert _check(ds, 2, 5, PQG_FP64, 1e-12) == 0.0
        # PQG_MIXED before the handle built its FP32 tables is refused, then works
        C = torch.zeros(ds.N_dets, dtype=torch.float64, device="cuda")
        st = load().pqg_ci_sigma_noalloc(ds.h, ct.c_void_p(C.data_ptr()), ds.N_dets, 1, ct.c_void_p(C.data_ptr()),
                                         ds.N_dets, PQG_MIXED)
        assert st == _lib.PQG_ERR_UNSUPPORTED
        ds.sigma(np.zeros(ds.N_dets), precision=PQG_MIXED)
        assert _check(ds, 1, 6, PQG_MIXED, 2e-5) <= 2e-5
        ds.close()
    # 3432^2 = 11,778,624 determinants x 2 vectors: the copies of C and S exceed the fixed scratch
    ds = DetSet(wl.small_problem(905, 14, 7, 7), ctx)
    N = ds.N_dets
    C = torch.zeros(2 * N, dtype=torch.float64, device="cuda")
    st = load().pqg_ci_sigma_noalloc(ds.h, ct.c_void_p(C.data_ptr()), N, 2, ct.c_void_p(C.data_ptr()), N, PQG_FP64)
    assert st == _lib.PQG_ERR_ARG
    assert b"fixed scratch" in load().pqg_last_error()
    # the refused call left the handle working
    check(load().pqg_ci_sigma(ds.h, ct.c_void_p(C.data_ptr()), N, 1, ct.c_void_p(C.data_ptr() + 8 * N), N, 1,
                              PQG_FP64))
    ds.close()
    ctx.close()
