"""GPU parity at the full sizes of the benchmarked workloads (the spaces the solver is timed on).

* water/6-31G full CI (13 orbitals 5a5b, 1,656,369 determinants): FP64 and mixed Davidson energy
  against the oracle's own Davidson root (tests/golden/config_energies.json, written by
  make_config_energies.py), Ritz residual on sampled rows through the oracle's exact sigma;
* electron + positron full CI 48 + 48 orbitals (110,592 determinants): FP64 and FP32 sigma
  against the oracle on strided rows; FP64 and mixed Davidson roots against the oracle's roots,
  the oracle's full sigma of the Ritz vectors (residual and Rayleigh quotients);
* electronic full CI 16 orbitals 8a8b (165,636,900 determinants): FP64 and FP32 sigma of four
  vectors against the oracle's exact sigma on 3000 sampled rows;
* AO->MO transform at n = 320 (4-fold output of 2.638e9 elements, past 32-bit indexing): FP64 and
  BF16x3 against the Cholesky closed form on sampled columns spread over the whole output;
* serving: a prefix of the 512-request Poisson mix in mixed precision, every result equal to the
  same request solved alone (bit for bit) and within 1e-6 Eh of its plain FP64 solve.

Bars: FP64 sigma 1e-12 relative to the largest |sigma|, FP32 sigma 2e-5; FP64 energies 1e-8 Eh,
mixed energies 1e-6 Eh against the FP64 oracle.
"""
import json
import math
import os

import numpy as np
import pytest

from polyquant_amd import workloads as wl

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config_energies.json")


@pytest.fixture(scope="module")
def ctx():
    from polyquant_amd.ci import Context

    c = Context()
    yield c
    c.close()


def _golden(name, prob):
    from tests.golden.make_config_energies import checksum

    g = json.load(open(GOLDEN))[name]
    cs = checksum(prob)
    # the oracle solved these very integrals (up to summation order of the checksum)
    assert abs(cs[0] - g["checksum"][0]) <= 1e-9 * max(1.0, abs(g["checksum"][0]))
    assert abs(cs[1] - g["checksum"][1]) <= 1e-9 * g["checksum"][1]
    return np.array(g["energies"])


def test_water_631g_fci_davidson_energy(ctx):
    """H2O/6-31G full CI: FP64 energy within 1e-8 Eh of the oracle's root, mixed within 1e-6 Eh;
    the FP64 Ritz vector's residual, recomputed with the oracle's exact sigma on every 211th row,
    stays at the solver's tolerance."""
    from oracle import refcpu
    from polyquant_amd._lib import PQG_MIXED
    from polyquant_amd.ci import DetSet

    prob = wl.water_problem("6-31g")
    print("[gpu] case config h2o_631g_fci davidson", flush=True)
    Eref = _golden("h2o_631g_fci", prob)
    ds = DetSet(prob, ctx)
    assert ds.N_dets == 1656369
    E, X, rep = ds.davidson(1, nvec_init=5, tol=1e-6)
    assert rep.converged == 1
    assert abs(E[0] - Eref[0]) < 1e-8, (E, Eref)
    Em, _, rep_m = ds.davidson(1, nvec_init=5, tol=1e-6, precision=PQG_MIXED, want_vectors=False)
    assert rep_m.converged == 1 and rep_m.sigma_calls_fp64 >= 1
    assert abs(Em[0] - Eref[0]) < 1e-6, (Em, Eref)
    oc = refcpu.OracleCI(prob)
    rows = np.arange(0, ds.N_dets, 211)
    R = oc.sigma_rows(X, rows) - E[0] * X[rows]
    assert np.max(np.abs(R)) < 1e-6


def test_ep_fci_48e_48p_sigma_matches_oracle_rows(ctx):
    """ep-FCI 48e (1a1b) + 48p (1 positron): FP64 and FP32 sigma of three vectors against the
    oracle on every 37th row (2,989 rows)."""
    from oracle import refcpu
    from polyquant_amd._lib import PQG_MIXED
    from polyquant_amd.ci import DetSet

    prob = wl.config2()
    print("[gpu] case config ep_fci_48e_48p sigma", flush=True)
    ds = DetSet(prob, ctx)
    assert ds.N_dets == 48 * 48 * 48
    oc = refcpu.OracleCI(prob)
    C = np.random.default_rng(21).standard_normal((ds.N_dets, 3))
    rows = np.arange(0, ds.N_dets, 37)
    ref = oc.sigma_rows(C, rows)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(ds.sigma(C)[rows] - ref)) < 1e-12 * scale
    assert np.max(np.abs(ds.sigma(C, precision=PQG_MIXED)[rows] - ref)) < 2e-5 * scale


def test_ep_fci_48e_48p_davidson_roots(ctx):
    """ep-FCI 48e + 48p, 3 roots with the benchmark's search space (16 start vectors, 64 at most):
    FP64 roots within 1e-8 Eh of the oracle's, mixed within 1e-6 Eh; with the oracle's full
    sigma the Ritz vectors have residual norms below 1e-5 and Rayleigh quotients equal to the
    energies within 1e-9."""
    from oracle import refcpu
    from polyquant_amd._lib import PQG_MIXED
    from polyquant_amd.ci import DetSet

    prob = wl.config2()
    print("[gpu] case config ep_fci_48e_48p davidson", flush=True)
    Eref = _golden("ep_fci_48e_48p", prob)
    ds = DetSet(prob, ctx)
    E, X, rep = ds.davidson(3, nvec_init=16, nvec_max=64, tol=1e-6)
    assert rep.converged == 1
    assert np.max(np.abs(E - Eref)) < 1e-8, (E, Eref)
    Em, _, rep_m = ds.davidson(3, nvec_init=16, nvec_max=64, tol=1e-6, precision=PQG_MIXED, want_vectors=False)
    assert rep_m.converged == 1
    assert np.max(np.abs(Em - Eref)) < 1e-6, (Em, Eref)
    oc = refcpu.OracleCI(prob)
    HX = oc.sigma(X)
    assert np.max(np.linalg.norm(HX - X * E[None, :], axis=0)) < 1e-5
    rq = np.einsum("ik,ik->k", X, HX) / np.einsum("ik,ik->k", X, X)
    assert np.max(np.abs(rq - E)) < 1e-9


def test_fci_16o_8a8b_sigma_matches_oracle_rows(ctx):
    """Full CI 16 orbitals 8a8b (165,636,900 determinants), four device-resident vectors: FP64 and
    FP32 sigma (the link-GEMM path with every coupling, double and single) against the oracle's
    exact sigma on 3000 sampled rows."""
    import torch

    from oracle import refcpu
    from polyquant_amd._lib import PQG_FP64, PQG_MIXED
    from polyquant_amd.ci import DetSet

    prob = wl.config3()
    print("[gpu] case config fci_16o_8a8b sigma", flush=True)
    ds = DetSet(prob, ctx)
    N, k = ds.N_dets, 4
    assert N == math.comb(16, 8) ** 2
    g = torch.Generator(device="cuda").manual_seed(33)
    C = torch.randn(k * N, dtype=torch.float64, device="cuda", generator=g)
    S = torch.empty_like(C)
    rows = np.sort(np.random.default_rng(34).choice(N, 3000, replace=False))
    ridx = torch.from_numpy(rows).cuda()
    ds.sigma_device(C.data_ptr(), N, k, S.data_ptr(), N, PQG_FP64)
    got64 = S.view(k, N)[:, ridx].T.cpu().numpy()
    ds.sigma_device(C.data_ptr(), N, k, S.data_ptr(), N, PQG_MIXED)
    got32 = S.view(k, N)[:, ridx].T.cpu().numpy()
    del S
    Ch = np.asfortranarray(C.view(k, N).T.cpu().numpy())
    del C
    torch.cuda.empty_cache()
    oc = refcpu.OracleCI(prob)
    ref = oc.sigma_rows(Ch, rows)
    scale = np.max(np.abs(ref))
    err64 = np.max(np.abs(got64 - ref))
    err32 = np.max(np.abs(got32 - ref))
    print(f"fci_16o_8a8b: FP64 {err64 / scale:.2e}, FP32 {err32 / scale:.2e} relative", flush=True)
    assert err64 < 1e-12 * scale
    assert err32 < 2e-5 * scale


def test_ao2mo_n320_matches_closed_form():
    """n_ao = n_mo = 320 with 1280 Cholesky vectors: the 8-fold packed input (10.55 GB) and the
    4-fold output (2.638e9 elements, 21.1 GB) on the GPU; 256 sampled output columns, spread over
    the whole matrix (so past element 2^31), against (pq|rs) = sum_k (C^T L_k C)_pq (C^T L_k C)_rs:
    FP64 within 1e-12 of the largest element, BF16x3 within 5e-5."""
    import torch

    from polyquant_amd.ci import Context
    from polyquant_amd.integral import BF16X3, FP64, cholesky_eri8_device, transform_device

    n, nvec = 320, 1280
    npr = n * (n + 1) // 2
    print("[gpu] case config ao2mo n=320", flush=True)
    L, C = wl.config4(n=n, nvec=nvec)
    eri8, Lg = cholesky_eri8_device(L)
    del L
    Cg = torch.from_numpy(np.ascontiguousarray(C.ravel(order="F"))).cuda()
    cols = np.sort(np.random.default_rng(40).choice(npr, 256, replace=False))
    assert (int(cols[-1]) + 1) * npr > 2**31
    # closed form on the sampled columns
    P = torch.from_numpy(wl.pair_index_matrix(n).astype(np.int64)).cuda()
    Cm = Cg.reshape(n, n).T.contiguous()  # row-major C[mu][i]
    ii, jj = np.tril_indices(n)
    flat = torch.from_numpy((ii * n + jj).astype(np.int64)).cuda()
    Lmo = torch.empty(nvec, npr, dtype=torch.float64, device="cuda")
    for k0 in range(0, nvec, 128):
        Lk = Lg[k0:k0 + 128][:, P.reshape(-1)].reshape(-1, n, n)
        Mk = torch.matmul(torch.matmul(Cm.T.contiguous(), Lk), Cm)
        Lmo[k0:k0 + 128] = Mk.reshape(-1, n * n)[:, flat]
        del Lk, Mk
    del Lg
    cidx = torch.from_numpy(cols).cuda()
    ref = (Lmo[:, cidx].T.contiguous() @ Lmo).cpu().numpy()
    del Lmo
    amax = np.max(np.abs(ref))
    ctx = Context(0, hbm_budget=120 << 30)
    out = torch.empty(npr * npr, dtype=torch.float64, device="cuda")
    errs = {}
    for mode, name in ((FP64, "fp64"), (BF16X3, "bf16x3")):
        transform_device(ctx, n, eri8.data_ptr(), Cg.data_ptr(), out.data_ptr(), mode)
        got = torch.stack([out[int(q) * npr:(int(q) + 1) * npr] for q in cols]).cpu().numpy()
        errs[name] = float(np.max(np.abs(got - ref)))
    print(f"ao2mo n=320: max |value| {amax:.3f}, FP64 err {errs['fp64']:.2e}, BF16x3 err {errs['bf16x3']:.2e}",
          flush=True)
    ctx.close()
    assert errs["fp64"] < 1e-12 * amax
    assert errs["bf16x3"] < 5e-5 * amax


def test_serving_prefix_mixed_alone_equals_mix():
    """The first 16 requests of the 512-request Poisson mix (6-13 orbitals, with and without a
    positron, 1-4 roots, up to 2.9e6 determinants), served in mixed precision with continuous
    batching: every result is bit-identical to the same request solved alone in mixed precision,
    and within 1e-6 Eh of its plain FP64 solve."""
    from polyquant_amd._lib import PQG_FP64, PQG_MIXED
    from polyquant_amd.ci import Context, DetSet
    from polyquant_amd.serve import Server

    reqs = wl.serving_requests()[:16]
    print(f"[gpu] case config serving prefix of {len(reqs)} requests", flush=True)
    ctx = Context(hbm_budget=200 << 30)
    opts = dict(nvec_init=8, tol=1e-6, perturb=1e-2)
    srv = Server(ctx, max_slots=16)
    ids = [srv.submit(p, precision=PQG_MIXED, **opts) for _, p in reqs]
    got = srv.drain(len(ids), timeout_s=300)
    srv.close()
    for rid, (_, p) in zip(ids, reqs):
        r = got[rid]
        assert r["status"] == 0, (rid, r["message"])
        ds = DetSet(p, ctx)
        Ea, _, _ = ds.davidson(p.nroots, precision=PQG_MIXED, want_vectors=False, **opts)
        E64, _, _ = ds.davidson(p.nroots, precision=PQG_FP64, want_vectors=False, **opts)
        ds.close()
        assert np.array_equal(r["energies"], Ea), (rid, r["energies"], Ea)
        assert np.max(np.abs(r["energies"] - E64)) < 1e-6, (rid, r["energies"], E64)
    ctx.close()
