"""Serving loop: every result equals the same request solved alone; admission by HBM budget;
failing requests are isolated."""
import numpy as np
import pytest

from polyquant_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _mix(n):
    probs = []
    for r in range(n):
        if r % 3 == 2:
            probs.append(wl.small_problem(500 + r, 6 + r % 3, 2, 1 + r % 2, positron=(5 + r % 2, 1, 0), nroots=1 + r % 3))
        else:
            probs.append(wl.small_problem(500 + r, 7 + r % 4, 3, 2 + r % 2, nroots=1 + r % 3))
    return probs


def test_results_identical_alone_and_in_the_mix():
    from polyquant_amd.ci import Context, DetSet
    from polyquant_amd.serve import Server

    ctx = Context()
    probs = _mix(12)
    alone = []
    for p in probs:
        ds = DetSet(p, ctx)
        E, _, rep = ds.davidson(p.nroots, tol=1e-8)
        alone.append(E)
        ds.close()
    srv = Server(ctx, max_slots=6)
    ids = [srv.submit(p, tol=1e-8) for p in probs]
    got = srv.drain(len(ids))
    for rid, p, Ea in zip(ids, probs, alone):
        r = got[rid]
        assert r["status"] == 0
        assert np.array_equal(r["energies"], Ea), (rid, r["energies"], Ea)
        assert r["t_done"] >= r["t_admit"] >= r["t_submit"]
    srv.close()


def test_admission_by_budget_and_failure_isolation():
    from polyquant_amd.ci import Context
    from polyquant_amd.serve import Server

    ctx = Context(hbm_budget=1 << 30)  # 1 GiB: a few small requests at a time
    srv = Server(ctx, max_slots=4)
    probs = _mix(8)
    ids = [srv.submit(p, tol=1e-8) for p in probs]
    huge = wl.config3()  # needs far more than 1 GiB
    bad_mem = srv.submit(huge)
    bad_conv = srv.submit(wl.small_problem(77, 8, 3, 3), maxit=2, tol=1e-12)
    got = srv.drain(len(ids) + 2)
    assert all(got[i]["status"] == 0 for i in ids), [(i, got[i]["status"], got[i]["message"]) for i in ids]
    assert got[bad_mem]["status"] == 3  # PQG_ERR_OOM
    assert got[bad_conv]["status"] == 4  # PQG_ERR_NOCONV
    q, res, b = srv.stats()
    assert q == 0 and res == 0 and b == 0
    srv.close()
