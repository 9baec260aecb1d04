"""QMCPACK multideterminant dump from the GPU determinant set: the strings and address unfold
come through the C-ABI (pqg_ci_dump_strings / pqg_ci_dump_unfold) and every /MultiDet dataset
equals the oracle's restatement of dump_post_mf_for_qmcpack (calculation.cpp:622-657,
hdf5_utilities.cpp:399-468), bit for bit."""
import numpy as np
import pytest

from polyquant_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from polyquant_amd.ci import Context

    c = Context()
    yield c
    c.close()


@pytest.mark.parametrize("case", ["ep_truncated_psh_layout", "fci_open_shell_8o_4a2b"])
def test_multidet_from_gpu_detset_matches_oracle(case, ctx):
    from oracle import refcpu
    from polyquant_amd import qmcpack
    from polyquant_amd.ci import DetSet

    prob = {"ep_truncated_psh_layout": lambda: wl.small_problem(108, 6, 1, 1, positron=(6, 1, 0), ex_level=(2, 2, 2),
                                                                pos_ex_level=(1, 0, 1), max_collective=2),
            "fci_open_shell_8o_4a2b": lambda: wl.small_problem(102, 8, 4, 2)}[case]()
    print(f"[gpu] case qmcpack multidet {case}", flush=True)
    ds = DetSet(prob, ctx)
    oc = refcpu.OracleCI(prob)
    assert ds.N_dets == oc.N
    E, X, rep = ds.davidson(2, nvec_init=6, tol=1e-8)
    n_orb = [sp.n_orb for sp in prob.species]
    for fc in ([0] * len(n_orb), [3] + [0] * (len(n_orb) - 1), [58] + [0] * (len(n_orb) - 1)):
        got = qmcpack.multidet_arrays(ds, X, n_orb, fc)
        ref = refcpu.qmcpack_multidet(oc, X, n_orb, fc, 2)
        for k, v in ref.items():
            g = got["MultiDet/" + k]
            if isinstance(v, np.ndarray):
                assert g.dtype == v.dtype and np.array_equal(g, v), (fc, k)
            else:
                assert int(g) == v, (fc, k)


def test_epci_dump_post_mf_for_qmcpack(ctx, tmp_path):
    """EPCI.dump_post_mf_for_qmcpack after a solve: the saved archive holds the oracle's datasets
    for the solve's own CI vectors, one coefficient set per state."""
    from oracle import refcpu
    from polyquant_amd.ci import EPCI

    prob = wl.small_problem(107, 5, 2, 2, positron=(4, 2, 1))
    print("[gpu] case qmcpack EPCI dump", flush=True)
    ci = EPCI(prob, ctx, num_states=2, num_subspace_vec=6, convergence_E=1e-8)
    ci.run()
    path = str(tmp_path / "multidet.npz")
    got = ci.dump_post_mf_for_qmcpack(path, frozen_core=[1, 0])
    ref = refcpu.qmcpack_multidet(refcpu.OracleCI(prob), ci.C_ci, [5, 4], [1, 0], 2)
    back = np.load(path)
    assert sorted(back.files) == sorted(got)
    for k, v in ref.items():
        g = back["MultiDet/" + k]
        if isinstance(v, np.ndarray):
            assert np.array_equal(g, v), k
        else:
            assert int(g) == v, k
