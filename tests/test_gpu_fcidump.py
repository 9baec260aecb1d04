"""An FCIDUMP file as a request: the GPU solve of the file's problem equals that of the problem
it was written from (water 6-31G full CI, 1,656,369 determinants)."""
import numpy as np
import pytest

from polyquant_amd import fcidump
from polyquant_amd import workloads as wl

pytestmark = pytest.mark.gpu


def test_water_fcidump_round_trip_davidson(tmp_path):
    from polyquant_amd.ci import Context, DetSet

    prob = wl.water_problem("6-31g")
    print("[gpu] case fcidump water 6-31g", flush=True)
    path = str(tmp_path / "FCIDUMP")
    fcidump.write_fcidump(path, prob, e_constant=prob.e_shift)
    back = fcidump.read_fcidump(path)
    assert back.e_shift == pytest.approx(prob.e_shift, abs=1e-14)
    ctx = Context()
    a, b = DetSet(prob, ctx), DetSet(back, ctx)
    try:
        E0, _, _ = a.davidson(1, nvec_init=5, tol=1e-7, want_vectors=False)
        E1, _, _ = b.davidson(1, nvec_init=5, tol=1e-7, want_vectors=False)
        assert abs(E0[0] - E1[0]) < 1e-10, (E0, E1)
    finally:
        a.close()
        b.close()
        ctx.close()
