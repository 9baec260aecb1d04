"""Writes tests/golden/config_energies.json: the oracle's Davidson roots of two full-size spaces.

The oracle (oracle/refcpu.cpp, the CPU restatement of the reference CI) solves, from the
reference's unit-vector start (epci.cpp:589-608 wiring, plain FP64):
  * water/6-31G full CI, 13 orbitals 5a5b, 1,656,369 determinants, 1 root (RHF orbitals from
    polyquant_amd/molint.py);
  * electron + positron full CI, 48 + 48 orbitals, 1a1b electrons and one positron, 110,592
    determinants, 3 roots (seeded Cholesky integrals, polyquant_amd/workloads.py config2).
Each entry stores the energies (without the constant shift), the residual norms and iteration
count of the oracle's solve, and checksums of the integrals it was given, so a test can tell
that the same inputs are in use.  CPU only, about an hour on 8 threads:

    OMP_NUM_THREADS=8 python tests/golden/make_config_energies.py [water|ep]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
OUT = os.path.join(HERE, "config_energies.json")


def checksum(prob):
    """Sum and sum of squares of every one- and two-body integral block (float64)."""
    s = 0.0
    q = 0.0
    for hs in prob.h:
        for h in hs:
            s += float(np.sum(h))
            q += float(np.sum(h * h))
    for key in sorted(prob.eri):
        v = prob.eri[key]
        s += float(np.sum(v))
        q += float(np.sum(v * v))
    return [s, q]


def solve(name, prob, nev, nvec_init, nvec_max, tol):
    from oracle import refcpu

    t0 = time.perf_counter()
    oc = refcpu.OracleCI(prob)
    ev, X, res, it = oc.davidson(nev, nvec_init=nvec_init, nvec_max=nvec_max, maxit=2000, tol=tol)
    dt = time.perf_counter() - t0
    print(f"{name}: E={ev.tolist()} residuals={res.tolist()} iterations={it} ({dt:.0f} s)", flush=True)
    assert it > 0, f"{name}: oracle Davidson did not converge"
    return {"energies": [float(x) for x in ev], "residuals": [float(x) for x in res], "iterations": int(it),
            "nvec_init": nvec_init, "nvec_max": nvec_max, "tol": tol, "n_dets": int(oc.N),
            "e_shift": float(prob.e_shift), "checksum": checksum(prob), "seconds": round(dt, 1)}


def main():
    from polyquant_amd import workloads as wl

    which = sys.argv[1:] or ["water", "ep"]
    out = json.load(open(OUT)) if os.path.exists(OUT) else {}
    if "water" in which:
        out["h2o_631g_fci"] = solve("h2o_631g_fci", wl.water_problem("6-31g"), 1, 5, 10, 1e-7)
        json.dump(out, open(OUT, "w"), indent=1)
    if "ep" in which:
        out["ep_fci_48e_48p"] = solve("ep_fci_48e_48p", wl.config2(), 3, 16, 64, 1e-7)
        json.dump(out, open(OUT, "w"), indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    sys.exit(main())
