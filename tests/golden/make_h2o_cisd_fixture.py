"""Writes tests/golden/h2o_sto3g_cisd_ham_elements.npz from the reference's own test data.

Source: tests/data/h2o_sto3glibrary_cisd/cisd_ham_elements.txt of Polyquant v0.7.0 (141 x 141
CISD Hamiltonian elements of H2O/STO-3G in the reference's determinant address order, one
"value phase" row per (i, j), i-major; the diagonal includes the HF determinant energy, see
tests/integration_tests/calculation/calculation_test.cpp:262-300).  Run in a container where the
reference is mounted at /root/reference:

    python tests/golden/make_h2o_cisd_fixture.py
"""
import os
import sys

import numpy as np

SRC = "/root/reference/tests/data/h2o_sto3glibrary_cisd/cisd_ham_elements.txt"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "h2o_sto3g_cisd_ham_elements.npz")


def main():
    rows = np.loadtxt(SRC)
    n = int(round(np.sqrt(rows.shape[0])))
    assert n * n == rows.shape[0] == 19881
    np.savez_compressed(OUT, value=rows[:, 0].reshape(n, n), phase=rows[:, 1].reshape(n, n),
                        e0=-75.01170307729812, e1=-74.59209776692875)
    print(f"wrote {OUT}: {n} x {n}")


if __name__ == "__main__":
    sys.exit(main())
