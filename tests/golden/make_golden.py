"""Writes tests/golden/lu_golden.npz: seeded inputs with host-LAPACK (SciPy/OpenBLAS)
factors and pivots, the oracle library the reference's own tests use
(test/sgetrs/data/gen_data.py:57,68; test/complex_solver/residual_test.py:26).

Inputs follow the reference's generators: integer entries 1..10 from seed 149 made
diagonally dominant and row-shuffled (test/sgetrs/data/gen_data.py:41-50), plus N(0,1)/sqrt(n)
matrices of the BASELINE shapes at small n. Run: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy
import scipy.linalg as sla

out = {}


def add(name, a, kind):
    if kind == "s":
        lu, piv, info = sla.lapack.sgetrf(a.astype(np.float32))
    elif kind == "d":
        lu, piv, info = sla.lapack.dgetrf(a.astype(np.float64))
    else:
        lu, piv, info = sla.lapack.cgetrf(a.astype(np.complex64))
    assert info == 0
    out[name + "_a"] = a
    out[name + "_kind"] = np.array(kind)
    out[name + "_lu"] = lu
    out[name + "_ipiv"] = (piv + 1).astype(np.int32)


np.random.seed(149)
n = 32
A = np.random.randint(1, 11, size=(n, n)).astype(np.float32)
for i in range(n):
    A[i, i] = np.sum(np.abs(A[i, :])) - np.abs(A[i, i]) + np.random.rand() * 10
np.random.shuffle(A)
add("gen149_s32", A, "s")
add("gen149_d32", A.astype(np.float64), "d")

rng = np.random.default_rng(20260512)
for n in (16, 48, 96):
    add(f"normal_s{n}", (rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32), "s")
for n in (16, 40):
    z = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * np.sqrt(0.5) / np.sqrt(n)
    add(f"complex_c{n}", (z + 1.5 * np.eye(n)).astype(np.complex64), "c")
add("tall_s80x24", (rng.standard_normal((80, 24)) / np.sqrt(24)).astype(np.float32), "s")

out["scipy_version"] = np.array(scipy.__version__)
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lu_golden.npz")
np.savez_compressed(path, **out)
print("wrote", path, len(out), "arrays")
