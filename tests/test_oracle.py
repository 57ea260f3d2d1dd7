"""Pins the CPU oracle (oracle/lu_oracle.c) before it is trusted as the checker.

Each check is chosen so a plausible mistake in the oracle (wrong tie rule, dropped
update term, transposed solve, wrong inverse order) fails it:
  * host LAPACK (SciPy/OpenBLAS, the reference's own test oracle,
    test/sgetrs/data/gen_data.py:57,68 and test/complex_solver/residual_test.py:26)
    on seeded tie-free inputs: ipiv identical, factors within rounding;
  * the committed golden fixtures written by tests/golden/make_golden.py;
  * the reference's inline known-answer tests (cited per test);
  * closed forms (diagonal inverse, hand-solved triangular systems) and brute force.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "lu_golden.npz")


def _rand(rng, m, n):
    return (rng.standard_normal((m, n)) / np.sqrt(max(n, 1))).astype(np.float32)


@pytest.mark.parametrize("m,n", [(1, 1), (2, 2), (7, 5), (5, 7), (64, 64), (130, 97), (256, 256)])
def test_sgetrf_matches_lapack(m, n):
    rng = np.random.default_rng(1000 + m * 7 + n)
    a = _rand(rng, m, n)
    lu, piv, info, margin = O.sgetrf(a, margins=True)
    ref_lu, ref_piv, ref_info = sla.lapack.sgetrf(a)
    assert info == ref_info == 0
    assert margin.min() > 1e-4  # seeds chosen tie-free
    np.testing.assert_array_equal(piv, ref_piv + 1)
    np.testing.assert_allclose(lu, ref_lu, rtol=0, atol=2e-5 * max(1, n) ** 0.5)
    assert O.lu_backward_error(a, lu, piv) <= 10 * max(m, n) * 2.0 ** -24


@pytest.mark.parametrize("n", [3, 33, 128])
def test_dgetrf_zgetrf_match_lapack(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    lu, piv, info = O.dgetrf(a)
    rl, rp, _ = sla.lapack.dgetrf(a)
    np.testing.assert_array_equal(piv, rp + 1)
    np.testing.assert_allclose(lu, rl, rtol=0, atol=1e-12 * n)
    c = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    zl, zp, _ = O.zgetrf(c)
    # LAPACK izamax uses |re|+|im| too (dcabs1), so pivots must agree.
    rzl, rzp, _ = sla.lapack.zgetrf(c)
    np.testing.assert_array_equal(zp, rzp + 1)
    np.testing.assert_allclose(zl, rzl, rtol=0, atol=1e-12 * n)


def test_pivot_rule_first_max_brute_force():
    # Ties between equal magnitudes go to the first row; brute force over all sign/tie
    # patterns of a 4x1 column (src/utils/kernel/real_lu.hpp:59-77).
    for bits in range(3 ** 4):
        col = [((bits // 3 ** i) % 3) - 1 for i in range(4)]
        a = np.array(col, np.float32).reshape(4, 1)
        _, piv, info = O.sgetrf(a)
        mags = [abs(v) for v in col]
        expect = mags.index(max(mags)) + 1
        assert piv[0] == expect
        assert info == (1 if max(mags) == 0 else 0)


def _fill_input(m, n, b, singular, input_lu=False):
    """FillInput of test/sgetrf/real_lu_test_impl.h:57-72 (row-major transport)."""
    bad = singular and b == 0
    out = np.zeros((m, n), np.float32)
    for r in range(m):
        for c in range(n):
            if r == c:
                out[r, c] = 0.0 if (bad and r == 1) else float(n + r + b + 2)
            elif bad:
                out[r, c] = 0.0
            elif input_lu and r > c:
                out[r, c] = 0.125
            else:
                out[r, c] = 0.25
    return out


@pytest.mark.parametrize("m,n,pivot,singular", [(7, 5, True, False), (3, 3, False, False), (2, 2, True, False),
                                                 (3, 3, True, True)])
def test_reference_basic_cases(m, n, pivot, singular):
    # RunBasicCases, test/sgetrf/real_lu_test_impl.h:306-323: reconstruction within
    # 1e-4*max(1,|expected|), singular case info == 2.
    a = _fill_input(m, n, 0, singular)
    if pivot and m > 1 and not singular:
        a[[0, 1]] = a[[1, 0]]
    lu, piv, info = O.sgetrf(a, pivot=pivot)
    if singular:
        assert info == 2
        return
    assert info == 0
    low, up = O.unpack_lu(lu.astype(np.float64))
    pa = O.apply_ipiv_rows(a.astype(np.float64), piv) if pivot else a.astype(np.float64)
    assert np.all(np.abs(low @ up - pa) <= 1e-4 * np.maximum(1, np.abs(pa)))
    if pivot and (m, n) == (2, 2):
        assert piv[0] == 2  # the pre-swapped 2x2 still selects row 2 (issue note at :315-317)


@pytest.mark.parametrize("m,n,vals", [
    (2, 2, [2e-39, 1e-39, 1e-39, 2e-39]),
    (2, 2, [4.0e-38, 2.0e-38, 2.0e-38, 4.0e-38]),
    (513, 1, [1e-39] * 513),
])
def test_subnormal_pivots(m, n, vals):
    # test/sgetrf/sgetrf_subnormal_pivot_test.cpp:157-169, relative residual <= 1e-5 (:142).
    # The vector is a row-major host transport there; transpose into column-major.
    a = np.array(vals, np.float32).reshape(m, n)
    for divide in (False, True):
        lu, piv, info = O.sgetrf(a, divide=divide)
        assert info == 0
        low, up = O.unpack_lu(lu.astype(np.float64))
        pa = O.apply_ipiv_rows(a.astype(np.float64), piv)
        res = np.abs(low @ up - pa).max() / np.abs(pa).max()
        assert res <= 1e-5


def test_complex_pivot_kats():
    # test/complex_solver/complex_solver_test.cpp:473-491.
    for e in (4, 126, 127):
        v = np.float32(2.0 ** e)
        # column-major {(v,v),(1.5v,1.5v)} then {(v,0),(-v,0)}
        a = np.array([[v + 1j * v, v], [1.5 * v + 1.5j * v, -v]], np.complex64)
        lu, piv, info = O.cgetrf(a)
        assert info == 0 and list(piv) == [2, 2]
        assert np.isfinite(lu).all()
        assert abs(float(lu[1, 0].real) - 2.0 / 3.0) <= 2.0 ** -24
        assert abs(np.ldexp(float(lu[1, 1].real), -e) - 5.0 / 3.0) <= 2.0 ** -22
    tiny = np.float32(np.finfo(np.float32).smallest_subnormal)
    for col, expect in (([tiny, 2 * tiny], 2), ([1 + 2j, 2 + 1j], 1), ([complex(-0.0, -0.0), tiny], 2)):
        a = np.array(col, np.complex64).reshape(2, 1)
        _, piv, info = O.cgetrf(a)
        assert piv[0] == expect


def test_complex_pivot_score_boundary():
    # test/complex_solver/complex_solver_test.cpp:452-471: 65x33, A(0,0) = (1.5*2^127)(1+i)
    # overflows |re|+|im|; A(m-1,0) = FLT_MAX(1+i) must still win.
    m, n = 65, 33
    a = np.zeros((m, n), np.complex64)
    for j in range(1, n):
        a[j, j] = 1
    a[0, 0] = complex(np.float32(1.5 * 2.0 ** 127), np.float32(1.5 * 2.0 ** 127))
    fmax = np.finfo(np.float32).max
    a[m - 1, 0] = complex(fmax, fmax)
    _, piv, _ = O.cgetrf(a)
    assert piv[0] == m
    assert O.cpivot_score(float("nan"), 0.0) == 0
    assert O.cpivot_score(0.0, -0.0) == O.cpivot_score(-0.0, 0.0) == 1


def test_getrs_hand_solved():
    # tests/host/test_solve_lu.py:65-93.
    lu = np.array([[1, 0, 0], [2, 1, 0], [3, 4, 1]], np.float32)
    piv = np.arange(1, 4, dtype=np.int32)
    # L-solve only: use an identity U by solving with lu whose upper part is I.
    l_only = np.tril(lu, -1) + np.eye(3, dtype=np.float32)
    y = O.sgetrs(l_only, piv, np.array([[1], [2], [3]], np.float32))
    np.testing.assert_allclose(y.ravel(), [1, 0, 0], atol=1e-6)
    up = np.array([[2, 1, 3], [0, 2, 1], [0, 0, 1]], np.float32)
    x = O.sgetrs(up, piv, np.array([[2], [2], [1]], np.float32))
    np.testing.assert_allclose(x.ravel(), [-0.75, 0.5, 1.0], atol=1e-6)


@pytest.mark.parametrize("n", [4, 50, 200])
def test_getrs_matches_lapack_all_trans(n):
    rng = np.random.default_rng(n + 5)
    a = _rand(rng, n, n)
    lu, piv, _ = O.sgetrf(a)
    b = rng.standard_normal((n, 3)).astype(np.float32)
    for t in (0, 1):
        x = O.sgetrs(lu, piv, b, trans=t)
        ref = sla.lu_solve((lu.astype(np.float64), piv - 1), b.astype(np.float64), trans=t)
        np.testing.assert_allclose(x, ref, rtol=0, atol=1e-4 * np.abs(ref).max())
    c = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64) / np.sqrt(n)
    clu, cp, _ = O.cgetrf(c)
    bc = (rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))).astype(np.complex64)
    for t in (0, 1, 2):
        x = O.cgetrs(clu, cp, bc, trans=t)
        ref = sla.lu_solve((clu.astype(np.complex128), cp - 1), bc.astype(np.complex128), trans=t)
        np.testing.assert_allclose(x, ref, rtol=0, atol=1e-4 * np.abs(ref).max())


@pytest.mark.parametrize("n", [1, 3, 64, 257])
def test_dgetrs_matches_lapack(n):
    """orc_dgetrs (the FP64 reference of every device getrs comparison) against SciPy's
    LAPACK dgetrs in float64, trans 0 and 1, on pivots from the oracle's own dgetrf and on a
    hand-made pivot sequence with repeated and self interchanges."""
    rng = np.random.default_rng(900 + n)
    a = rng.standard_normal((n, n)) / np.sqrt(n)
    lu, piv, info = O.dgetrf(a)
    assert info == 0
    b = rng.standard_normal((n, 4))
    for t in (0, 1):
        x = O.dgetrs(lu, piv, b, trans=t)
        ref = sla.lu_solve((lu, piv - 1), b, trans=t)
        np.testing.assert_allclose(x, ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))
        # and the solve actually inverts op(A)
        opa = a if t == 0 else a.T
        assert np.abs(opa @ x - b).max() <= 1e-12 * n * max(1.0, np.abs(x).max())
    # a planted transposition mistake (trans 0 result used for trans 1) must fail the pin
    if n >= 3:
        wrong = O.dgetrs(lu, piv, b, trans=0)
        ref1 = sla.lu_solve((lu, piv - 1), b, trans=1)
        assert np.abs(wrong - ref1).max() > 1e-6
    # pivot sequences with repeats (ipiv[k] = n for every k) and identity entries
    if n >= 2:
        piv2 = np.full(n, n, dtype=np.int32)
        piv2[-1] = n
        lu2 = np.tril(rng.standard_normal((n, n)), -1) * 0.1 + np.triu(rng.standard_normal((n, n))) * 0.1
        lu2[np.diag_indices(n)] = 1.0 + np.abs(lu2[np.diag_indices(n)])
        for t in (0, 1):
            x = O.dgetrs(lu2, piv2, b, trans=t)
            ref = sla.lu_solve((lu2, piv2 - 1), b, trans=t)
            np.testing.assert_allclose(x, ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))


def test_inverse_closed_forms():
    # docs/zh/smatinv_batched.md example: diag(2, 4) -> diag(0.5, 0.25), exact.
    c, piv, info = O.smatinv(np.diag([2.0, 4.0]).astype(np.float32))
    assert info == 0
    np.testing.assert_array_equal(c, np.diag([0.5, 0.25]).astype(np.float32))
    # singular: info = first zero pivot
    _, _, info = O.smatinv(np.array([[1, 2], [2, 4]], np.float32))
    assert info == 2
    # complex upper-right KAT of complex_solver_test.cpp:540-552: LU [[1+i, -a],[0, 1]]
    # -> inverse (0,1) entry is a/(1+i).
    for re in (2.0 ** 126, 2.0 ** 127, -(2.0 ** 127)):
        a = complex(re, re)
        lu = np.array([[1 + 1j, -a], [0, 1]], np.complex64)
        inv, info = O.cgetri(lu, None)
        expect = np.complex128(a) / (1 + 1j)
        assert info == 0
        assert abs(inv[0, 1].real - expect.real) <= 4 * abs(np.spacing(np.float32(expect.real)))


@pytest.mark.parametrize("n", [1, 2, 3, 8, 32, 64])
def test_inverse_matches_numpy(n):
    rng = np.random.default_rng(n)
    a = (rng.standard_normal((n, n)) / np.sqrt(n) + 2 * np.eye(n)).astype(np.float32)
    c, piv, info = O.smatinv(a)
    assert info == 0
    np.testing.assert_allclose(c, np.linalg.inv(a.astype(np.float64)), rtol=0, atol=1e-5)
    assert O.inverse_residual(a, c) <= 5
    z = ((rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * np.sqrt(0.5) / np.sqrt(n)
         + 1.5 * np.eye(n)).astype(np.complex64)
    zc, _, info = O.cmatinv(z)
    assert info == 0
    np.testing.assert_allclose(zc, np.linalg.inv(z.astype(np.complex128)), rtol=0, atol=2e-5)
    assert O.inverse_residual(z, zc) <= 5


def test_golden_fixtures():
    g = np.load(GOLDEN)
    for key in [k for k in g.files if k.endswith("_a")]:
        name = key[:-2]
        a = g[key]
        kind = str(g[name + "_kind"])
        if kind == "s":
            lu, piv, info = O.sgetrf(a)
        elif kind == "c":
            lu, piv, info = O.cgetrf(a)
        else:
            lu, piv, info = O.dgetrf(a)
        np.testing.assert_array_equal(piv, g[name + "_ipiv"])
        tol = 1e-12 if kind == "d" else 3e-5
        np.testing.assert_allclose(lu, g[name + "_lu"], rtol=0, atol=tol * np.sqrt(a.shape[0]))


def test_planted_mistakes_are_caught():
    """A wrong tie rule or a dropped update term in the oracle would fail the checks above."""
    rng = np.random.default_rng(3)
    a = _rand(rng, 32, 32)
    lu, piv, _ = O.sgetrf(a)
    # dropped term: zero one multiplier and recompute the residual -> large
    bad = lu.copy()
    bad[20, 3] = 0
    assert O.lu_backward_error(a, bad, piv) > 1e3 * O.lu_backward_error(a, lu, piv)
    # wrong tie rule: a last-max column selects a different row than the checker
    col = np.array([[1.0], [-1.0], [0.5]], np.float32)
    _, p, _ = O.sgetrf(col)
    assert p[0] == 1


# ---------------------------------------------------------------------------------------
# Tournament pivot selection (no reference counterpart: the library's communication-
# avoiding factorization). Pinned by (1) the single-node case, which is partial pivoting;
# (2) a hand-worked two-level tree whose answer differs both from partial pivoting and
# from a tree that stacked factored instead of original rows; (3) planted winners.
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("m,w,seed", [(5, 3, 1), (100, 32, 2), (256, 32, 3), (256, 7, 4), (40, 40, 5)])
def test_tslu_single_node_is_partial_pivoting(m, w, seed):
    rng = np.random.default_rng(seed)
    a = _rand(rng, m, w + 3)
    piv, info = O.stslu_pivots(a, w)
    _, ref, rinfo = O.sgetrf(a)
    np.testing.assert_array_equal(piv, ref[:min(m, w)])
    assert info == 0 and rinfo == 0


def test_tslu_two_level_hand_example():
    """node_rows 4, group 2, w 2 on rows (col0, col1):
    r0 (7,-7) r1 (2,9) r2 (4,-1) r3 (-6,9) | r4 (0,-4) r5 (8,-6) r6 (-5,-6) r7 (9,3).
    Node {r0..r3}: pivot r0, then col1 - (col0/7)(-7) = 11, 3, 3 -> keeps r0, r1.
    Node {r4..r7}: pivot r7, then col1 - col0/3 = -4, -8.67, -4.33 -> keeps r7, r5.
    Root on the ORIGINAL rows r0 r1 r7 r5: pivot r7, then col1 - col0/3 = -9.33, 8.33,
    -8.67 -> r0. Interchanges: r7 from position 7 (ipiv 8), then r0, now at position 7
    (ipiv 8). Partial pivoting over all rows would take r3 (11) second; stacking the
    factored rows would take r1."""
    a = np.array([[7, -7], [2, 9], [4, -1], [-6, 9], [0, -4], [8, -6], [-5, -6], [9, 3]], np.float32)
    blk = np.asfortranarray(a)
    ipiv = np.zeros(2, np.int32)
    info = O.lib().orc_stslu_pivots(8, 2, blk.ctypes.data, 8, 4, 2, ipiv.ctypes.data)
    assert info == 0
    np.testing.assert_array_equal(ipiv, [8, 8])
    _, ref, _ = O.sgetrf(a)
    assert ref[1] != 8  # partial pivoting chooses differently


def test_tslu_planted_winners():
    """Row p_j holds 100 (j + 1) in column j and small entries elsewhere: every node keeps
    the planted rows it holds, so the root's pivot order is p_0, p_1, ..."""
    rng = np.random.default_rng(7)
    m, w = 5000, 32
    a = (rng.standard_normal((m, w)) * 0.01).astype(np.float32)
    rows = rng.choice(m, w, replace=False)
    for j, r in enumerate(rows):
        a[r, j] = 100.0 * (j + 1)
    piv, info = O.stslu_pivots(a, w)
    assert info == 0
    perm = np.arange(m)
    for j, x in enumerate(piv - 1):
        perm[[j, x]] = perm[[x, j]]
    np.testing.assert_array_equal(perm[:w], rows)
