"""GPU parity of the complex64 LU path (cgetrf, cgetrs N/T/C, cgetri, batched inverses)
through the C ABI against the CPU oracle and the reference's complex known answers.

Bars: ipiv identical to the oracle up to the first step whose |re|+|im| tie margin is
< 1e-4; LU residual ||PA - LU||_1 / (n ||A||_1 eps32) <= 5 and inverse residual
||I - C A||_1 / (n ||A|| ||C|| eps32) <= 5 (test/complex_solver/complex_solver_test.cpp:56-103).
"""
import ctypes

import numpy as np
import pytest

import aclsolver as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS32 = 2.0 ** -24


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def handle(torch):
    h = A.Handle()
    yield h
    h.close()


def _crand(rng, m, n, shift=0.0):
    z = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))) * np.sqrt(0.5) / np.sqrt(max(n, 1))
    if shift:
        z = z + shift * np.eye(m, n)
    return z.astype(np.complex64)


def _cgetrf(torch, handle, a, lda=None, pivot=True):
    m, n = a.shape
    lda = lda or max(m, 1)
    d = A.to_device_cm(a, ld=lda)
    lw = ctypes.c_int(0)
    assert A.lib().aclsolverCgetrf_bufferSize(handle.h, m, n, A.ptr(d), lda, ctypes.byref(lw)) == A.SUCCESS
    work = torch.zeros(max(lw.value, 1), dtype=torch.complex64, device="cuda")
    ipiv = torch.zeros(max(min(m, n), 1), dtype=torch.int32, device="cuda") if pivot else None
    info = torch.full((1,), -99, dtype=torch.int32, device="cuda")
    st = A.lib().aclsolverCgetrf(handle.h, m, n, A.ptr(d), lda, A.ptr(work), A.ptr(ipiv), A.ptr(info))
    torch.cuda.synchronize()
    return st, A.from_device_cm(d, m, n), (ipiv[: min(m, n)].cpu().numpy() if pivot else None), int(info.item())


def _lu_residual(a, lu, piv):
    n = max(a.shape)
    return O.lu_backward_error(a, lu, piv) * n / n / EPS32


@pytest.mark.parametrize("m,n,seed", [(64, 64, 1), (300, 300, 2), (129, 129, 3), (1000, 100, 4), (800, 64, 5),
                                      (1000, 40, 6), (5, 9, 7), (1, 1, 8), (700, 700, 9), (2200, 300, 10),
                                      (5000, 140, 11), (9000, 33, 12)])
def test_cgetrf_parity(torch, handle, m, n, seed):
    rng = np.random.default_rng(seed)
    a = _crand(rng, m, n)
    st, lu, piv, info = _cgetrf(torch, handle, a, lda=m + 2)
    assert st == A.SUCCESS and info == 0
    ref_lu, ref_piv, _, margin = O.cgetrf(a, margins=True)
    bad = np.nonzero(margin < 1e-4)[0]
    upto = int(bad[0]) if len(bad) else len(piv)
    np.testing.assert_array_equal(piv[:upto], ref_piv[:upto])
    # complex_solver_test.cpp:56-84: ||PA-LU||_1 / (n ||A||_1 eps) <= 5
    assert O.lu_backward_error(a, lu, piv) / EPS32 <= 5


@pytest.mark.parametrize("m,n,seed", [(5000, 200, 21), (2600, 2600, 22)])
def test_ctall_leaf_exchange_paths_agree(torch, handle, m, n, seed, monkeypatch):
    """The complex tall leaves' two exchange paths (spread over one XCD's L2, the default where
    the device allows it, and the device-coherent path, OPS_SOLVER_LEAF_XCD=0) give identical
    factors, with the oracle's pivots and the reference's residual bar."""
    rng = np.random.default_rng(seed)
    a = _crand(rng, m, n)
    st, lu, piv, info = _cgetrf(torch, handle, a)
    monkeypatch.setenv("OPS_SOLVER_LEAF_XCD", "0")
    st0, lu0, piv0, info0 = _cgetrf(torch, handle, a)
    monkeypatch.delenv("OPS_SOLVER_LEAF_XCD")
    assert st == st0 == A.SUCCESS and info == info0 == 0
    assert np.array_equal(lu.view(np.uint32), lu0.view(np.uint32))
    assert np.array_equal(piv, piv0)
    # the first kp pivots depend only on the first kp columns
    kp = min(m, n, 192)
    _, ref_piv, _, margin = O.cgetrf(np.ascontiguousarray(a[:, :kp]), margins=True)
    bad = np.nonzero(margin < 1e-4)[0]
    upto = int(bad[0]) if len(bad) else kp
    np.testing.assert_array_equal(piv[:upto], ref_piv[:upto])
    assert O.lu_backward_error(a, lu, piv) / EPS32 <= 5


def test_cgetrf_tall_zero_column(torch, handle):
    """Tall panels (leaves spread over several workgroups): an exactly zero column gives
    info at that column, as the reference's first zero pivot."""
    m, n = 4500, 48
    rng = np.random.default_rng(13)
    a = _crand(rng, m, n)
    a[:, 21] = 0
    st, lu, piv, info = _cgetrf(torch, handle, a)
    _, ref_piv, ref_info, margin = O.cgetrf(a, margins=True)
    assert st == A.SUCCESS and info == ref_info == 22
    bad = np.nonzero(margin[:21] < 1e-4)[0]
    upto = int(bad[0]) if len(bad) else 21
    np.testing.assert_array_equal(piv[:upto], ref_piv[:upto])


def test_complex_pivot_kats(torch, handle):
    # complex_solver_test.cpp:473-491 (exponents 4/126/127): ipiv = {2,2}, L21 = 2/3, U22 = 5/3 * 2^e
    for e in (4, 126, 127):
        v = np.float32(2.0 ** e)
        a = np.array([[v + 1j * v, v], [1.5 * v + 1.5j * v, -v]], np.complex64)
        st, lu, piv, info = _cgetrf(torch, handle, a)
        assert st == A.SUCCESS and info == 0 and list(piv) == [2, 2]
        assert np.isfinite(lu).all()
        assert abs(float(lu[1, 0].real) - 2.0 / 3.0) <= 2.0 ** -24
        assert abs(np.ldexp(float(lu[1, 1].real), -e) - 5.0 / 3.0) <= 2.0 ** -22
    tiny = np.float32(np.finfo(np.float32).smallest_subnormal)
    for col, expect in (([tiny, 2 * tiny], 2), ([1 + 2j, 2 + 1j], 1), ([complex(-0.0, -0.0), tiny], 2)):
        a = np.array(col, np.complex64).reshape(2, 1)
        st, _, piv, _ = _cgetrf(torch, handle, a)
        assert piv[0] == expect
    # :452-471 65x33 pivot-score overflow boundary: first pivot = m
    m, n = 65, 33
    a = np.zeros((m, n), np.complex64)
    for j in range(1, n):
        a[j, j] = 1
    a[0, 0] = complex(np.float32(1.5 * 2.0 ** 127), np.float32(1.5 * 2.0 ** 127))
    fmax = np.finfo(np.float32).max
    a[m - 1, 0] = complex(fmax, fmax)
    st, _, piv, info = _cgetrf(torch, handle, a)
    assert piv[0] == m


def test_division_range_kat(torch, handle):
    # complex_solver_test.cpp:520-539: [[1+i],[a]] without pivoting -> L21 = a/(1+i) within 4 ulp
    for real in (2.0 ** 126, 2.0 ** 127, -(2.0 ** 127), float(np.finfo(np.float32).max)):
        for imag in (real, -real, 2.0 ** -126):
            col = np.array([[1 + 1j], [complex(real, imag)]], np.complex64)
            st, lu, _, info = _cgetrf(torch, handle, col, pivot=False)
            assert st == A.SUCCESS and info == 0
            exp = np.complex128(complex(np.float32(real), np.float32(imag))) / (1 + 1j)
            for got, want in ((lu[1, 0].real, exp.real), (lu[1, 0].imag, exp.imag)):
                want32 = np.float32(want)
                assert np.isfinite(got)
                assert abs(np.float64(got) - np.float64(want32)) <= 4 * abs(float(np.spacing(want32)))


@pytest.mark.parametrize("trans", [0, 1, 2])
@pytest.mark.parametrize("n,nrhs", [(6, 2), (300, 5)])
def test_cgetrs_device_and_host(torch, handle, trans, n, nrhs):
    rng = np.random.default_rng(n + trans)
    a = _crand(rng, n, n, shift=1.5)
    b = _crand(rng, n, nrhs)
    lu, piv, _ = O.cgetrf(a)
    op = a if trans == 0 else (a.T if trans == 1 else a.conj().T)
    dA = A.to_device_cm(lu)
    dB = A.to_device_cm(b)
    dpiv = torch.from_numpy(piv).cuda()
    dinfo = torch.full((1,), 9, dtype=torch.int32, device="cuda")
    st = A.lib().aclsolverCgetrsDevice(handle.h, trans, n, nrhs, A.ptr(dA), n, A.ptr(dpiv), A.ptr(dB), n,
                                       A.ptr(dinfo))
    torch.cuda.synchronize()
    assert st == A.SUCCESS and dinfo.item() == 0
    x = A.from_device_cm(dB, n, nrhs)
    assert O.solve_backward_error(op, x, b) <= 10 * n * EPS32
    # host row-major contract
    lda, ldb = n + 3, nrhs + 1
    ha = np.zeros((n, lda), np.complex64)
    ha[:, :n] = lu
    hb = np.zeros((n, ldb), np.complex64)
    hb[:, :nrhs] = b
    hp = piv.astype(np.int32).copy()
    info = np.array([-99], np.int32)
    r = A.lib().aclsolverCgetrs(handle.h, trans, n, nrhs, ha.ctypes.data, lda, hp.ctypes.data, hb.ctypes.data, ldb,
                                info.ctypes.data)
    assert r == 0 and info[0] == 0
    assert O.solve_backward_error(op, hb[:, :nrhs], b) <= 10 * n * EPS32


def _ptr_array(torch, tensors):
    return torch.tensor([t.data_ptr() for t in tensors], dtype=torch.int64, device="cuda")


@pytest.mark.parametrize("n,batch", [(32, 64), (5, 17), (16, 40), (1, 3), (64, 6), (40, 70)])
def test_cmatinv_batched(torch, handle, n, batch):
    rng = np.random.default_rng(n * 31 + batch)
    mats = [_crand(rng, n, n, shift=1.5) for _ in range(batch)]
    dA = [A.to_device_cm(m) for m in mats]
    dC = [torch.zeros((n, n), dtype=torch.complex64, device="cuda") for _ in range(batch)]
    info = torch.full((batch,), -7, dtype=torch.int32, device="cuda")
    pa, pc = _ptr_array(torch, dA), _ptr_array(torch, dC)  # kept alive across the call
    st = A.lib().aclsolverCmatinvBatched(handle.h, n, A.ptr(pa), n, A.ptr(pc), n, A.ptr(info), batch)
    torch.cuda.synchronize()
    assert st == A.SUCCESS and (info.cpu().numpy() == 0).all()
    for m, c in zip(mats, dC):
        inv = A.from_device_cm(c, n, n)
        assert O.inverse_residual(m, inv) <= 5


def test_cmatinv_batched_singular_member(torch, handle):
    """One-wave-per-matrix path (33 <= n <= 64): a member with an exactly zero column reports
    info = that column + 1 (the first zero pivot of the elimination); its neighbours in the
    same workgroup are unaffected."""
    n, batch = 50, 9
    rng = np.random.default_rng(404)
    mats = [_crand(rng, n, n, shift=1.5) for _ in range(batch)]
    mats[5][:, 17] = 0
    dA = [A.to_device_cm(m) for m in mats]
    dC = [torch.zeros((n, n), dtype=torch.complex64, device="cuda") for _ in range(batch)]
    info = torch.full((batch,), -7, dtype=torch.int32, device="cuda")
    pa, pc = _ptr_array(torch, dA), _ptr_array(torch, dC)
    st = A.lib().aclsolverCmatinvBatched(handle.h, n, A.ptr(pa), n, A.ptr(pc), n, A.ptr(info), batch)
    torch.cuda.synchronize()
    got = info.cpu().numpy()
    assert st == A.SUCCESS and got[5] == 18 and (np.delete(got, 5) == 0).all()
    for i, (m, c) in enumerate(zip(mats, dC)):
        if i != 5:
            assert O.inverse_residual(m, A.from_device_cm(c, n, n)) <= 5


@pytest.mark.parametrize("n,batch", [(32, 20), (7, 9), (48, 3), (64, 11)])
def test_cgetri_batched_and_single(torch, handle, n, batch):
    rng = np.random.default_rng(n + 5)
    mats = [_crand(rng, n, n, shift=1.5) for _ in range(batch)]
    facts = [O.cgetrf(m) for m in mats]
    dA = [A.to_device_cm(f[0]) for f in facts]
    piv = torch.from_numpy(np.concatenate([f[1] for f in facts]).astype(np.int32)).cuda()
    dC = [torch.zeros((n, n), dtype=torch.complex64, device="cuda") for _ in range(batch)]
    info = torch.full((batch,), -7, dtype=torch.int32, device="cuda")
    pa, pc = _ptr_array(torch, dA), _ptr_array(torch, dC)
    st = A.lib().aclsolverCgetriBatched(handle.h, n, A.ptr(pa), n, A.ptr(piv), A.ptr(pc), n, A.ptr(info), batch)
    torch.cuda.synchronize()
    assert st == A.SUCCESS and (info.cpu().numpy() == 0).all()
    for m, c in zip(mats, dC):
        assert O.inverse_residual(m, A.from_device_cm(c, n, n)) <= 5
    # single in-place cgetri
    lu, p, _ = facts[0]
    d = A.to_device_cm(lu)
    lw = ctypes.c_int(0)
    assert A.lib().aclsolverCgetri_bufferSize(handle.h, n, A.ptr(d), n, ctypes.byref(lw)) == A.SUCCESS
    work = torch.zeros(lw.value, dtype=torch.complex64, device="cuda")
    i1 = torch.full((1,), -9, dtype=torch.int32, device="cuda")
    st = A.lib().aclsolverCgetri(handle.h, n, A.ptr(d), n, A.ptr(torch.from_numpy(p).cuda()), A.ptr(work), lw.value,
                                 A.ptr(i1))
    torch.cuda.synchronize()
    assert st == A.SUCCESS and i1.item() == 0
    assert O.inverse_residual(mats[0], A.from_device_cm(d, n, n)) <= 5


def test_cgetrf_invalid_shapes(torch, handle):
    info = torch.full((1,), 5, dtype=torch.int32, device="cuda")
    L = A.lib()
    L.aclsolverSetStrictLimits(handle.h, 1)
    # m <= 4096 (src/cgetrf/cgetrf_host.cpp:24-30): devInfo = -1; n <= 8192: -2
    assert L.aclsolverCgetrf(handle.h, 4097, 4, None, 4097, None, None, A.ptr(info)) == A.INVALID_VALUE
    torch.cuda.synchronize()
    assert info.item() == -1
    assert L.aclsolverCgetrf(handle.h, 4, 8193, None, 4, None, None, A.ptr(info)) == A.INVALID_VALUE
    torch.cuda.synchronize()
    assert info.item() == -2
    # cmatinv n <= 32 under the reference limits
    assert L.aclsolverCmatinvBatched(handle.h, 33, None, 33, None, 33, A.ptr(info), 1) == A.INVALID_VALUE
    L.aclsolverSetStrictLimits(handle.h, 0)


@pytest.mark.parametrize("n,nrhs", [(500, 3), (2100, 2)])
def test_cgetrf_cgetrs_forward_error_vs_zgetrf(torch, handle, n, nrhs):
    """The complex64 factor-and-solve against the FP64 oracle (the oracle's zgetrf, solved in
    complex128): relative forward error within 10 n eps32 times the infinity-norm condition
    number of A (the C5 distribution: re/im N(0, 1/2)/sqrt(n) + 1.5 I)."""
    import scipy.linalg as sla

    rng = np.random.default_rng(7000 + n)
    a = ((rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * np.sqrt(0.5 / n)
         + 1.5 * np.eye(n)).astype(np.complex64)
    b = ((rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))) * np.sqrt(0.5)).astype(np.complex64)
    dA = A.to_device_cm(a)
    dB = A.to_device_cm(b)
    lw = ctypes.c_int(0)
    assert A.lib().aclsolverCgetrf_bufferSize(handle.h, n, n, A.ptr(dA), n, ctypes.byref(lw)) == A.SUCCESS
    work = torch.empty(max(lw.value, 1), dtype=torch.complex64, device="cuda")
    ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    info = torch.full((1,), -9, dtype=torch.int32, device="cuda")
    assert A.lib().aclsolverCgetrf(handle.h, n, n, A.ptr(dA), n, A.ptr(work), A.ptr(ipiv), A.ptr(info)) == A.SUCCESS
    assert A.lib().aclsolverCgetrsDevice(handle.h, 0, n, nrhs, A.ptr(dA), n, A.ptr(ipiv), A.ptr(dB), n,
                                         A.ptr(info)) == A.SUCCESS
    torch.cuda.synchronize()
    assert info.item() == 0
    x = A.from_device_cm(dB, n, nrhs).astype(np.complex128)
    lu64, piv64, zinfo = O.zgetrf(a.astype(np.complex128))
    assert zinfo == 0
    x64 = sla.lu_solve((lu64, piv64 - 1), b.astype(np.complex128))
    kappa = np.linalg.cond(a.astype(np.complex128), np.inf)
    ferr = np.abs(x - x64).max() / np.abs(x64).max()
    assert ferr <= 10 * n * EPS32 * kappa, (ferr, kappa)
