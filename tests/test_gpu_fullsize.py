"""GPU checks at the headline sizes: the mixed-precision dense solve at n = 16384,
cond 1e6 (SB, SH and SS to a backward error <= 1e-6), the drop-in sgetrf at n = 16384 and
cgetrf at n = 8192 against the oracle on a pivot prefix, with the LU residual in FP64.

Why a prefix: the first k pivots of partial pivoting depend only on the first k columns, so
the oracle factors A[:, :k] (seconds on the CPU) and its ipiv must equal the GPU's ipiv[:k]
on every step before the first pivot tie (oracle margin < 1e-4). The residual
||PA - LU||_1 / (n ||A||_1) <= 10 n eps32 is computed from the full factors as an FP64
(complex128) product on the GPU. Backward errors use
||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) in FP64 on the GPU.
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


def _ipiv_to_perm(piv):
    perm = np.arange(len(piv))
    for k, p in enumerate(piv):
        p -= 1
        if p != k:
            perm[k], perm[p] = perm[p], perm[k]
    return perm


def _lu_residual(torch, a_rows, lu_t, piv, dtype):
    """||P A - L U||_1 / (n ||A||_1) in dtype on the GPU. a_rows / lu_t: device tensors whose
    rows are the matrix columns (column-major storage)."""
    n = a_rows.shape[0]
    perm = torch.from_numpy(_ipiv_to_perm(piv)).cuda()
    a = a_rows.T.to(dtype)
    lu = lu_t.T.to(dtype)
    low = torch.tril(lu, -1) + torch.eye(n, dtype=dtype, device="cuda")
    up = torch.triu(lu)
    del lu
    r = a[perm] - low @ up
    del low, up
    res = (r.abs().sum(0).max() / (n * a.abs().sum(0).max())).item()
    del r, a
    torch.cuda.empty_cache()
    return res


def _prefix_check(piv, ref_piv, margin, min_steps):
    bad = np.nonzero(margin < 1e-4)[0]
    upto = int(bad[0]) if len(bad) else len(ref_piv)
    np.testing.assert_array_equal(piv[:upto], ref_piv[:upto])
    assert upto >= min_steps, f"only {upto} tie-free steps compared"
    return upto


def _qr_system(torch, n, cond, seed):
    """A = Q1 diag(s) Q2^T, s log-spaced 1..cond, Q from QR of seeded N(0,1); b = A x_true."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    q1, _ = torch.linalg.qr(torch.randn(n, n, device="cuda", generator=g))
    q2, _ = torch.linalg.qr(torch.randn(n, n, device="cuda", generator=g))
    s = torch.logspace(0, float(np.log10(cond)), n, device="cuda")
    at = ((q1 * s) @ q2.T).T.contiguous()  # rows of at = columns of A
    del q1, q2
    x_true = torch.randn(n, device="cuda", generator=g, dtype=torch.float64)
    b = (at.T.double() @ x_true).float().contiguous()
    return at, b, x_true


def test_mixed_precision_solve_n16384_cond1e6(torch, handle):
    n, cond = 16384, 1e6
    at, b, x_true = _qr_system(torch, n, cond, seed=16384)
    L = A.lib()
    L.aclsolverSetRefinement(handle.h, ctypes.c_double(1e-6), 20, 60)
    sz = ctypes.c_size_t(0)
    assert L.aclsolverSBgesv_bufferSize(handle.h, n, 1, ctypes.byref(sz)) == A.SUCCESS
    work = torch.empty(sz.value, dtype=torch.uint8, device="cuda")
    x = torch.empty(n, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    a64 = at.T.double()
    anorm = a64.abs().sum(1).max()
    for name in ("SB", "SH", "SS"):
        niter = ctypes.c_int(-100)
        st = getattr(L, f"aclsolver{name}gesv")(handle.h, n, 1, A.ptr(at), n, A.ptr(ipiv), A.ptr(b), n, A.ptr(x), n,
                                                A.ptr(work), sz.value, ctypes.byref(niter), A.ptr(info))
        torch.cuda.synchronize()
        assert st == A.SUCCESS and info.item() == 0, name
        x64 = x.double()
        r = b.double() - a64 @ x64
        berr = (r.abs().max() / (anorm * x64.abs().max() + b.double().abs().max())).item()
        ferr = ((x64 - x_true).abs().max() / x_true.abs().max()).item()
        print(f"{name}: berr {berr:.3e} ferr {ferr:.3e} niter {niter.value}", flush=True)
        assert berr <= 1e-6, (name, berr)
        assert ferr <= cond * 1e-6 * 10, (name, ferr)
        if name != "SS":
            assert niter.value >= 0, (name, niter.value)
        piv = ipiv.cpu().numpy()
        assert ((piv >= np.arange(1, n + 1)) & (piv <= n)).all()
    del a64, work
    torch.cuda.empty_cache()


def test_sgetrf_n16384_prefix_and_residual(torch, handle):
    n, k = 16384, 1024
    rng = np.random.default_rng(163840)
    at_h = rng.standard_normal((n, n), dtype=np.float32) / np.float32(np.sqrt(n))  # rows = columns of A
    at = torch.from_numpy(at_h).cuda()
    d = at.clone()
    L = A.lib()
    lw = ctypes.c_int(0)
    assert L.aclsolverSgetrf_bufferSize(handle.h, n, n, A.ptr(d), n, ctypes.byref(lw)) == A.SUCCESS
    work = torch.zeros(lw.value, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    st = L.aclsolverSgetrf(handle.h, n, n, A.ptr(d), n, A.ptr(work), A.ptr(ipiv), A.ptr(info))
    torch.cuda.synchronize()
    assert st == A.SUCCESS and info.item() == 0
    piv = ipiv.cpu().numpy()
    res = _lu_residual(torch, at, d, piv, torch.float64)
    print(f"sgetrf n={n}: ||PA-LU||_1/(n||A||_1) = {res:.3e}", flush=True)
    assert res <= 10 * n * EPS32
    del d, work
    torch.cuda.empty_cache()
    _, ref_piv, ref_info, margin = O.sgetrf(np.ascontiguousarray(at_h[:k].T), margins=True)
    assert ref_info == 0
    upto = _prefix_check(piv[:k], ref_piv, margin, min_steps=128)
    print(f"sgetrf n={n}: ipiv identical to the oracle over the first {upto} of {k} steps", flush=True)


def test_cgetrf_n8192_prefix_and_residual(torch, handle):
    n, k = 8192, 512
    rng = np.random.default_rng(81920)
    scale = np.float32(np.sqrt(0.5 / n))
    re = rng.standard_normal((n, n), dtype=np.float32) * scale
    im = rng.standard_normal((n, n), dtype=np.float32) * scale
    at_h = (re + 1j * im).astype(np.complex64)  # rows = columns of A
    del re, im
    at = torch.from_numpy(at_h).cuda()
    d = at.clone()
    L = A.lib()
    lw = ctypes.c_int(0)
    assert L.aclsolverCgetrf_bufferSize(handle.h, n, n, A.ptr(d), n, ctypes.byref(lw)) == A.SUCCESS
    work = torch.zeros(max(lw.value, 1), dtype=torch.complex64, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    st = L.aclsolverCgetrf(handle.h, n, n, A.ptr(d), n, A.ptr(work), A.ptr(ipiv), A.ptr(info))
    torch.cuda.synchronize()
    assert st == A.SUCCESS and info.item() == 0
    piv = ipiv.cpu().numpy()
    res = _lu_residual(torch, at, d, piv, torch.complex128)
    print(f"cgetrf n={n}: ||PA-LU||_1/(n||A||_1) = {res:.3e}", flush=True)
    assert res <= 10 * n * EPS32
    del d, work
    torch.cuda.empty_cache()
    _, ref_piv, ref_info, margin = O.cgetrf(np.ascontiguousarray(at_h[:k].T), margins=True)
    assert ref_info == 0
    upto = _prefix_check(piv[:k], ref_piv, margin, min_steps=64)
    print(f"cgetrf n={n}: ipiv identical to the oracle over the first {upto} of {k} steps", flush=True)


def test_complex_solve_n8192_vs_fp64(torch, handle):
    """The complex configuration (n = 8192, 16 RHS, +1.5 on the diagonal): backward error in
    complex128 on the GPU and forward error against a complex128 solve of the same system."""
    n, nrhs = 8192, 16
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    a0 = torch.complex(torch.randn(n, n, device="cuda", generator=g), torch.randn(n, n, device="cuda", generator=g))
    a0 = a0 * (0.5 ** 0.5 / n ** 0.5) + 1.5 * torch.eye(n, device="cuda", dtype=torch.complex64)
    b0 = torch.complex(torch.randn(nrhs, n, device="cuda", generator=g),
                       torch.randn(nrhs, n, device="cuda", generator=g))
    a, b = a0.clone(), b0.clone()
    L = A.lib()
    lw = ctypes.c_int(0)
    assert L.aclsolverCgetrf_bufferSize(handle.h, n, n, A.ptr(a), n, ctypes.byref(lw)) == A.SUCCESS
    work = torch.empty(max(lw.value, 1), dtype=torch.complex64, device="cuda")
    ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    info = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    assert L.aclsolverCgetrf(handle.h, n, n, A.ptr(a), n, A.ptr(work), A.ptr(ipiv), A.ptr(info)) == A.SUCCESS
    assert L.aclsolverCgetrsDevice(handle.h, 0, n, nrhs, A.ptr(a), n, A.ptr(ipiv), A.ptr(b), n, A.ptr(info)) == A.SUCCESS
    torch.cuda.synchronize()
    assert info.item() == 0
    a64 = a0.T.to(torch.complex128)
    x = b.T.to(torch.complex128)
    rhs = b0.T.to(torch.complex128)
    berr = ((rhs - a64 @ x).abs().max() / (a64.abs().sum(1).max() * x.abs().max() + rhs.abs().max())).item()
    x64 = torch.linalg.solve(a64, rhs)
    ferr = ((x - x64).abs().max() / x64.abs().max()).item()
    print(f"cgetrf+cgetrs n={n}: berr {berr:.3e}, error vs complex128 solve {ferr:.3e}", flush=True)
    assert berr <= 10 * n * EPS32
    assert ferr <= 1e-4  # cond(A) is small (diagonal shift 1.5): FP32-level accuracy
