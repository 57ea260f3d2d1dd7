"""GPU checks of the mixed-precision gesv family (SH: split-FP16, SB: split-BF16 MFMA trailing
updates + GMRES-IR with FP64 residuals; SS: plain FP32) through the C ABI.

Bar: backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= 1e-6 (the
refinement target), forward error against the known FP64 solution within cond * 1e-6,
pivots of the refined solves reported, info 0.
"""
import ctypes

import numpy as np
import pytest

import aclsolver as A

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _system(torch, n, cond, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    q1, _ = torch.linalg.qr(torch.randn(n, n, device="cuda", generator=g, dtype=torch.float64))
    q2, _ = torch.linalg.qr(torch.randn(n, n, device="cuda", generator=g, dtype=torch.float64))
    s = torch.logspace(0, np.log10(cond), n, device="cuda", dtype=torch.float64)
    a = ((q1 * s) @ q2.T).float()
    at = a.T.contiguous()
    x_true = torch.randn(n, device="cuda", generator=g, dtype=torch.float64)
    b = (a.double() @ x_true).float()
    return a, at, b, x_true


@pytest.mark.parametrize("name,tournament", [("SB", 0), ("SB", 1), ("SH", 0), ("SH", 1), ("SS", 0)])
@pytest.mark.parametrize("n,cond", [(1024, 1e4), (2048, 1e6), (1100, 1e4), (4700, 1e5)])
def test_gesv_refined(torch, name, tournament, n, cond):
    a, at, b, x_true = _system(torch, n, cond, seed=n)
    with A.Handle() as h:
        L = A.lib()
        L.aclsolverSetRefinement(h.h, ctypes.c_double(1e-6), 30, 30)
        assert L.aclsolverSetTournamentPivoting(h.h, tournament) == A.SUCCESS
        lw = ctypes.c_size_t(0)
        assert getattr(L, f"aclsolver{name}gesv_bufferSize")(h.h, n, 1, ctypes.byref(lw)) == A.SUCCESS
        work = torch.empty(lw.value, dtype=torch.uint8, device="cuda")
        x = torch.empty(n, device="cuda")
        ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        info = torch.full((1,), -5, dtype=torch.int32, device="cuda")
        niter = ctypes.c_int(-100)
        st = getattr(L, f"aclsolver{name}gesv")(h.h, n, 1, A.ptr(at), n, A.ptr(ipiv), A.ptr(b), n, A.ptr(x), n,
                                                A.ptr(work), lw.value, ctypes.byref(niter), A.ptr(info))
        torch.cuda.synchronize()
    assert st == A.SUCCESS and info.item() == 0
    r = b.double() - a.double() @ x.double()
    berr = (r.abs().max() / (a.double().abs().sum(1).max() * x.double().abs().max() + b.double().abs().max())).item()
    ferr = ((x.double() - x_true).abs().max() / x_true.abs().max()).item()
    piv = ipiv.cpu().numpy()
    assert ((piv >= np.arange(1, n + 1)) & (piv <= n)).all()
    if name == "SS":
        assert niter.value == 0
        assert berr <= 10 * n * 2.0 ** -24
    else:
        assert niter.value >= 0, f"refinement did not converge ({niter.value})"
        assert berr <= 1e-6
    assert ferr <= max(cond * 1e-6, 1e-5) * 10


@pytest.mark.parametrize("split", [0, 1, 2])
def test_sbgesv_split_levels(torch, split):
    """SBgesv at each bf16 operand split (0: plain, 1: two parts, 2: three parts) reaches
    the refinement target on a moderately conditioned system; bad levels are rejected."""
    n, cond = 1024, 1e3
    a, at, b, x_true = _system(torch, n, cond, seed=77 + split)
    with A.Handle() as h:
        L = A.lib()
        assert L.aclsolverSetBf16Split(h.h, 3) == A.INVALID_VALUE
        assert L.aclsolverSetBf16Split(h.h, split) == A.SUCCESS
        L.aclsolverSetRefinement(h.h, ctypes.c_double(1e-6), 30, 30)
        lw = ctypes.c_size_t(0)
        assert L.aclsolverSBgesv_bufferSize(h.h, n, 1, ctypes.byref(lw)) == A.SUCCESS
        work = torch.empty(lw.value, dtype=torch.uint8, device="cuda")
        x = torch.empty(n, device="cuda")
        info = torch.full((1,), -5, dtype=torch.int32, device="cuda")
        niter = ctypes.c_int(-100)
        st = L.aclsolverSBgesv(h.h, n, 1, A.ptr(at), n, None, A.ptr(b), n, A.ptr(x), n, A.ptr(work), lw.value,
                               ctypes.byref(niter), A.ptr(info))
        torch.cuda.synchronize()
    assert st == A.SUCCESS and info.item() == 0
    assert niter.value >= 0, f"refinement did not converge ({niter.value})"
    r = b.double() - a.double() @ x.double()
    berr = (r.abs().max() / (a.double().abs().sum(1).max() * x.double().abs().max() + b.double().abs().max())).item()
    assert berr <= 1e-6


def test_sbgesv_butterfly_option(torch):
    """The unpivoted random-butterfly form (opt-in) still reaches the refinement target on a
    moderately conditioned system, and the ipiv it reports is the identity or a valid
    pivot sequence from the pivoted fallback."""
    n, cond = 1024, 1e3
    a, at, b, x_true = _system(torch, n, cond, seed=91)
    with A.Handle() as h:
        L = A.lib()
        assert L.aclsolverSetButterfly(h.h, 1) == A.SUCCESS
        L.aclsolverSetRefinement(h.h, ctypes.c_double(1e-6), 30, 30)
        lw = ctypes.c_size_t(0)
        assert L.aclsolverSBgesv_bufferSize(h.h, n, 1, ctypes.byref(lw)) == A.SUCCESS
        work = torch.empty(lw.value, dtype=torch.uint8, device="cuda")
        x = torch.empty(n, device="cuda")
        ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        info = torch.full((1,), -5, dtype=torch.int32, device="cuda")
        niter = ctypes.c_int(-100)
        st = L.aclsolverSBgesv(h.h, n, 1, A.ptr(at), n, A.ptr(ipiv), A.ptr(b), n, A.ptr(x), n, A.ptr(work),
                               lw.value, ctypes.byref(niter), A.ptr(info))
        torch.cuda.synchronize()
    assert st == A.SUCCESS and info.item() == 0 and niter.value >= 0
    r = b.double() - a.double() @ x.double()
    berr = (r.abs().max() / (a.double().abs().sum(1).max() * x.double().abs().max() + b.double().abs().max())).item()
    assert berr <= 1e-6
    piv = ipiv.cpu().numpy()
    assert ((piv >= np.arange(1, n + 1)) & (piv <= n)).all()
