"""Tournament-pivoting LU (aclsolverSgetrfTournament; the tall-panel leaves of SH/SB).

Pinned against the oracle's restatement of the tournament (oracle stslu_pivots: 256-row
blocks, 8 children per tree node, partial pivoting with the blocked scaling rule inside
every node): the first leaf's 32 pivots depend only on the original first 32 columns and
the node arithmetic is the oracle's (fused updates, rounded reciprocal), so ipiv[:32] must
match bit for bit. The whole factorization is checked by its residual
||P A - L U||_1 / (n ||A||_1) <= 10 n eps32 in FP64 on the GPU, a growth bound on L, and,
where no panel is taller than 1024 rows, bit-identity with aclsolverSgetrf.
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


def _ipiv_to_perm(piv, m):
    perm = np.arange(m)
    for k, p in enumerate(piv):
        p -= 1
        if p != k:
            perm[k], perm[p] = perm[p], perm[k]
    return perm


def _factor(torch, handle, a, tournament):
    m, n = a.shape
    L = A.lib()
    d = A.to_device_cm(a)
    lw = ctypes.c_int(0)
    assert L.aclsolverSgetrf_bufferSize(handle.h, m, n, A.ptr(d), m, ctypes.byref(lw)) == A.SUCCESS
    work = torch.zeros(lw.value, device="cuda")
    ipiv = torch.zeros(min(m, n), dtype=torch.int32, device="cuda")
    info = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    fn = L.aclsolverSgetrfTournament if tournament else L.aclsolverSgetrf
    st = fn(handle.h, m, n, A.ptr(d), m, A.ptr(work), A.ptr(ipiv), A.ptr(info))
    torch.cuda.synchronize()
    assert st == A.SUCCESS
    return d, ipiv.cpu().numpy(), int(info.item())


def _residual(torch, a, d, piv):
    """||P A - L U||_1 / (n ||A||_1) in FP64 on the GPU; also max |L|."""
    m, n = a.shape
    k = min(m, n)
    a64 = torch.from_numpy(a).cuda().double()
    lu = d[:n, :m].T.double()
    low = torch.tril(lu[:, :k], -1)
    low[:k, :k] += torch.eye(k, dtype=torch.float64, device="cuda")
    up = torch.triu(lu[:k, :])
    perm = torch.from_numpy(_ipiv_to_perm(piv, m)).cuda()
    r = a64[perm] - low @ up
    res = (r.abs().sum(0).max() / (n * a64.abs().sum(0).max())).item()
    return res, low.abs().max().item()


@pytest.mark.parametrize("m,n,seed", [(1100, 1100, 1), (2100, 300, 2), (4100, 96, 3), (16384, 64, 4), (16384, 40, 5)])
def test_first_leaf_pivots_match_oracle_tournament(torch, handle, m, n, seed):
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((m, n)) / np.sqrt(m)).astype(np.float32)
    print(f"tslu first leaf m={m} n={n} seed={seed}", flush=True)
    d, piv, info = _factor(torch, handle, a, True)
    assert info == 0
    w = min(32, n)
    ref, rinfo = O.stslu_pivots(a, w)
    assert rinfo == 0
    np.testing.assert_array_equal(piv[:w], ref)
    k = min(m, n)
    assert ((piv >= np.arange(1, k + 1)) & (piv <= m)).all()
    res, lmax = _residual(torch, a, d, piv)
    print(f"  residual {res:.3e}, max|L| {lmax:.3f}", flush=True)
    assert res <= 10 * max(m, n) * EPS32
    assert lmax <= 8.0


@pytest.mark.parametrize("n,seed", [(2100, 11), (4100, 12)])
def test_square_residual_and_growth(torch, handle, n, seed):
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32)
    print(f"tslu square n={n} seed={seed}", flush=True)
    d, piv, info = _factor(torch, handle, a, True)
    assert info == 0
    res, lmax = _residual(torch, a, d, piv)
    print(f"  residual {res:.3e}, max|L| {lmax:.3f}", flush=True)
    assert res <= 10 * n * EPS32
    assert lmax <= 8.0
    # deterministic: a second factorization gives the same bits
    d2, piv2, _ = _factor(torch, handle, a, True)
    np.testing.assert_array_equal(piv2, piv)
    assert torch.equal(d2, d)


def test_short_panels_equal_partial_pivoting(torch, handle):
    """No panel taller than 1024 rows: no tournament leaf runs, the factors are sgetrf's."""
    rng = np.random.default_rng(21)
    a = (rng.standard_normal((1024, 700)) / 32).astype(np.float32)
    d1, p1, i1 = _factor(torch, handle, a, True)
    d2, p2, i2 = _factor(torch, handle, a, False)
    assert i1 == i2 == 0
    np.testing.assert_array_equal(p1, p2)
    assert torch.equal(d1, d2)


def test_zero_column_reports_info(torch, handle):
    rng = np.random.default_rng(31)
    m = 2100
    a = (rng.standard_normal((m, m)) / np.sqrt(m)).astype(np.float32)
    a[:, 40] = 0.0
    d, piv, info = _factor(torch, handle, a, True)
    assert info == 41
    _, _, pinfo = _factor(torch, handle, a, False)
    assert pinfo == 41
