"""GPU parity of the real LU path through the C ABI against the CPU oracle.

Bars (written per test): ipiv identical to the oracle on tie-free inputs (the oracle
reports each step's tie margin; steps at or beyond the first margin < 1e-4 are excluded);
backward error ||PA - LU||_1 / (n ||A||_1) <= 10 n eps32; solutions against the FP64
oracle; info and status codes exactly as the reference contract; bitwise determinism
over repeats (test/sgetrf/real_lu_test_impl.h:153-172).
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


def _sgetrf(torch, handle, a, lda=None, pivot=True):
    m, n = a.shape
    lda = lda or max(m, 1)
    d = A.to_device_cm(a, ld=lda)
    lw = ctypes.c_int(0)
    assert A.lib().aclsolverSgetrf_bufferSize(handle.h, m, n, A.ptr(d), lda, ctypes.byref(lw)) == A.SUCCESS
    work = torch.zeros(max(lw.value, 1), dtype=torch.float32, device="cuda")
    ipiv = torch.zeros(max(min(m, n), 1), dtype=torch.int32, device="cuda") if pivot else None
    info = torch.full((1,), -99, dtype=torch.int32, device="cuda")
    st = A.lib().aclsolverSgetrf(handle.h, m, n, A.ptr(d), lda, A.ptr(work), A.ptr(ipiv), A.ptr(info))
    torch.cuda.synchronize()
    lu = A.from_device_cm(d, m, n)
    piv = ipiv[: min(m, n)].cpu().numpy() if pivot else None
    return st, lu, piv, int(info.item()), d


def _check_pivots(piv, ref_piv, margin):
    bad = np.nonzero(margin < 1e-4)[0]
    upto = int(bad[0]) if len(bad) else len(piv)
    np.testing.assert_array_equal(piv[:upto], ref_piv[:upto])
    return upto


@pytest.mark.parametrize("m,n,seed", [(512, 512, 1), (64, 64, 2), (300, 200, 3), (200, 300, 4), (1, 1, 5),
                                      (5, 7, 6), (1000, 1000, 7), (2100, 2100, 8), (33, 1, 9), (4100, 256, 10),
                                      (8200, 300, 11), (16384, 72, 12),
                                      # thin: m <= 32 < n, n > 128 (the reference's thin kernel shape)
                                      (16, 300, 13), (32, 1000, 14), (7, 129, 15), (1, 200, 16)])
def test_sgetrf_parity(torch, handle, m, n, seed):
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((m, n)) / np.sqrt(n)).astype(np.float32)
    st, lu, piv, info, _ = _sgetrf(torch, handle, a, lda=m + 3)
    assert st == A.SUCCESS and info == 0
    ref_lu, ref_piv, ref_info, margin = O.sgetrf(a, margins=True)
    upto = _check_pivots(piv, ref_piv, margin)
    if upto == len(piv):
        # same pivots -> the factors agree to within accumulated rounding
        scale = max(1.0, float(np.abs(ref_lu).max()))
        np.testing.assert_allclose(lu, ref_lu, rtol=0, atol=1e-4 * scale * np.sqrt(max(m, n) / 64))
    # backward error bound 10 * n * eps32
    assert O.lu_backward_error(a, lu, piv) <= 10 * max(m, n) * EPS32


def test_sgetrf_config1_solve_vs_fp64(torch, handle):
    """Config: n=512, A = N(0,1)/sqrt(n), one RHS b ~ N(0,1), partial pivoting. Seed 513 is
    tie-free (every step's top-2 pivot candidates differ by > 1e-4 relative), so ipiv must
    match the oracle over all 512 steps."""
    n = 512
    rng = np.random.default_rng(513)
    a = (rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32)
    b = rng.standard_normal((n, 1)).astype(np.float32)
    st, lu, piv, info, d = _sgetrf(torch, handle, a)
    assert st == A.SUCCESS and info == 0
    ref_lu, ref_piv, _, margin = O.sgetrf(a, margins=True)
    assert _check_pivots(piv, ref_piv, margin) == n
    db = A.to_device_cm(b)
    dpiv = torch.from_numpy(piv).cuda()
    dinfo = torch.zeros(1, dtype=torch.int32, device="cuda")
    st = A.lib().aclsolverSgetrsDevice(handle.h, 0, n, 1, A.ptr(d), n, A.ptr(dpiv), A.ptr(db), n, A.ptr(dinfo))
    torch.cuda.synchronize()
    assert st == A.SUCCESS and dinfo.item() == 0
    x = A.from_device_cm(db, n, 1)
    x64 = np.linalg.solve(a.astype(np.float64), b.astype(np.float64))
    assert O.solve_backward_error(a, x, b) <= 10 * n * EPS32
    cond = np.linalg.cond(a.astype(np.float64))
    assert np.abs(x - x64).max() / np.abs(x64).max() <= 10 * n * EPS32 * cond


def test_sgetrf_no_pivoting(torch, handle):
    n = 96
    rng = np.random.default_rng(11)
    a = (rng.standard_normal((n, n)) / np.sqrt(n) + 4 * np.eye(n)).astype(np.float32)
    st, lu, piv, info, _ = _sgetrf(torch, handle, a, pivot=False)
    assert st == A.SUCCESS and info == 0
    ref_lu, _, _ = O.sgetrf(a, pivot=False)
    np.testing.assert_allclose(lu, ref_lu, rtol=0, atol=1e-5)


def _fill_input(m, n, b, singular):
    out = np.zeros((m, n), np.float32)
    bad = singular and b == 0
    for r in range(m):
        for c in range(n):
            if r == c:
                out[r, c] = 0.0 if (bad and r == 1) else float(n + r + b + 2)
            elif bad:
                out[r, c] = 0.0
            else:
                out[r, c] = 0.25
    return out


@pytest.mark.parametrize("m,n,pivot,singular", [(7, 5, True, False), (3, 3, False, False), (2, 2, True, False),
                                                 (3, 3, True, True)])
def test_reference_basic_cases(torch, handle, m, n, pivot, singular):
    # test/sgetrf/real_lu_test_impl.h:306-323 with lda = m + 8 (:39).
    a = _fill_input(m, n, 0, singular)
    if pivot and m > 1 and not singular:
        a[[0, 1]] = a[[1, 0]]
    st, lu, piv, info, _ = _sgetrf(torch, handle, a, lda=m + 8, pivot=pivot)
    assert st == A.SUCCESS
    if singular:
        assert info == 2
        return
    assert info == 0
    low, up = O.unpack_lu(lu.astype(np.float64))
    pa = O.apply_ipiv_rows(a.astype(np.float64), piv) if pivot else a.astype(np.float64)
    assert np.all(np.abs(low @ up - pa) <= 1e-4 * np.maximum(1, np.abs(pa)))
    ref_lu, ref_piv, _ = O.sgetrf(a, pivot=pivot)
    if pivot:
        np.testing.assert_array_equal(piv, ref_piv)


def test_subnormal_pivot_kats(torch, handle):
    # test/sgetrf/sgetrf_subnormal_pivot_test.cpp:157-169, relative residual <= 1e-5.
    for m, n, vals in [(2, 2, [2e-39, 1e-39, 1e-39, 2e-39]), (2, 2, [4e-38, 2e-38, 2e-38, 4e-38]),
                       (513, 1, [1e-39] * 513)]:
        a = np.array(vals, np.float32).reshape(m, n)
        st, lu, piv, info, _ = _sgetrf(torch, handle, a)
        assert st == A.SUCCESS and info == 0
        low, up = O.unpack_lu(lu.astype(np.float64))
        pa = O.apply_ipiv_rows(a.astype(np.float64), piv)
        assert np.abs(low @ up - pa).max() / np.abs(pa).max() <= 1e-5


def test_singular_and_zero_columns(torch, handle):
    n = 70
    rng = np.random.default_rng(3)
    a = rng.standard_normal((n, n)).astype(np.float32)
    a[:, 40] = 0.0
    st, lu, piv, info, _ = _sgetrf(torch, handle, a)
    _, _, ref_info = O.sgetrf(a)
    assert st == A.SUCCESS and info == ref_info == 41


def test_tall_zero_column_and_ties(torch, handle):
    """Tall panels (leaves spread over several workgroups): an exactly zero column gives
    info at that column and the elimination skips it; equal-magnitude candidates in
    different row blocks resolve to the lowest row, as LAPACK's first-maximum rule."""
    m, n = 9000, 40
    rng = np.random.default_rng(21)
    a = rng.standard_normal((m, n)).astype(np.float32)
    a[:, 17] = 0.0
    # column 0: the same largest magnitude at rows 7000, 600 and 8999 -> row 600 wins
    a[:, 0] = np.clip(a[:, 0], -3.0, 3.0)
    a[7000, 0], a[600, 0], a[8999, 0] = -8.0, 8.0, -8.0
    st, lu, piv, info, _ = _sgetrf(torch, handle, a)
    ref_lu, ref_piv, ref_info, margin = O.sgetrf(a, margins=True)
    assert st == A.SUCCESS and info == ref_info == 18
    assert piv[0] == ref_piv[0] == 601
    _check_pivots(piv[:17], ref_piv[:17], np.minimum(margin[:17], 1.0) + (np.arange(17) == 0))


def test_determinism(torch, handle):
    n = 700
    rng = np.random.default_rng(99)
    a = rng.standard_normal((n, n)).astype(np.float32)
    outs = [_sgetrf(torch, handle, a)[1:4] for _ in range(3)]
    for lu, piv, info in outs[1:]:
        assert np.array_equal(lu.view(np.uint32), outs[0][0].view(np.uint32))
        assert np.array_equal(piv, outs[0][1]) and info == outs[0][2]
    # a tall panel, whose leaves exchange pivots between workgroups
    a = rng.standard_normal((6000, 200)).astype(np.float32)
    outs = [_sgetrf(torch, handle, a)[1:4] for _ in range(3)]
    for lu, piv, info in outs[1:]:
        assert np.array_equal(lu.view(np.uint32), outs[0][0].view(np.uint32))
        assert np.array_equal(piv, outs[0][1]) and info == outs[0][2]


@pytest.mark.parametrize("m,n,seed", [(3000, 3000, 7), (4500, 1300, 8), (1300, 2000, 9)])
def test_deferred_interchanges_agree(torch, handle, m, n, seed, monkeypatch):
    """The interchanges of the columns left of each panel, applied once at the end through
    the composed row map (the default) or panel by panel (OPS_SOLVER_DEFER_SWAPS=0), give
    identical factors; unshifted N(0,1) entries pivot on nearly every column."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((m, n)).astype(np.float32)
    st, lu, piv, info, _ = _sgetrf(torch, handle, a)
    monkeypatch.setenv("OPS_SOLVER_DEFER_SWAPS", "0")
    st0, lu0, piv0, info0, _ = _sgetrf(torch, handle, a)
    monkeypatch.delenv("OPS_SOLVER_DEFER_SWAPS")
    assert st == st0 == A.SUCCESS and info == info0 == 0
    assert np.array_equal(piv, piv0)
    assert np.array_equal(lu.view(np.uint32), lu0.view(np.uint32))
    assert O.lu_backward_error(a, lu, piv) <= 10 * max(m, n) * EPS32


@pytest.mark.parametrize("m,n,seed", [(6000, 200, 5), (3000, 3000, 6)])
def test_tall_leaf_exchange_paths_agree(torch, handle, m, n, seed, monkeypatch):
    """The tall leaves' two exchange paths (spread over one XCD's L2, the default where the
    device allows it, and the device-coherent path, OPS_SOLVER_LEAF_XCD=0) give identical
    factors; both match the oracle's pivots and meet the backward-error bar."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((m, n)).astype(np.float32)
    st, lu, piv, info, _ = _sgetrf(torch, handle, a)
    monkeypatch.setenv("OPS_SOLVER_LEAF_XCD", "0")
    st0, lu0, piv0, info0, _ = _sgetrf(torch, handle, a)
    monkeypatch.delenv("OPS_SOLVER_LEAF_XCD")
    assert st == st0 == A.SUCCESS and info == info0 == 0
    assert np.array_equal(lu.view(np.uint32), lu0.view(np.uint32))
    assert np.array_equal(piv, piv0)
    # the first kp pivots depend only on the first kp columns
    kp = min(m, n, 256)
    _, ref_piv, _, margin = O.sgetrf(np.ascontiguousarray(a[:, :kp]), margins=True)
    _check_pivots(piv[:kp], ref_piv, margin)
    assert O.lu_backward_error(a, lu, piv) <= 10 * max(m, n) * EPS32


def test_invalid_arguments(torch, handle):
    L = A.lib()
    info = torch.full((1,), 5, dtype=torch.int32, device="cuda")
    # RunBasicCases CheckInvalid: m = -1 -> INVALID_VALUE, devInfo = -1
    assert L.aclsolverSgetrf(handle.h, -1, 2, None, 1, None, None, A.ptr(info)) == A.INVALID_VALUE
    torch.cuda.synchronize()
    assert info.item() == -1
    a = torch.zeros((4, 4), device="cuda")
    work = torch.zeros(4096, device="cuda")
    assert L.aclsolverSgetrf(handle.h, 4, 4, A.ptr(a), 3, A.ptr(work), None, A.ptr(info)) == A.INVALID_VALUE
    torch.cuda.synchronize()
    assert info.item() == -4
    # non-finite input: devInfo = -3 (src/sgetrf/sgetrf_host.cpp:118)
    a[1, 2] = float("nan")
    assert L.aclsolverSgetrf(handle.h, 4, 4, A.ptr(a), 4, A.ptr(work), None, A.ptr(info)) == A.INVALID_VALUE
    torch.cuda.synchronize()
    assert info.item() == -3
    # empty problem writes info = 0
    assert L.aclsolverSgetrf(handle.h, 0, 0, None, 1, None, None, A.ptr(info)) == A.SUCCESS
    torch.cuda.synchronize()
    assert info.item() == 0
    # strict reference limits: m > 8192 rejected with devInfo = -1
    L.aclsolverSetStrictLimits(handle.h, 1)
    assert L.aclsolverSgetrf(handle.h, 8193, 4, None, 8193, None, None, A.ptr(info)) == A.INVALID_VALUE
    L.aclsolverSetStrictLimits(handle.h, 0)
    torch.cuda.synchronize()
    assert info.item() == -1


@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("n,nrhs", [(5, 3), (300, 7), (1100, 2), (512, 1), (777, 1), (2049, 1)])
def test_getrs_device_and_host(torch, handle, trans, n, nrhs):
    rng = np.random.default_rng(n + trans)
    a = (rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32)
    b = rng.standard_normal((n, nrhs)).astype(np.float32)
    lu, piv, _ = O.sgetrf(a)
    ref = O.dgetrs(lu.astype(np.float64), piv, b.astype(np.float64), trans=trans)
    # Device column-major
    dA = A.to_device_cm(lu)
    dB = A.to_device_cm(b, ld=n + 2)
    dpiv = torch.from_numpy(piv).cuda()
    dinfo = torch.full((1,), 9, dtype=torch.int32, device="cuda")
    st = A.lib().aclsolverSgetrsDevice(handle.h, trans, n, nrhs, A.ptr(dA), n, A.ptr(dpiv), A.ptr(dB), n + 2,
                                       A.ptr(dinfo))
    torch.cuda.synchronize()
    assert st == A.SUCCESS and dinfo.item() == 0
    x = A.from_device_cm(dB, n, nrhs)
    tol = 50 * n * EPS32 * np.abs(ref).max() * max(1.0, np.linalg.cond(lu.astype(np.float64)) ** 0.5)
    np.testing.assert_allclose(x, ref, rtol=0, atol=tol)
    # Host row-major (reference contract): A[r*lda + c], B[r*ldb + c]
    lda, ldb = n + 1, nrhs + 2
    ha = np.zeros((n, lda), np.float32)
    ha[:, :n] = lu
    hb = np.zeros((n, ldb), np.float32)
    hb[:, :nrhs] = b
    hp = piv.astype(np.int32).copy()
    info = np.array([-99], np.int32)
    r = A.lib().aclsolverSgetrs(handle.h, trans, n, nrhs, ha.ctypes.data, lda, hp.ctypes.data, hb.ctypes.data, ldb,
                                info.ctypes.data)
    assert r == 0 and info[0] == 0
    np.testing.assert_allclose(hb[:, :nrhs], ref, rtol=0, atol=tol)


def _ptr_array(torch, tensors):
    return torch.tensor([t.data_ptr() for t in tensors], dtype=torch.int64, device="cuda")


@pytest.mark.parametrize("n,batch", [(32, 256), (1, 5), (2, 8), (7, 33), (16, 100), (64, 20), (3, 8)])
def test_smatinv_batched(torch, handle, n, batch):
    rng = np.random.default_rng(n * 1000 + batch)
    mats = [(rng.standard_normal((n, n)) / np.sqrt(n) + 2 * np.eye(n)).astype(np.float32) for _ in range(batch)]
    dA = [A.to_device_cm(m) for m in mats]
    dC = [torch.zeros((n, n), device="cuda") for _ in range(batch)]
    info = torch.full((batch,), -7, dtype=torch.int32, device="cuda")
    # keep the pointer arrays alive: a freed temporary would be recycled by the allocator
    pa, pc = _ptr_array(torch, dA), _ptr_array(torch, dC)
    st = A.lib().aclsolverSmatinvBatched(handle.h, n, A.ptr(pa), n, A.ptr(pc), n, A.ptr(info), batch)
    torch.cuda.synchronize()
    assert st == A.SUCCESS
    assert (info.cpu().numpy() == 0).all()
    for m, c in zip(mats, dC):
        inv = A.from_device_cm(c, n, n)
        ref, _, _ = O.smatinv(m)
        # A * inv(A) ~ I with rtol = atol = 5e-3 (test/utils/verify_utils.py:26-90) and the
        # residual bound ||I - C A||_1 / (n ||A|| ||C|| eps) <= 5 (complex_solver_test.cpp:86-103)
        np.testing.assert_allclose(m.astype(np.float64) @ inv, np.eye(n), rtol=5e-3, atol=5e-3)
        assert O.inverse_residual(m, inv) <= 5
        np.testing.assert_allclose(inv, ref, rtol=0, atol=1e-4 * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("n,bad,offset", [(32, float("nan"), 0), (32, float("inf"), 0), (32, float("-inf"), 1),
                                          (5, float("nan"), 0), (7, float("inf"), 3), (64, float("nan"), 2)])
def test_smatinv_batched_rejects_nonfinite(torch, handle, n, bad, offset):
    """The drop-in batched inverse scans its input first: one NaN or Inf anywhere in the
    batch gives INVALID_VALUE with info = -2 for every matrix and leaves the outputs
    untouched. offset (floats) shifts every matrix off 16-byte alignment, so both the
    vector and the scalar form of the scan run; orders 5 and 7 are not multiples of 4."""
    batch = 41
    rng = np.random.default_rng(n * 10 + offset)
    per = n * n + offset
    pool = torch.zeros(batch * per + 4, device="cuda")
    mats = []
    for b in range(batch):
        m = (rng.standard_normal((n, n)) / np.sqrt(n) + 2 * np.eye(n)).astype(np.float32)
        mats.append(m)
        pool[b * per + offset: b * per + offset + n * n] = torch.from_numpy(np.ascontiguousarray(m.T).reshape(-1))
    pool[23 * per + offset + (n * n) // 2 + 1] = bad
    dC = [torch.full((n, n), 7.0, device="cuda") for _ in range(batch)]
    pa = torch.tensor([pool.data_ptr() + 4 * (b * per + offset) for b in range(batch)], dtype=torch.int64,
                      device="cuda")
    pc = _ptr_array(torch, dC)
    info = torch.full((batch,), 9, dtype=torch.int32, device="cuda")
    st = A.lib().aclsolverSmatinvBatched(handle.h, n, A.ptr(pa), n, A.ptr(pc), n, A.ptr(info), batch)
    torch.cuda.synchronize()
    assert st == A.INVALID_VALUE
    assert (info.cpu().numpy() == -2).all()
    assert all(bool((c == 7.0).all().item()) for c in dC)
    # the same batch with the bad value removed goes through
    pool[23 * per + offset + (n * n) // 2 + 1] = 0.5
    st = A.lib().aclsolverSmatinvBatched(handle.h, n, A.ptr(pa), n, A.ptr(pc), n, A.ptr(info), batch)
    torch.cuda.synchronize()
    assert st == A.SUCCESS and (info.cpu().numpy() == 0).all()


def test_smatinv_strided_ipiv_and_singular(torch, handle):
    n, batch = 32, 1000
    rng = np.random.default_rng(5)
    mats = (rng.standard_normal((batch, n, n)) / np.sqrt(n)).astype(np.float32)
    mats[7, :, 5] = 0  # singular: zero column -> info = first zero pivot
    d = torch.from_numpy(np.ascontiguousarray(mats.transpose(0, 2, 1))).cuda()  # column-major per matrix
    c = torch.zeros_like(d)
    ipiv = torch.zeros((batch, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    st = A.lib().aclsolverSmatinvStridedBatched(handle.h, n, A.ptr(d), n, n * n, A.ptr(c), n, n * n, A.ptr(ipiv),
                                                A.ptr(info), batch)
    torch.cuda.synchronize()
    assert st == A.SUCCESS
    hinfo, hpiv, hc = info.cpu().numpy(), ipiv.cpu().numpy(), c.cpu().numpy()
    for b in range(batch):
        lu, ref_piv, ref_info, margin = O.sgetrf(mats[b], margins=True)
        assert hinfo[b] == ref_info
        if ref_info == 0:
            _check_pivots(hpiv[b], ref_piv, margin)
            inv = hc[b].T
            assert O.inverse_residual(mats[b], inv) <= 5


@pytest.mark.parametrize("n,batch", [(32, 64), (5, 9), (64, 8), (100, 2), (65, 33), (128, 20), (200, 7), (512, 3),
                                     (600, 2)])
def test_sgetri_batched(torch, handle, n, batch):
    rng = np.random.default_rng(n + 77)
    mats = [(rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32) for _ in range(batch)]
    facts = [O.sgetrf(m) for m in mats]
    dA = [A.to_device_cm(f[0]) for f in facts]
    piv = torch.from_numpy(np.concatenate([f[1] for f in facts]).astype(np.int32)).cuda()
    dC = [torch.zeros((n, n + 1), device="cuda") for _ in range(batch)]
    info = torch.full((batch,), -7, dtype=torch.int32, device="cuda")
    pa, pc = _ptr_array(torch, dA), _ptr_array(torch, dC)
    st = A.lib().aclsolverSgetriBatched(handle.h, n, A.ptr(pa), n, A.ptr(piv), A.ptr(pc), n + 1, A.ptr(info), batch)
    torch.cuda.synchronize()
    assert st == A.SUCCESS and (info.cpu().numpy() == 0).all()
    for m, c in zip(mats, dC):
        inv = A.from_device_cm(c, n, n)
        assert O.inverse_residual(m, inv) <= 5


def test_sgetri_batched_blocked_singular_and_unpivoted(torch, handle):
    """Orders 65..512 (blocked batched inverse): a singular factor reports info = first zero
    U(k, k) + 1 while the others are inverted; factors without pivots (PivotArray null)."""
    n, batch = 130, 5
    rng = np.random.default_rng(1300)
    mats = [(rng.standard_normal((n, n)) / np.sqrt(n) + 3 * np.eye(n)).astype(np.float32) for _ in range(batch)]
    facts = [O.sgetrf(m, pivot=False) for m in mats]
    lus = [f[0].copy() for f in facts]
    lus[2][97, 97] = 0.0  # matrix 2: U(97, 97) = 0
    dA = [A.to_device_cm(lu) for lu in lus]
    dC = [torch.zeros((n, n), device="cuda") for _ in range(batch)]
    info = torch.full((batch,), -7, dtype=torch.int32, device="cuda")
    pa, pc = _ptr_array(torch, dA), _ptr_array(torch, dC)
    st = A.lib().aclsolverSgetriBatched(handle.h, n, A.ptr(pa), n, None, A.ptr(pc), n, A.ptr(info), batch)
    torch.cuda.synchronize()
    assert st == A.SUCCESS
    got = info.cpu().numpy()
    assert got[2] == 98 and (np.delete(got, 2) == 0).all()
    for b in (0, 1, 3, 4):
        inv = A.from_device_cm(dC[b], n, n)
        assert O.inverse_residual(mats[b], inv) <= 5


def test_sgetri_single(torch, handle):
    n = 150
    rng = np.random.default_rng(150)
    a = (rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32)
    lu, piv, _ = O.sgetrf(a)
    d = A.to_device_cm(lu)
    dpiv = torch.from_numpy(piv).cuda()
    lw = ctypes.c_int(0)
    assert A.lib().aclsolverSgetri_bufferSize(handle.h, n, A.ptr(d), n, ctypes.byref(lw)) == A.SUCCESS
    work = torch.zeros(lw.value, device="cuda")
    info = torch.full((1,), -9, dtype=torch.int32, device="cuda")
    st = A.lib().aclsolverSgetri(handle.h, n, A.ptr(d), n, A.ptr(dpiv), A.ptr(work), lw.value, A.ptr(info))
    torch.cuda.synchronize()
    assert st == A.SUCCESS and info.item() == 0
    inv = A.from_device_cm(d, n, n)
    assert O.inverse_residual(a, inv) <= 5
