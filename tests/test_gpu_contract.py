"""GPU checks of the boundary contract: size limits, the architecture guard, the
bounded-wait timeout report, the reference's NaN pivot rule, pointer-array validation of
the batched entry points and the thin (m <= 32 < n) factorization shape.

Bars (per test): status codes and devInfo exactly as the contract states
(include/cann_ops_solver.h; reference src/sgetrf/sgetrf_host.cpp:46-73,
src/utils/real_lu_common.h:383-410, src/utils/complex_solver_host.h:50-53); ipiv identical
to the oracle (which follows src/utils/kernel/real_lu.hpp:59-77, :100-109 for NaN) on every
step that is not a finite pivot tie (margin < 1e-4); backward error <= 10 n eps32.
"""
import ctypes
import os

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


def _info(torch, v=7):
    return torch.full((1,), v, dtype=torch.int32, device="cuda")


# ---------------------------------------------------------------------------------------
# limits
# ---------------------------------------------------------------------------------------
def test_engine_limits_give_invalid_value(torch, handle):
    """m or n beyond the engine's 16384 gives INVALID_VALUE with devInfo = -(argument) --
    never EXECUTION_FAILED -- also with the reference's strict limits off."""
    L = A.lib()
    L.aclsolverSetStrictLimits(handle.h, 0)
    info = _info(torch)
    a = torch.zeros(16, device="cuda")
    work = torch.zeros(16, device="cuda")
    assert L.aclsolverSgetrf(handle.h, 16385, 4, A.ptr(a), 16385, A.ptr(work), None, A.ptr(info)) == A.INVALID_VALUE
    torch.cuda.synchronize()
    assert info.item() == -1
    lw = ctypes.c_int(0)
    assert L.aclsolverSgetrf_bufferSize(handle.h, 16385, 4, None, 16385, ctypes.byref(lw)) == A.INVALID_VALUE
    # device getrs: n = 16385 -> devInfo = -2
    info.fill_(7)
    assert L.aclsolverSgetrsDevice(handle.h, 0, 16385, 1, A.ptr(a), 16385, None, A.ptr(a), 16385,
                                   A.ptr(info)) == A.INVALID_VALUE
    torch.cuda.synchronize()
    assert info.item() == -2
    # getri: n = 16385 -> devInfo = -1
    info.fill_(7)
    assert L.aclsolverSgetri(handle.h, 16385, A.ptr(a), 16385, None, A.ptr(work), 16, A.ptr(info)) == A.INVALID_VALUE
    torch.cuda.synchronize()
    assert info.item() == -1
    # complex: m = 16385 -> devInfo = -1
    info.fill_(7)
    assert L.aclsolverCgetrf(handle.h, 16385, 4, A.ptr(a), 16385, A.ptr(work), None, A.ptr(info)) == A.INVALID_VALUE
    torch.cuda.synchronize()
    assert info.item() == -1
    # host getrs: n = 16385 rejected before any device work
    ha = np.zeros(4, np.float32)
    hp = np.ones(4, np.int32)
    hinfo = np.zeros(1, np.int32)
    assert L.aclsolverSgetrs(handle.h, 0, 16385, 1, ha.ctypes.data, 16385, hp.ctypes.data, ha.ctypes.data, 1,
                             hinfo.ctypes.data) == A.ACL_ERROR_INVALID_PARAM
    # 16384 itself passes the argument checks (m = 16384, n = 1)
    m = 16384
    col = torch.randn(m, device="cuda")
    lw = ctypes.c_int(0)
    assert L.aclsolverSgetrf_bufferSize(handle.h, m, 1, A.ptr(col), m, ctypes.byref(lw)) == A.SUCCESS
    work = torch.zeros(lw.value, device="cuda")
    ipiv = torch.zeros(1, dtype=torch.int32, device="cuda")
    info.fill_(7)
    assert L.aclsolverSgetrf(handle.h, m, 1, A.ptr(col), m, A.ptr(work), A.ptr(ipiv), A.ptr(info)) == A.SUCCESS
    torch.cuda.synchronize()
    assert info.item() == 0


def test_arch_mismatch_guard(torch):
    """A handle created on a device that does not report gfx950 answers ARCH_MISMATCH on
    every compute entry point (the override variable stands in for another device)."""
    os.environ["OPS_SOLVER_ARCH_OVERRIDE"] = "gfx942:sramecc+:xnack-"
    try:
        h = A.Handle()
    finally:
        del os.environ["OPS_SOLVER_ARCH_OVERRIDE"]
    L = A.lib()
    info = _info(torch)
    a = torch.eye(4, device="cuda")
    work = torch.zeros(4096, device="cuda")
    lw = ctypes.c_int(0)
    try:
        assert L.aclsolverSgetrf_bufferSize(h.h, 4, 4, A.ptr(a), 4, ctypes.byref(lw)) == A.ARCH_MISMATCH
        assert L.aclsolverSgetrf(h.h, 4, 4, A.ptr(a), 4, A.ptr(work), None, A.ptr(info)) == A.ARCH_MISMATCH
        assert L.aclsolverCgetrf(h.h, 2, 2, A.ptr(a), 2, A.ptr(work), None, A.ptr(info)) == A.ARCH_MISMATCH
        assert L.aclsolverSgetrsDevice(h.h, 0, 4, 1, A.ptr(a), 4, None, A.ptr(work), 4,
                                       A.ptr(info)) == A.ARCH_MISMATCH
        sz = ctypes.c_size_t(0)
        assert L.aclsolverSBgesv_bufferSize(h.h, 4, 1, ctypes.byref(sz)) == A.ARCH_MISMATCH
        torch.cuda.synchronize()
        assert info.item() == 7  # nothing was launched
    finally:
        h.close()
    # a handle made without the override works on this gfx950 device
    with A.Handle() as h2:
        assert L.aclsolverSgetrf_bufferSize(h2.h, 4, 4, A.ptr(a), 4, ctypes.byref(lw)) == A.SUCCESS


# ---------------------------------------------------------------------------------------
# bounded waits: forced from the host with a zero poll budget, every kernel still exits
# ---------------------------------------------------------------------------------------
def test_leaf_wait_timeout_is_reported(torch, handle):
    L = A.lib()
    m = n = 2304  # panels taller than 2048 rows run the multi-workgroup leaves
    rng = np.random.default_rng(44)
    a = (rng.standard_normal((m, n)) / np.sqrt(n)).astype(np.float32)
    ca = (a + 1j * rng.standard_normal((m, n)).astype(np.float32) / np.sqrt(n)).astype(np.complex64)
    d = A.to_device_cm(a)
    cd = A.to_device_cm(ca)
    lw = ctypes.c_int(0)
    assert L.aclsolverSgetrf_bufferSize(handle.h, m, n, A.ptr(d), m, ctypes.byref(lw)) == A.SUCCESS
    work = torch.zeros(lw.value, device="cuda")
    assert L.aclsolverCgetrf_bufferSize(handle.h, m, n, A.ptr(cd), m, ctypes.byref(lw)) == A.SUCCESS
    cwork = torch.zeros(max(lw.value, 1), dtype=torch.complex64, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = _info(torch)
    try:
        assert L.aclsolverDebugSetSpinLimit(0) == A.SUCCESS
        st = L.aclsolverSgetrf(handle.h, m, n, A.ptr(d), m, A.ptr(work), A.ptr(ipiv), A.ptr(info))
        torch.cuda.synchronize()
        assert st == A.SUCCESS and info.item() == A.INFO_WAIT_TIMEOUT
        info.fill_(7)
        st = L.aclsolverCgetrf(handle.h, m, n, A.ptr(cd), m, A.ptr(cwork), A.ptr(ipiv), A.ptr(info))
        torch.cuda.synchronize()
        assert st == A.SUCCESS and info.item() == A.INFO_WAIT_TIMEOUT
        # the refined gesv family turns it into a failed call, the plain one into dinfo
        sz = ctypes.c_size_t(0)
        assert L.aclsolverSBgesv_bufferSize(handle.h, n, 1, ctypes.byref(sz)) == A.SUCCESS
        gw = torch.empty(sz.value, dtype=torch.uint8, device="cuda")
        at = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
        b = torch.randn(n, device="cuda")
        x = torch.empty(n, device="cuda")
        niter = ctypes.c_int(0)
        info.fill_(7)
        st = L.aclsolverSBgesv(handle.h, n, 1, A.ptr(at), n, A.ptr(ipiv), A.ptr(b), n, A.ptr(x), n, A.ptr(gw),
                               sz.value, ctypes.byref(niter), A.ptr(info))
        torch.cuda.synchronize()
        assert st == A.EXECUTION_FAILED
        info.fill_(7)
        st = L.aclsolverSSgesv(handle.h, n, 1, A.ptr(at), n, A.ptr(ipiv), A.ptr(b), n, A.ptr(x), n, A.ptr(gw),
                               sz.value, ctypes.byref(niter), A.ptr(info))
        torch.cuda.synchronize()
        assert st == A.SUCCESS and info.item() == A.INFO_WAIT_TIMEOUT
    finally:
        L.aclsolverDebugSetSpinLimit(A.DEFAULT_SPIN_LIMIT)
    # with the default budget the same factorization succeeds again (the flag is per call)
    d = A.to_device_cm(a)
    info.fill_(7)
    st = L.aclsolverSgetrf(handle.h, m, n, A.ptr(d), m, A.ptr(work), A.ptr(ipiv), A.ptr(info))
    torch.cuda.synchronize()
    assert st == A.SUCCESS and info.item() == 0
    lu = A.from_device_cm(d, m, n)
    assert O.lu_backward_error(a, lu, ipiv.cpu().numpy()) <= 10 * n * EPS32


# ---------------------------------------------------------------------------------------
# NaN pivot rule (reference: a NaN on the diagonal position is the pivot, a NaN below it
# never wins) on the paths without a finite scan
# ---------------------------------------------------------------------------------------
def _compare_pivots_with_nan(piv, ref_piv, margin, ref_lu):
    """Compare every step before the first finite pivot tie (margin < 1e-4 with a finite
    chosen pivot); NaN steps are exact by the rule and always compared."""
    n = len(ref_piv)
    upto = n
    for k in range(n):
        if margin[k] < 1e-4 and np.isfinite(ref_lu[k, k]):
            upto = k
            break
    np.testing.assert_array_equal(piv[:upto], ref_piv[:upto])
    return upto


@pytest.mark.parametrize("case", ["diag0", "below0", "diag_later", "column"])
def test_nan_pivot_rule_uneven_batch(torch, case):
    rng = np.random.default_rng({"diag0": 1, "below0": 2, "diag_later": 3, "column": 4}[case])
    sizes = [8, 40, 64, 100, 150, 230]
    nrhs = 2
    mats, rhs = [], []
    for n in sizes:
        a = (rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32)
        if case == "diag0":
            a[0, 0] = np.nan
        elif case == "below0":
            a[n // 2, 0] = np.nan
        elif case == "diag_later":
            a[n // 3, n // 3] = np.nan
        else:
            a[:, 1] = np.nan
        mats.append(a)
        rhs.append(rng.standard_normal((n, nrhs)).astype(np.float32))
    dA = [A.to_device_cm(m) for m in mats]
    dB = [A.to_device_cm(b) for b in rhs]
    pa = torch.tensor([t.data_ptr() for t in dA], dtype=torch.int64, device="cuda")
    pb = torch.tensor([t.data_ptr() for t in dB], dtype=torch.int64, device="cuda")
    ipiv = torch.zeros(sum(sizes), dtype=torch.int32, device="cuda")
    info = torch.full((len(sizes),), -9, dtype=torch.int32, device="cuda")
    ns = np.array(sizes, np.int32)
    with A.Handle() as h:
        st = A.lib().aclsolverSgesvVbatched(h.h, len(sizes), ns.ctypes.data, nrhs, A.ptr(pa), A.ptr(pb),
                                            A.ptr(ipiv), A.ptr(info))
        torch.cuda.synchronize()
    assert st == A.SUCCESS
    hpiv = ipiv.cpu().numpy()
    off = 0
    for i, n in enumerate(sizes):
        ref_lu, ref_piv, ref_info, margin = O.sgetrf(mats[i], margins=True)
        piv = hpiv[off:off + n]
        off += n
        upto = _compare_pivots_with_nan(piv, ref_piv, margin, ref_lu)
        if case == "diag0":
            assert piv[0] == 1 and upto == n
        if case == "below0":
            assert piv[0] != n // 2 + 1


@pytest.mark.parametrize("n", [512, 3000])
def test_nan_pivot_rule_dense_gesv(torch, handle, n):
    """A NaN at (0, 0) is the first pivot (then every later column is NaN, so every pivot
    is its own diagonal); a NaN below the diagonal of column 0 is never chosen. Panels of
    up to 2048 rows run the one-workgroup leaf, taller ones the multi-workgroup leaf. The
    refinement reports a NaN solution as not converged (negative niter)."""
    L = A.lib()
    rng = np.random.default_rng(n)
    a = (rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32)
    sz = ctypes.c_size_t(0)
    assert L.aclsolverSBgesv_bufferSize(handle.h, n, 1, ctypes.byref(sz)) == A.SUCCESS
    work = torch.empty(sz.value, dtype=torch.uint8, device="cuda")
    x = torch.empty(n, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = _info(torch)
    b = torch.randn(n, device="cuda")

    def run(name, mat):
        at = torch.from_numpy(np.ascontiguousarray(mat.T)).cuda()
        niter = ctypes.c_int(0)
        info.fill_(7)
        st = getattr(L, f"aclsolver{name}gesv")(handle.h, n, 1, A.ptr(at), n, A.ptr(ipiv), A.ptr(b), n, A.ptr(x), n,
                                                A.ptr(work), sz.value, ctypes.byref(niter), A.ptr(info))
        torch.cuda.synchronize()
        return st, niter.value, int(info.item()), ipiv.cpu().numpy()

    m1 = a.copy()
    m1[0, 0] = np.nan
    for name in ("SB", "SS"):
        st, niter, inf, piv = run(name, m1)
        assert st == A.SUCCESS and inf == 0
        np.testing.assert_array_equal(piv, np.arange(1, n + 1))
        if name == "SB":
            assert niter < 0
    m2 = a.copy()
    col0 = np.abs(m2[:, 0])
    m2[int(np.argmax(col0)), 0] = np.nan  # the would-be pivot becomes NaN
    st, niter, inf, piv = run("SS", m2)
    assert st == A.SUCCESS
    col0[int(np.argmax(col0))] = -1
    assert piv[0] == int(np.argmax(col0)) + 1
    # NaN in b: SUCCESS, info 0, refinement not converged
    bn = b.clone()
    b[5] = float("nan")
    st, niter, inf, piv = run("SB", a)
    b.copy_(bn)
    assert st == A.SUCCESS and inf == 0 and niter < 0


def test_butterfly_nonfinite_falls_back(torch, handle):
    """The unpivoted butterfly form on an input whose solve is non-finite: the refinement
    sees a NaN backward error, refactors with partial pivoting once, and then reports no
    convergence -- without a hang and with a valid pivot sequence."""
    L = A.lib()
    n = 1024
    rng = np.random.default_rng(12)
    a = (rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32)
    a[100, 200] = np.nan
    at = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    b = torch.randn(n, device="cuda")
    x = torch.empty(n, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = _info(torch)
    sz = ctypes.c_size_t(0)
    assert L.aclsolverSetButterfly(handle.h, 1) == A.SUCCESS
    try:
        assert L.aclsolverSBgesv_bufferSize(handle.h, n, 1, ctypes.byref(sz)) == A.SUCCESS
        work = torch.empty(sz.value, dtype=torch.uint8, device="cuda")
        niter = ctypes.c_int(0)
        st = L.aclsolverSBgesv(handle.h, n, 1, A.ptr(at), n, A.ptr(ipiv), A.ptr(b), n, A.ptr(x), n, A.ptr(work),
                               sz.value, ctypes.byref(niter), A.ptr(info))
        torch.cuda.synchronize()
    finally:
        L.aclsolverSetButterfly(handle.h, 0)
    assert st == A.SUCCESS and niter.value < 0
    piv = ipiv.cpu().numpy()
    assert ((piv >= np.arange(1, n + 1)) & (piv <= n)).all()


# ---------------------------------------------------------------------------------------
# batched pointer-array validation (src/utils/real_lu_common.h:196-428)
# ---------------------------------------------------------------------------------------
def _ptrs(torch, tensors):
    return torch.tensor([t.data_ptr() for t in tensors], dtype=torch.int64, device="cuda")


def test_batched_pointer_array_rejections(torch, handle):
    L = A.lib()
    n, batch = 8, 4
    mats = [torch.eye(n, device="cuda") * 2 for _ in range(batch)]
    outs = [torch.zeros((n, n), device="cuda") for _ in range(batch)]
    info = torch.full((batch,), 7, dtype=torch.int32, device="cuda")
    pa = _ptrs(torch, mats)

    def matinv(pc):
        info.fill_(7)
        st = L.aclsolverSmatinvBatched(handle.h, n, A.ptr(pa), n, A.ptr(pc), n, A.ptr(info), batch)
        torch.cuda.synchronize()
        return st, info.cpu().numpy()

    # a good call
    st, inf = matinv(_ptrs(torch, outs))
    assert st == A.SUCCESS and (inf == 0).all()
    # two outputs alias each other -> Ainv (argument 4)
    st, inf = matinv(_ptrs(torch, [outs[0], outs[1], outs[1], outs[3]]))
    assert st == A.INVALID_VALUE and (inf == -4).all()
    # an output overlaps an input -> argument 4
    st, inf = matinv(_ptrs(torch, [outs[0], mats[2], outs[2], outs[3]]))
    assert st == A.INVALID_VALUE and (inf == -4).all()
    # an output overlaps the info array -> argument 4
    big = torch.zeros(n * n + batch, device="cuda")
    pc = _ptrs(torch, [outs[0], outs[1], outs[2], big])
    st = L.aclsolverSmatinvBatched(handle.h, n, A.ptr(pa), n, A.ptr(pc), n, A.ptr(big) + 4 * (n * n - 2), batch)
    torch.cuda.synchronize()
    assert st == A.INVALID_VALUE
    # an input that is host memory -> A (argument 2)
    host = np.zeros(n * n, np.float32)
    pa_bad = torch.tensor([mats[0].data_ptr(), host.ctypes.data, mats[2].data_ptr(), mats[3].data_ptr()],
                          dtype=torch.int64, device="cuda")
    info.fill_(7)
    st = L.aclsolverSmatinvBatched(handle.h, n, A.ptr(pa_bad), n, A.ptr(_ptrs(torch, outs)), n, A.ptr(info), batch)
    torch.cuda.synchronize()
    assert st == A.INVALID_VALUE and (info.cpu().numpy() == -2).all()
    # the pointer array itself in host memory -> argument 2
    host_arr = np.array([t.data_ptr() for t in mats], np.int64)
    info.fill_(7)
    st = L.aclsolverSmatinvBatched(handle.h, n, host_arr.ctypes.data, n, A.ptr(_ptrs(torch, outs)), n, A.ptr(info),
                                   batch)
    torch.cuda.synchronize()
    assert st == A.INVALID_VALUE and (info.cpu().numpy() == -2).all()
    # getri_batched: a pivot outside [k+1, n] -> PivotArray (argument 4); aliasing C -> 5
    piv = torch.arange(1, n + 1, dtype=torch.int32, device="cuda").repeat(batch)
    info.fill_(7)
    st = L.aclsolverSgetriBatched(handle.h, n, A.ptr(pa), n, A.ptr(piv), A.ptr(_ptrs(torch, outs)), n, A.ptr(info),
                                  batch)
    torch.cuda.synchronize()
    assert st == A.SUCCESS and (info.cpu().numpy() == 0).all()
    bad = piv.clone()
    bad[n + 3] = 1  # matrix 1, step 3: pivot row above the step
    info.fill_(7)
    st = L.aclsolverSgetriBatched(handle.h, n, A.ptr(pa), n, A.ptr(bad), A.ptr(_ptrs(torch, outs)), n, A.ptr(info),
                                  batch)
    torch.cuda.synchronize()
    assert st == A.INVALID_VALUE and (info.cpu().numpy() == -4).all()
    info.fill_(7)
    st = L.aclsolverSgetriBatched(handle.h, n, A.ptr(pa), n, A.ptr(piv),
                                  A.ptr(_ptrs(torch, [outs[0], outs[0], outs[2], outs[3]])), n, A.ptr(info), batch)
    torch.cuda.synchronize()
    assert st == A.INVALID_VALUE and (info.cpu().numpy() == -5).all()
    # complex matinv: aliasing outputs -> argument 4
    cm = [torch.eye(n, dtype=torch.complex64, device="cuda") * 2 for _ in range(batch)]
    co = [torch.zeros((n, n), dtype=torch.complex64, device="cuda") for _ in range(batch)]
    info.fill_(7)
    st = L.aclsolverCmatinvBatched(handle.h, n, A.ptr(_ptrs(torch, cm)), n,
                                   A.ptr(_ptrs(torch, [co[0], co[0], co[2], co[3]])), n, A.ptr(info), batch)
    torch.cuda.synchronize()
    assert st == A.INVALID_VALUE and (info.cpu().numpy() == -4).all()


# ---------------------------------------------------------------------------------------
# thin shapes (reference src/sgetrf/sgetrf_kernel.cpp:307-512: m <= 32 < n, n > 128)
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("m,n,seed", [(32, 129, 1), (16, 300, 2), (1, 200, 3), (7, 1000, 4), (32, 4096, 5)])
def test_thin_shape_parity(torch, handle, m, n, seed):
    L = A.lib()
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((m, n)) / np.sqrt(m)).astype(np.float32)
    d = A.to_device_cm(a, ld=m + 1)
    lw = ctypes.c_int(0)
    assert L.aclsolverSgetrf_bufferSize(handle.h, m, n, A.ptr(d), m + 1, ctypes.byref(lw)) == A.SUCCESS
    work = torch.zeros(max(lw.value, 1), device="cuda")
    ipiv = torch.zeros(m, dtype=torch.int32, device="cuda")
    info = _info(torch)
    st = L.aclsolverSgetrf(handle.h, m, n, A.ptr(d), m + 1, A.ptr(work), A.ptr(ipiv), A.ptr(info))
    torch.cuda.synchronize()
    assert st == A.SUCCESS and info.item() == 0
    lu = A.from_device_cm(d, m, n)
    piv = ipiv.cpu().numpy()
    ref_lu, ref_piv, ref_info, margin = O.sgetrf(a, margins=True)
    bad = np.nonzero(margin < 1e-4)[0]
    upto = int(bad[0]) if len(bad) else m
    np.testing.assert_array_equal(piv[:upto], ref_piv[:upto])
    if upto == m:
        np.testing.assert_allclose(lu, ref_lu, rtol=0, atol=2e-5 * max(1.0, float(np.abs(ref_lu).max())))
    assert O.lu_backward_error(a, lu, piv) <= 10 * n * EPS32
