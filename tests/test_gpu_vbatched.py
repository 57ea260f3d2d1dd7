"""GPU checks of the uneven batched LU solve (getrf + getrs per matrix in one persistent
launch) through the C ABI, including hipGraph capture of the plan's execute step.

Bars: per-matrix ipiv identical to the oracle's sgetrf up to the first pivot tie
(margin < 1e-4); packed LU within rounding of the oracle when pivots agree; the solution's
backward error <= 10 n eps32 per right-hand side; info = first zero pivot for a singular
matrix; graph replays identical to the eager result.
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


def _batch(rng, sizes, nrhs, singular=()):
    mats, rhs = [], []
    for i, n in enumerate(sizes):
        a = (rng.standard_normal((n, n)) / np.sqrt(n) + 2 * np.eye(n)).astype(np.float32)
        if i in singular:
            a[:, n // 2] = 0
        mats.append(a)
        rhs.append(rng.standard_normal((n, nrhs)).astype(np.float32))
    return mats, rhs


def _upload(torch, mats, rhs):
    dA = [A.to_device_cm(m) for m in mats]
    dB = [A.to_device_cm(b) for b in rhs]
    pa = torch.tensor([t.data_ptr() for t in dA], dtype=torch.int64, device="cuda")
    pb = torch.tensor([t.data_ptr() for t in dB], dtype=torch.int64, device="cuda")
    return dA, dB, pa, pb


@pytest.mark.parametrize("seed,count", [(1, 300), (2, 64)])
def test_vbatched_parity(torch, seed, count):
    rng = np.random.default_rng(seed)
    sizes = [int(v) for v in rng.integers(1, 257, size=count)]
    sizes[:6] = [1, 64, 65, 128, 129, 256]
    nrhs = 4
    mats, rhs = _batch(rng, sizes, nrhs, singular=(7,))
    dA, dB, pa, pb = _upload(torch, mats, rhs)
    total = sum(sizes)
    ipiv = torch.zeros(total, dtype=torch.int32, device="cuda")
    info = torch.full((count,), -9, dtype=torch.int32, device="cuda")
    ns = np.array(sizes, np.int32)
    with A.Handle() as h:
        st = A.lib().aclsolverSgesvVbatched(h.h, count, ns.ctypes.data, nrhs, A.ptr(pa), A.ptr(pb), A.ptr(ipiv),
                                            A.ptr(info))
        torch.cuda.synchronize()
    assert st == A.SUCCESS
    hinfo = info.cpu().numpy()
    hpiv = ipiv.cpu().numpy()
    off = 0
    for i, n in enumerate(sizes):
        lu = A.from_device_cm(dA[i], n, n)
        x = A.from_device_cm(dB[i], n, nrhs)
        ref_lu, ref_piv, ref_info, margin = O.sgetrf(mats[i], margins=True)
        assert hinfo[i] == ref_info, (i, n)
        piv = hpiv[off:off + n]
        off += n
        if ref_info:
            continue
        bad = np.nonzero(margin < 1e-4)[0]
        upto = int(bad[0]) if len(bad) else n
        np.testing.assert_array_equal(piv[:upto], ref_piv[:upto])
        if upto == n:
            np.testing.assert_allclose(lu, ref_lu, rtol=0, atol=2e-5 * max(1.0, np.abs(ref_lu).max()) * np.sqrt(n))
        assert O.lu_backward_error(mats[i], lu, piv) <= 10 * n * EPS32
        assert O.solve_backward_error(mats[i], x, rhs[i]) <= 10 * n * EPS32


def test_vbatched_plan_graph_replay(torch):
    rng = np.random.default_rng(11)
    sizes = [int(v) for v in rng.integers(8, 257, size=200)]
    nrhs = 4
    mats, rhs = _batch(rng, sizes, nrhs)
    dA, dB, pa, pb = _upload(torch, mats, rhs)
    a0 = [t.clone() for t in dA]
    b0 = [t.clone() for t in dB]
    ipiv = torch.zeros(sum(sizes), dtype=torch.int32, device="cuda")
    info = torch.full((len(sizes),), -9, dtype=torch.int32, device="cuda")
    ns = np.array(sizes, np.int32)
    stream = torch.cuda.Stream()
    h = A.Handle(stream=stream.cuda_stream)
    plan = ctypes.c_void_p(None)
    assert A.lib().aclsolverSgesvVbatched_createPlan(h.h, len(sizes), ns.ctypes.data, nrhs,
                                                     ctypes.byref(plan)) == A.SUCCESS

    def restore():
        for t, s in zip(dA, a0):
            t.copy_(s)
        for t, s in zip(dB, b0):
            t.copy_(s)

    # eager reference
    torch.cuda.synchronize()
    restore()
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        assert A.lib().aclsolverSgesvVbatched_execute(h.h, plan, A.ptr(pa), A.ptr(pb), A.ptr(ipiv),
                                                      A.ptr(info)) == A.SUCCESS
    stream.synchronize()
    eager = [t.clone() for t in dB]
    eager_piv = ipiv.clone()
    # capture
    g = torch.cuda.CUDAGraph()
    restore()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=stream):
        assert A.lib().aclsolverSgesvVbatched_execute(h.h, plan, A.ptr(pa), A.ptr(pb), A.ptr(ipiv),
                                                      A.ptr(info)) == A.SUCCESS
    for rep in range(3):
        restore()
        ipiv.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for t, e in zip(dB, eager):
            assert torch.equal(t, e), f"replay {rep} differs from eager"
        assert torch.equal(ipiv, eager_piv)
        assert (info == 0).all()
    A.lib().aclsolverSgesvVbatched_destroyPlan(plan)
    h.close()


@pytest.mark.parametrize("nrhs", [0, 1, 3])
def test_vbatched_class_boundaries(torch, nrhs):
    """Orders at the edges of the size classes (one wave per system up to 64, then the
    column-cyclic classes 65..128, 129..192, 193..256), fewer than four right-hand sides,
    and one singular system per class (info = its first zero pivot)."""
    rng = np.random.default_rng(100 + nrhs)
    sizes = [63, 64, 65, 66, 127, 128, 129, 130, 191, 192, 193, 194, 255, 256, 64, 128, 192, 256]
    singular = (14, 15, 16, 17)
    mats, rhs = _batch(rng, sizes, max(nrhs, 1), singular=singular)
    if nrhs == 0:
        rhs = [np.zeros((n, 1), np.float32) for n in sizes]  # a buffer the solve must leave alone
    dA, dB, pa, pb = _upload(torch, mats, rhs)
    b_before = [t.clone() for t in dB]
    ipiv = torch.zeros(sum(sizes), dtype=torch.int32, device="cuda")
    info = torch.full((len(sizes),), -9, dtype=torch.int32, device="cuda")
    ns = np.array(sizes, np.int32)
    with A.Handle() as h:
        st = A.lib().aclsolverSgesvVbatched(h.h, len(sizes), ns.ctypes.data, nrhs, A.ptr(pa), A.ptr(pb),
                                            A.ptr(ipiv), A.ptr(info))
        torch.cuda.synchronize()
    assert st == A.SUCCESS
    hinfo = info.cpu().numpy()
    hpiv = ipiv.cpu().numpy()
    off = 0
    for i, n in enumerate(sizes):
        lu = A.from_device_cm(dA[i], n, n)
        ref_lu, ref_piv, ref_info, margin = O.sgetrf(mats[i], margins=True)
        piv = hpiv[off:off + n]
        off += n
        assert hinfo[i] == ref_info, (i, n, hinfo[i], ref_info)
        if i in singular:
            assert ref_info == n // 2 + 1
            continue
        bad = np.nonzero(margin < 1e-4)[0]
        upto = int(bad[0]) if len(bad) else n
        np.testing.assert_array_equal(piv[:upto], ref_piv[:upto])
        assert O.lu_backward_error(mats[i], lu, piv) <= 10 * n * EPS32, (i, n)
        if nrhs == 0:
            assert torch.equal(dB[i], b_before[i]), "B must not be touched when nrhs = 0"
        else:
            x = A.from_device_cm(dB[i], n, nrhs)
            assert O.solve_backward_error(mats[i], x, rhs[i]) <= 10 * n * EPS32, (i, n)


def _check_against_oracle(sizes, mats, rhs, dA, dB, hpiv, hinfo, nrhs):
    off = 0
    for i, n in enumerate(sizes):
        lu = A.from_device_cm(dA[i], n, n)
        x = A.from_device_cm(dB[i], n, nrhs)
        ref_lu, ref_piv, ref_info, margin = O.sgetrf(mats[i], margins=True)
        piv = hpiv[off:off + n]
        off += n
        assert hinfo[i] == ref_info, (i, n, hinfo[i], ref_info)
        if ref_info:
            continue
        bad = np.nonzero(margin < 1e-4)[0]
        upto = int(bad[0]) if len(bad) else n
        np.testing.assert_array_equal(piv[:upto], ref_piv[:upto])
        assert O.lu_backward_error(mats[i], lu, piv) <= 10 * n * EPS32, (i, n)
        assert O.solve_backward_error(mats[i], x, rhs[i]) <= 10 * n * EPS32, (i, n)


CFG_MAX_N = {1: 128, 2: 128, 3: 192, 4: 192, 5: 256}


@pytest.mark.parametrize("cfg", [1, 2, 3, 4, 5])
def test_vbatched_forced_schedule(torch, cfg):
    """Each configuration of the persistent launch on every order it can take: the bins it
    fits use it, the bins above use the 16-wave configuration. Sizes leave partial items
    (fewer matrices than an item holds) and one singular system."""
    rng = np.random.default_rng(500 + cfg)
    top = CFG_MAX_N[cfg]
    sizes = [int(v) for v in rng.integers(65, top + 1, size=41)] + [1, 33, 64, 65, top, 250]
    sizes += [int(v) for v in rng.integers(1, 257, size=13)]
    nrhs = 3
    mats, rhs = _batch(rng, sizes, nrhs, singular=(5,))
    dA, dB, pa, pb = _upload(torch, mats, rhs)
    ipiv = torch.zeros(sum(sizes), dtype=torch.int32, device="cuda")
    info = torch.full((len(sizes),), -9, dtype=torch.int32, device="cuda")
    ns = np.array(sizes, np.int32)
    schedule = np.array([cfg if 64 + 16 * (b + 1) <= top else 5 for b in range(12)], np.int32)
    with A.Handle() as h:
        plan = ctypes.c_void_p(None)
        assert A.lib().aclsolverSgesvVbatched_createPlanWithSchedule(
            h.h, len(sizes), ns.ctypes.data, nrhs, schedule.ctypes.data, ctypes.byref(plan)) == A.SUCCESS
        got = np.zeros(12, np.int32)
        assert A.lib().aclsolverSgesvVbatched_getSchedule(plan, got.ctypes.data) == A.SUCCESS
        np.testing.assert_array_equal(got, schedule)
        # two executions in a row: the queue re-arms itself after each launch
        for rep in range(2):
            if rep:
                for t, m in zip(dA, mats):
                    t.copy_(A.to_device_cm(m))
                for t, b in zip(dB, rhs):
                    t.copy_(A.to_device_cm(b))
                info.fill_(-9)
            assert A.lib().aclsolverSgesvVbatched_execute(h.h, plan, A.ptr(pa), A.ptr(pb), A.ptr(ipiv),
                                                          A.ptr(info)) == A.SUCCESS
            torch.cuda.synchronize()
            _check_against_oracle(sizes, mats, rhs, dA, dB, ipiv.cpu().numpy(), info.cpu().numpy(), nrhs)
        assert A.lib().aclsolverSgesvVbatched_destroyPlan(plan) == A.SUCCESS


def test_vbatched_autotuned_schedule_and_bad_schedules(torch):
    """The autotuned schedule fits every bin; schedules naming no configuration, or one too
    small for its bin, are refused."""
    ns = np.array([70, 200, 10], np.int32)
    with A.Handle() as h:
        plan = ctypes.c_void_p(None)
        assert A.lib().aclsolverSgesvVbatched_createPlan(h.h, 3, ns.ctypes.data, 1, ctypes.byref(plan)) == A.SUCCESS
        got = np.zeros(12, np.int32)
        assert A.lib().aclsolverSgesvVbatched_getSchedule(plan, got.ctypes.data) == A.SUCCESS
        for b, c in enumerate(got):
            assert 1 <= c <= 5 and CFG_MAX_N[int(c)] >= 64 + 16 * (b + 1), (b, c)
        A.lib().aclsolverSgesvVbatched_destroyPlan(plan)
        for bad in ([0] * 12, [6] * 12, [1] * 12, [5] * 11 + [4]):
            sched = np.array(bad, np.int32)
            p2 = ctypes.c_void_p(None)
            assert A.lib().aclsolverSgesvVbatched_createPlanWithSchedule(
                h.h, 3, ns.ctypes.data, 1, sched.ctypes.data, ctypes.byref(p2)) == A.INVALID_VALUE
