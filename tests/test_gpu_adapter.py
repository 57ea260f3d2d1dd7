"""The operator adapters (protocol v2) driven end to end on the GPU, as an external test
package would: create -> reset -> execute -> (stream sync) -> readback -> destroy, with
compact row-major host transport. Results are checked against the oracle: sgetrf's LU and
1-based ipiv (bit-exact pivots on tie-free input), inverse residuals for sgetri,
sgetri_batched and smatinv_batched; the info outputs; a second reset/execute round on the
same state gives the same bytes."""
import ctypes

import numpy as np
import pytest

import adapter_v2 as AV
import aclsolver as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu


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


def _stream(handle):
    s = ctypes.c_void_p(None)
    assert A.lib().aclsolverGetStream(handle.h, ctypes.byref(s)) == A.SUCCESS
    return s


def _run(torch, handle, op, case, a_rm, ipiv_in=None, rounds=2):
    lib = AV.load(op)
    state = ctypes.c_void_p(None)
    rc = lib.solver_adapter_create_v2(ctypes.byref(case), handle.h, _stream(handle), ctypes.byref(state))
    assert rc == AV.OK, rc
    outs = []
    try:
        for _ in range(rounds):
            a = np.ascontiguousarray(a_rm, dtype=np.float32)
            piv = None if ipiv_in is None else np.ascontiguousarray(ipiv_in, dtype=np.int32)
            rc = lib.solver_adapter_reset_v2(state, a.ctypes.data, a.nbytes, None, 0,
                                             None if piv is None else piv.ctypes.data, 0 if piv is None else piv.size,
                                             None, 0)
            assert rc == AV.OK, rc
            st = ctypes.c_int(-1)
            assert lib.solver_adapter_execute(state, ctypes.byref(st)) == AV.OK
            assert st.value == A.SUCCESS, st.value
            torch.cuda.synchronize()
            out = np.empty_like(a)
            info = np.full(case.base.info_count, -7, np.int32)
            piv_out = np.full(case.out_ipiv_count, -7, np.int32)
            rc = lib.solver_adapter_readback_v2(state, out.ctypes.data, out.nbytes, info.ctypes.data, info.size,
                                                piv_out.ctypes.data if piv_out.size else None, piv_out.size, None, 0)
            assert rc == AV.OK, rc
            outs.append((out, info, piv_out))
    finally:
        assert lib.solver_adapter_destroy(state) == AV.OK
    assert all(np.array_equal(outs[0][0], o[0]) and np.array_equal(outs[0][2], o[2]) for o in outs[1:])
    return outs[0]


def test_sgetrf_adapter(torch, handle):
    m, n, lda = 300, 200, 310
    rng = np.random.default_rng(71)
    a = (rng.standard_normal((m, n)) / np.sqrt(n)).astype(np.float32)
    case = AV.make_case("sgetrf", n=n, m=m, lda=lda, out_ipiv=min(m, n))
    lu, info, piv = _run(torch, handle, "sgetrf", case, a)
    ref_lu, ref_piv, ref_info, margin = O.sgetrf(a, margins=True)
    assert info[0] == ref_info == 0
    upto = int(np.argmax(margin < 1e-4)) if (margin < 1e-4).any() else len(ref_piv)
    np.testing.assert_array_equal(piv[:upto], ref_piv[:upto])
    assert O.lu_backward_error(a, lu.reshape(m, n), piv) <= 10 * max(m, n) * 2.0 ** -24


def test_sgetri_adapter(torch, handle):
    n, lda = 100, 104
    rng = np.random.default_rng(72)
    a = (rng.standard_normal((n, n)) / np.sqrt(n) + 2 * np.eye(n)).astype(np.float32)
    lu, piv, info = O.sgetrf(a)
    case = AV.make_case("sgetri", n=n, lda=lda, in_ipiv=n)
    inv, info_out, _ = _run(torch, handle, "sgetri", case, lu, ipiv_in=piv)
    assert info_out[0] == 0
    assert O.inverse_residual(a, inv.reshape(n, n)) <= 5


@pytest.mark.parametrize("op,n,batch,lda,ldc", [("sgetri_batched", 20, 50, 22, 24),
                                                ("smatinv_batched", 32, 1000, 33, 35),
                                                ("smatinv_batched", 5, 7, 5, 5)])
def test_batched_adapters(torch, handle, op, n, batch, lda, ldc):
    rng = np.random.default_rng(73 + n)
    mats = (rng.standard_normal((batch, n, n)) / np.sqrt(n) + 2 * np.eye(n)).astype(np.float32)
    pivs = None
    src = mats
    if op == "sgetri_batched":
        lus, pv = [], []
        for b in range(batch):
            lu, piv, info = O.sgetrf(mats[b])
            lus.append(lu)
            pv.append(piv)
        src, pivs = np.stack(lus), np.concatenate(pv)
    case = AV.make_case(op, n=n, batch=batch, lda=lda, ldc=ldc, in_ipiv=0 if pivs is None else pivs.size)
    out, info, _ = _run(torch, handle, op, case, src, ipiv_in=pivs)
    assert (info == 0).all()
    out = out.reshape(batch, n, n)
    for b in range(0, batch, max(1, batch // 17)):
        assert O.inverse_residual(mats[b], out[b]) <= 5, b
