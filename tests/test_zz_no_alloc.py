"""aclsolverSgetrsDeviceNoAlloc: the device column-major solve whose staging lives in a fixed
per-GPU device array instead of handle scratch. Two calls in a row (the second smaller)
must give the same bits as aclsolverSgetrsDevice on the same factors, and both must agree
with the FP64 oracle solve within the getrs tolerance; sizes beyond the array are refused
with INVALID_VALUE and devInfo = -2 / -3."""
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


def _system(n, nrhs, seed):
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32)
    b = rng.standard_normal((n, nrhs)).astype(np.float32)
    lu, piv, info = O.sgetrf(a)
    assert info == 0
    return lu, piv, b


def _solve(torch, handle, fn, trans, lu, piv, b, ldb_pad):
    n, nrhs = b.shape
    dA = A.to_device_cm(lu)
    dB = A.to_device_cm(b, ld=n + ldb_pad)
    dpiv = torch.from_numpy(piv).cuda()
    dinfo = torch.full((1,), 9, dtype=torch.int32, device="cuda")
    st = getattr(A.lib(), fn)(handle.h, trans, n, nrhs, A.ptr(dA), n, A.ptr(dpiv), A.ptr(dB), n + ldb_pad,
                              A.ptr(dinfo))
    torch.cuda.synchronize()
    assert st == A.SUCCESS and dinfo.item() == 0, (fn, st, dinfo.item())
    return A.from_device_cm(dB, n, nrhs)


# (trans, first (n, nrhs), second smaller (n, nrhs)): one right-hand side of order >= 512
# takes the chained single-launch solve, the others the blocked solve.
CASES = [
    (0, (2049, 1), (600, 1)),
    (1, (1100, 5), (300, 3)),
    (0, (777, 9), (33, 2)),
]


@pytest.mark.parametrize("trans,first,second", CASES)
def test_noalloc_matches_allocating_entry_point(torch, handle, trans, first, second):
    for k, (n, nrhs) in enumerate((first, second)):
        print(f"noalloc case trans={trans} n={n} nrhs={nrhs} call={k}", flush=True)
        lu, piv, b = _system(n, nrhs, 1000 * trans + n + k)
        ref_dev = _solve(torch, handle, "aclsolverSgetrsDevice", trans, lu, piv, b, 3)
        got = _solve(torch, handle, "aclsolverSgetrsDeviceNoAlloc", trans, lu, piv, b, 3)
        # same kernels on the same data: the same bits
        np.testing.assert_array_equal(got, ref_dev)
        ref = O.dgetrs(lu.astype(np.float64), piv, b.astype(np.float64), trans=trans)
        tol = 50 * n * EPS32 * np.abs(ref).max() * max(1.0, np.linalg.cond(lu.astype(np.float64)) ** 0.5)
        np.testing.assert_allclose(got, ref, rtol=0, atol=tol)


def test_noalloc_refuses_sizes_beyond_its_array(torch, handle):
    L = A.lib()
    n = 16384
    per_rhs = 4 * n
    nrhs = A.NOALLOC_SCRATCH_BYTES // per_rhs + 1  # n * nrhs floats alone exceed the array
    dA = torch.zeros(1, device="cuda")
    dB = torch.zeros(1, device="cuda")
    dinfo = torch.full((1,), 9, dtype=torch.int32, device="cuda")
    st = L.aclsolverSgetrsDeviceNoAlloc(handle.h, 0, n, nrhs, A.ptr(dA), n, None, A.ptr(dB), n, A.ptr(dinfo))
    torch.cuda.synchronize()
    assert st == A.INVALID_VALUE and dinfo.item() == -3
    # empty problems succeed without touching the array
    dinfo.fill_(9)
    st = L.aclsolverSgetrsDeviceNoAlloc(handle.h, 0, 0, 1, None, 1, None, None, 1, A.ptr(dinfo))
    torch.cuda.synchronize()
    assert st == A.SUCCESS and dinfo.item() == 0
