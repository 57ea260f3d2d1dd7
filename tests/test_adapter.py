"""CPU checks of the operator adapters (protocol v2, include/solver_adapter_v2.h): each
per-operator library loads, describes its operator (name, float32, device, column-major,
the library's own function address) and rejects malformed cases before touching a device;
destroy accepts a null or partially built state. The full protocol runs in
test_gpu_adapter.py."""
import ctypes

import pytest

import aclsolver
import adapter_v2 as AV


@pytest.mark.parametrize("op", AV.OPS)
def test_descriptor(op):
    lib = AV.load(op)
    d = lib.solver_adapter_descriptor().contents
    assert d.version == AV.VERSION_V2 and d.struct_size == ctypes.sizeof(AV.Desc)
    assert d.op.decode() == op.replace("_", "")
    assert (d.dtype, d.dut_abi, d.dut_layout) == (b"float32", b"device", b"column_major")
    assert d.info_kind == (b"array" if "batched" in op else b"scalar")
    target = {"sgetrf": "aclsolverSgetrf", "sgetri": "aclsolverSgetri", "sgetri_batched": "aclsolverSgetriBatched",
              "smatinv_batched": "aclsolverSmatinvBatched"}[op]
    fn = getattr(ctypes.CDLL(aclsolver.LIB_PATH), target)
    assert d.target_symbol == ctypes.cast(fn, ctypes.c_void_p).value


def _bad_cases(op):
    good = dict(n=8, m=8 if op == "sgetrf" else 0, batch=3 if "batched" in op else 1,
                ldc=8 if "batched" in op else 0)
    cases = []
    c = AV.make_case(op, **good)
    c.base.version = 1
    cases.append(("version", c))
    c = AV.make_case(op, **good)
    c.base.struct_size -= 8
    cases.append(("struct_size", c))
    c = AV.make_case(op, **good)
    c.base.lda = 4
    cases.append(("lda", c))
    c = AV.make_case(op, **good)
    c.base.a_bytes += 4
    cases.append(("a_bytes", c))
    c = AV.make_case(op, **good)
    c.base.b_bytes = 4
    cases.append(("b_bytes", c))
    c = AV.make_case(op, **good)
    c.in_tau_count = 1
    cases.append(("tau", c))
    c = AV.make_case(op, **good)
    c.base.info_count += 1
    cases.append(("info_count", c))
    c = AV.make_case(op, **good)
    c.out_ipiv_count = 5  # neither 0 nor min(m, n) * batch
    cases.append(("pivots", c))
    if "batched" in op:
        c = AV.make_case(op, **good)
        c.ldc = 4
        cases.append(("ldc", c))
    else:
        c = AV.make_case(op, **good)
        c.base.batch = 2
        c.base.a_bytes = c.base.out_bytes = 8 * 8 * 2 * 4
        cases.append(("batch", c))
    if op == "smatinv_batched":
        cases.append(("order", AV.make_case(op, n=65, batch=1, ldc=65)))
    return cases


@pytest.mark.parametrize("op", AV.OPS)
def test_malformed_cases_are_rejected_and_state_is_destroyable(op):
    lib = AV.load(op)
    fake_handle = ctypes.c_void_p(0x1000)  # never dereferenced: validation comes first
    for name, case in _bad_cases(op):
        state = ctypes.c_void_p(None)
        rc = lib.solver_adapter_create_v2(ctypes.byref(case), fake_handle, None, ctypes.byref(state))
        assert rc == AV.INVALID_CASE, name
        assert state.value is not None, name  # published before validation
        assert lib.solver_adapter_destroy(state) == AV.OK, name


@pytest.mark.parametrize("op", AV.OPS)
def test_null_arguments(op):
    lib = AV.load(op)
    case = AV.make_case(op, n=4, m=4 if op == "sgetrf" else 0)
    state = ctypes.c_void_p(None)
    assert lib.solver_adapter_create_v2(ctypes.byref(case), None, None, ctypes.byref(state)) == AV.INVALID_CASE
    assert state.value is None
    assert lib.solver_adapter_create_v2(ctypes.byref(case), ctypes.c_void_p(0x1000), None, None) == AV.INVALID_CASE
    assert lib.solver_adapter_destroy(None) == AV.OK
    st = ctypes.c_int(7)
    assert lib.solver_adapter_execute(None, ctypes.byref(st)) == AV.INVALID_CASE and st.value == 7
    assert lib.solver_adapter_reset_v2(None, None, 0, None, 0, None, 0, None, 0) == AV.INVALID_CASE
    assert lib.solver_adapter_readback_v2(None, None, 0, None, 0, None, 0, None, 0) == AV.INVALID_CASE
