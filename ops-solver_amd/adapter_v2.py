"""ctypes view of the operator-adapter protocol v2 (include/solver_adapter_v2.h) and a
loader for the per-operator adapter libraries (ops-solver_amd/adapter/libadapter_<op>.so).
Used by the tests; a harness in another language binds the same C symbols."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
OPS = ("sgetrf", "sgetri", "sgetri_batched", "smatinv_batched")
VERSION_V2 = 2
OK, INVALID_CASE, ALLOC_FAILED, RUNTIME_FAILED, QUERY_FAILED = 0, 1, 2, 3, 4


class Desc(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint32), ("struct_size", ctypes.c_uint32), ("op", ctypes.c_char_p),
                ("dtype", ctypes.c_char_p), ("dut_abi", ctypes.c_char_p), ("dut_layout", ctypes.c_char_p),
                ("info_kind", ctypes.c_char_p), ("target_symbol", ctypes.c_void_p)]


class Case(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint32), ("struct_size", ctypes.c_uint32), ("n", ctypes.c_int64),
                ("nrhs", ctypes.c_int64), ("batch", ctypes.c_int64), ("lda", ctypes.c_int64), ("ldb", ctypes.c_int64),
                ("uplo", ctypes.c_char), ("info_probe", ctypes.c_char_p), ("input_kind", ctypes.c_char_p),
                ("a_bytes", ctypes.c_size_t), ("b_bytes", ctypes.c_size_t), ("out_bytes", ctypes.c_size_t),
                ("info_count", ctypes.c_size_t)]


class CaseV2(ctypes.Structure):
    _fields_ = [("base", Case), ("m", ctypes.c_int64), ("k", ctypes.c_int64), ("ldc", ctypes.c_int64),
                ("ldx", ctypes.c_int64), ("side", ctypes.c_char), ("trans", ctypes.c_char),
                ("in_ipiv_count", ctypes.c_size_t), ("in_tau_count", ctypes.c_size_t),
                ("out_ipiv_count", ctypes.c_size_t), ("out_tau_count", ctypes.c_size_t)]


def path(op):
    return os.path.join(HERE, "adapter", f"libadapter_{op}.so")


def load(op):
    lib = ctypes.CDLL(path(op))
    P, S = ctypes.c_void_p, ctypes.c_size_t
    lib.solver_adapter_descriptor.restype = ctypes.POINTER(Desc)
    lib.solver_adapter_create_v2.argtypes = [ctypes.POINTER(CaseV2), P, P, ctypes.POINTER(P)]
    lib.solver_adapter_reset_v2.argtypes = [P, P, S, P, S, P, S, P, S]
    lib.solver_adapter_execute.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    lib.solver_adapter_readback_v2.argtypes = [P, P, S, P, S, P, S, P, S]
    lib.solver_adapter_destroy.argtypes = [P]
    return lib


def make_case(op, n, m=0, batch=1, lda=None, ldc=0, in_ipiv=0, out_ipiv=0):
    """A version-2 case for operator op with compact row-major float32 transport."""
    rows = m if op == "sgetrf" else n
    lda = lda if lda is not None else max(1, rows)
    c = CaseV2()
    c.base.version = VERSION_V2
    c.base.struct_size = ctypes.sizeof(CaseV2)
    c.base.n, c.base.nrhs, c.base.batch, c.base.lda, c.base.ldb = n, 0, batch, lda, 0
    c.base.uplo = b"N"
    c.base.a_bytes = c.base.out_bytes = rows * n * batch * 4
    c.base.b_bytes = 0
    c.base.info_count = batch if op in ("sgetri_batched", "smatinv_batched") else 1
    c.m, c.k, c.ldc, c.ldx = m, 0, ldc, 0
    c.side, c.trans = b"L", b"N"
    c.in_ipiv_count, c.out_ipiv_count = in_ipiv, out_ipiv
    return c
