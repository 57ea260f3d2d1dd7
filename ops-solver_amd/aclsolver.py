"""Python binding of libops_solver_hip.so (the MI355X LU solver C ABI) over ctypes.

The entry points keep the reference's names and argument conventions
(include/cann_ops_solver.h of ops-solver); this module only marshals arguments. Device
memory comes from PyTorch (torch.cuda tensors), used purely as an allocator and stream
source. Column-major matrices are held as torch tensors of shape (cols, ld): element
(r, c) is t[c, r].

Importing this module does not touch the GPU. The library must have been built
(`make -C ops-solver_amd`); a missing library raises instead of falling back.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libops_solver_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cann_ops_solver.h")

SUCCESS = 0
INVALID_VALUE = 3
EXECUTION_FAILED = 5
ARCH_MISMATCH = 8
HANDLE_IS_NULLPTR = 9
INFO_WAIT_TIMEOUT = -1000
DEFAULT_SPIN_LIMIT = 1 << 20
NOALLOC_SCRATCH_BYTES = 8 << 20
ACL_ERROR_INVALID_PARAM = 100000

_lib = None


class LibraryMissing(RuntimeError):
    pass


def declared_symbols():
    """Entry points declared in include/cann_ops_solver.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"(?:aclsolverStatus_t|aclError)\s+(aclsolver\w+)\s*\(", text)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} is not built; run `make -C {HERE}`")
    L = ctypes.CDLL(LIB_PATH)
    P, I, I64, U64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64
    S = ctypes.c_int  # status enum
    sig = {
        "aclsolverCreate": ([P], S),
        "aclsolverDestroy": ([P], S),
        "aclsolverSetStream": ([P, P], S),
        "aclsolverGetStream": ([P, P], S),
        "aclsolverSetStrictLimits": ([P, I], S),
        "aclsolverSetRefinement": ([P, ctypes.c_double, I, I], S),
        "aclsolverSetBf16Split": ([P, I], S),
        "aclsolverSetButterfly": ([P, I], S),
        "aclsolverSgetrf_bufferSize": ([P, I, I, P, I, P], S),
        "aclsolverSgetrf": ([P, I, I, P, I, P, P, P], S),
        "aclsolverSgetri_bufferSize": ([P, I, P, I, P], S),
        "aclsolverSgetri": ([P, I, P, I, P, P, I, P], S),
        "aclsolverSgetriBatched": ([P, I, P, I, P, P, I, P, I], S),
        "aclsolverSmatinvBatched": ([P, I, P, I, P, I, P, I], S),
        "aclsolverSmatinvStridedBatched": ([P, I, P, I, ctypes.c_longlong, P, I, ctypes.c_longlong, P, P, I], S),
        "aclsolverSgetrs": ([P, U64, I64, I64, P, I64, P, P, I64, P], I),
        "aclsolverSgetrsDevice": ([P, I, I, I, P, I, P, P, I, P], S),
        "aclsolverSgetrsDeviceNoAlloc": ([P, I, I, I, P, I, P, P, I, P], S),
        "aclsolverSetTournamentPivoting": ([P, I], S),
        "aclsolverSgetrfTournament": ([P, I, I, P, I, P, P, P], S),
        "aclsolverCgetrf_bufferSize": ([P, I, I, P, I, P], S),
        "aclsolverCgetrf": ([P, I, I, P, I, P, P, P], S),
        "aclsolverCgetri_bufferSize": ([P, I, P, I, P], S),
        "aclsolverCgetri": ([P, I, P, I, P, P, I, P], S),
        "aclsolverCgetriBatched": ([P, I, P, I, P, P, I, P, I], S),
        "aclsolverCmatinvBatched": ([P, I, P, I, P, I, P, I], S),
        "aclsolverCgetrs": ([P, U64, I64, I64, P, I64, P, P, I64, P], I),
        "aclsolverCgetrsDevice": ([P, I, I, I, P, I, P, P, I, P], S),
        "aclsolverSHgesv_bufferSize": ([P, I, I, P], S),
        "aclsolverSBgesv_bufferSize": ([P, I, I, P], S),
        "aclsolverSSgesv_bufferSize": ([P, I, I, P], S),
        "aclsolverSHgesv": ([P, I, I, P, I, P, P, I, P, I, P, ctypes.c_size_t, P, P], S),
        "aclsolverSBgesv": ([P, I, I, P, I, P, P, I, P, I, P, ctypes.c_size_t, P, P], S),
        "aclsolverSSgesv": ([P, I, I, P, I, P, P, I, P, I, P, ctypes.c_size_t, P, P], S),
        "aclsolverSgesvVbatched_createPlan": ([P, I, P, I, P], S),
        "aclsolverSgesvVbatched_execute": ([P, P, P, P, P, P], S),
        "aclsolverSgesvVbatched_destroyPlan": ([P], S),
        "aclsolverSgesvVbatched_createPlanWithSchedule": ([P, I, P, I, P, P], S),
        "aclsolverSgesvVbatched_getSchedule": ([P, P], S),
        "aclsolverSgesvVbatched": ([P, I, P, I, P, P, P, P], S),
        "aclsolverProfileBegin": ([], S),
        "aclsolverProfileEnd": ([I, P, P, P], S),
        "aclsolverDebugSetSpinLimit": ([ctypes.c_uint], S),
    }
    for name, (args, res) in sig.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
    _lib = L
    return L


def ptr(t):
    """Device (or host numpy) address of a tensor/array, None for None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


class Handle:
    """aclsolverHandle_t bound to torch's current stream."""

    def __init__(self, stream=None):
        import torch

        self.h = ctypes.c_void_p(None)
        st = lib().aclsolverCreate(ctypes.byref(self.h))
        if st != SUCCESS:
            raise RuntimeError(f"aclsolverCreate failed: {st}")
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        lib().aclsolverSetStream(self.h, ctypes.c_void_p(s))

    def close(self):
        if self.h:
            lib().aclsolverDestroy(self.h)
            self.h = ctypes.c_void_p(None)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------------------
# numpy <-> device column-major helpers
# ---------------------------------------------------------------------------------------


def to_device_cm(a, ld=None, dtype=None):
    """numpy (rows, cols) -> torch tensor (cols, ld) holding the column-major matrix."""
    import numpy as np
    import torch

    a = np.asarray(a)
    rows, cols = a.shape
    ld = ld or max(rows, 1)
    tdtype = dtype or (torch.complex64 if np.iscomplexobj(a) else torch.float32)
    host = torch.zeros((max(cols, 1), ld), dtype=tdtype)
    if rows and cols:
        host[:cols, :rows] = torch.from_numpy(np.ascontiguousarray(a.T))
    return host.cuda()


def from_device_cm(t, rows, cols):
    """torch tensor (cols, ld) -> numpy (rows, cols)."""
    return t[:cols, :rows].cpu().numpy().T.copy()
