"""CPU checks of the C-ABI library: it is built, loads without a GPU, and exports every
entry point that include/cann_ops_solver.h declares (no compute calls here)."""
import ctypes
import os
import subprocess

import pytest

import aclsolver


def test_library_built_and_exports_declared_symbols():
    assert os.path.exists(aclsolver.LIB_PATH), "build with `make -C ops-solver_amd` first"
    lib = ctypes.CDLL(aclsolver.LIB_PATH)
    declared = aclsolver.declared_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", aclsolver.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared) <= exported


def test_declarations_cite_reference():
    text = open(aclsolver.HEADER).read()
    # every non-extension declaration carries a reference line citation
    assert text.count("/* :") >= 15
    assert "EXTENSION" in text


def test_null_handle_contract_without_gpu():
    """Null-handle paths return before any HIP call (reference aclsolver_auxiliary.cpp)."""
    lib = aclsolver.lib()
    assert lib.aclsolverCreate(None) == aclsolver.HANDLE_IS_NULLPTR
    assert lib.aclsolverDestroy(None) == aclsolver.HANDLE_IS_NULLPTR
    assert lib.aclsolverSetStream(None, None) == aclsolver.HANDLE_IS_NULLPTR
    assert lib.aclsolverSgetrf(None, 4, 4, None, 4, None, None, None) == aclsolver.HANDLE_IS_NULLPTR
    lw = ctypes.c_int(0)
    assert lib.aclsolverSgetrf_bufferSize(None, 4, 4, None, 4, ctypes.byref(lw)) == aclsolver.HANDLE_IS_NULLPTR
    assert lib.aclsolverSmatinvBatched(None, 4, None, 4, None, 4, None, 1) == aclsolver.HANDLE_IS_NULLPTR


def test_host_getrs_rejects_bad_arguments_before_device_work():
    """aclsolverSgetrs validates on the host (src/sgetrs/sgetrs_host.cpp:139-151,
    src/utils/pivot_check.h:27-42): bad sizes -> INVALID_PARAM with info untouched, an
    ipiv entry outside [1, n] -> info = -4. No GPU is touched on these paths."""
    import numpy as np

    lib = aclsolver.lib()
    n = 4
    a = np.eye(n, dtype=np.float32)
    b = np.ones((n, 1), np.float32)
    info = np.array([7], np.int32)
    ipiv = np.array([1, 2, 3, 4], np.int32)
    # nrhs = 0 -> rejected, info untouched
    r = lib.aclsolverSgetrs(None, 0, n, 0, a.ctypes.data, n, ipiv.ctypes.data, b.ctypes.data, 1, info.ctypes.data)
    assert r == aclsolver.ACL_ERROR_INVALID_PARAM and info[0] == 7
    # trans = 2 -> rejected
    r = lib.aclsolverSgetrs(None, 2, n, 1, a.ctypes.data, n, ipiv.ctypes.data, b.ctypes.data, 1, info.ctypes.data)
    assert r == aclsolver.ACL_ERROR_INVALID_PARAM and info[0] == 7
    # lda - n > 16383 -> rejected
    r = lib.aclsolverSgetrs(None, 0, n, 1, a.ctypes.data, n + 16384, ipiv.ctypes.data, b.ctypes.data, 1,
                            info.ctypes.data)
    assert r == aclsolver.ACL_ERROR_INVALID_PARAM and info[0] == 7
    bad = np.array([1, 5, 3, 4], np.int32)
    r = lib.aclsolverSgetrs(None, 0, n, 1, a.ctypes.data, n, bad.ctypes.data, b.ctypes.data, 1, info.ctypes.data)
    assert r == aclsolver.ACL_ERROR_INVALID_PARAM and info[0] == -4
