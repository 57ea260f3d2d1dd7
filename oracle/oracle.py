"""ctypes front end of the CPU oracle (oracle/lu_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
import this module to check or time results; the shipped library never calls it.

Arrays are numpy, column-major (Fortran order). ipiv is 1-based int32, as in the
reference's getrf output (include/cann_ops_solver.h:82-99).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblu_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        p = ctypes.c_void_p
        i = ctypes.c_int
        _lib.orc_sgetrf.argtypes = [i, i, p, i, p, i, p]
        _lib.orc_sgetrf.restype = i
        _lib.orc_stslu_pivots.argtypes = [i, i, p, i, i, i, p]
        _lib.orc_stslu_pivots.restype = i
        _lib.orc_sgetrs.argtypes = [i, i, i, p, i, p, p, i]
        _lib.orc_sgetri.argtypes = [i, p, i, p]
        _lib.orc_sgetri.restype = i
        _lib.orc_smatinv.argtypes = [i, p, i, p, i, p]
        _lib.orc_smatinv.restype = i
        _lib.orc_dgetrf.argtypes = [i, i, p, i, p]
        _lib.orc_dgetrf.restype = i
        _lib.orc_dgetrs.argtypes = [i, i, i, p, i, p, p, i]
        _lib.orc_cgetrf.argtypes = [i, i, p, i, p, p]
        _lib.orc_cgetrf.restype = i
        _lib.orc_cgetrs.argtypes = [i, i, i, p, i, p, p, i]
        _lib.orc_cgetri.argtypes = [i, p, i, p]
        _lib.orc_cgetri.restype = i
        _lib.orc_cmatinv.argtypes = [i, p, i, p, i, p]
        _lib.orc_cmatinv.restype = i
        _lib.orc_zgetrf.argtypes = [i, i, p, i, p]
        _lib.orc_zgetrf.restype = i
        _lib.orc_cpivot_score.argtypes = [ctypes.c_float, ctypes.c_float]
        _lib.orc_cpivot_score.restype = ctypes.c_uint32
    return _lib


def _ptr(x):
    return None if x is None else x.ctypes.data


def _fortran(a, dtype):
    return np.array(a, dtype=dtype, order="F", copy=True)


def sgetrf(a, pivot=True, divide=False, margins=False):
    """Returns (lu, ipiv or None, info[, margin]) for a float32 matrix (m, n)."""
    lu = _fortran(a, np.float32)
    m, n = lu.shape
    k = min(m, n)
    ipiv = np.zeros(max(k, 1), np.int32) if pivot else None
    margin = np.zeros(max(k, 1), np.float32) if margins else None
    info = lib().orc_sgetrf(m, n, _ptr(lu), max(m, 1), _ptr(ipiv), int(divide), _ptr(margin))
    ipiv = ipiv[:k] if pivot else None
    if margins:
        return lu, ipiv, info, margin[:k]
    return lu, ipiv, info


def stslu_pivots(a, w, node_rows=256, group=8):
    """Tournament pivots of the leaf a[:, :w] (float32, m x w): (ipiv[0:w] 1-based, info).
    The tree of the library's tournament leaves: 256-row blocks, 8 children per node."""
    blk = _fortran(np.asarray(a)[:, :w], np.float32)
    m = blk.shape[0]
    ipiv = np.zeros(max(w, 1), np.int32)
    info = lib().orc_stslu_pivots(m, w, _ptr(blk), max(m, 1), node_rows, group, _ptr(ipiv))
    return ipiv[: min(m, w)], info


def dgetrf(a, pivot=True):
    lu = _fortran(a, np.float64)
    m, n = lu.shape
    k = min(m, n)
    ipiv = np.zeros(max(k, 1), np.int32) if pivot else None
    info = lib().orc_dgetrf(m, n, _ptr(lu), max(m, 1), _ptr(ipiv))
    return lu, (ipiv[:k] if pivot else None), info


def sgetrs(lu, ipiv, b, trans=0):
    lu = _fortran(lu, np.float32)
    x = _fortran(b, np.float32)
    if x.ndim == 1:
        x = x.reshape(-1, 1, order="F")
    n = lu.shape[0]
    piv = None if ipiv is None else np.ascontiguousarray(ipiv, np.int32)
    lib().orc_sgetrs(int(trans), n, x.shape[1], _ptr(lu), max(n, 1), _ptr(piv), _ptr(x), max(n, 1))
    return x


def dgetrs(lu, ipiv, b, trans=0):
    lu = _fortran(lu, np.float64)
    x = _fortran(b, np.float64)
    if x.ndim == 1:
        x = x.reshape(-1, 1, order="F")
    n = lu.shape[0]
    piv = None if ipiv is None else np.ascontiguousarray(ipiv, np.int32)
    lib().orc_dgetrs(int(trans), n, x.shape[1], _ptr(lu), max(n, 1), _ptr(piv), _ptr(x), max(n, 1))
    return x


def sgetri(lu, ipiv):
    """In-place LAPACK-order inverse from LU; returns (inverse, info)."""
    a = _fortran(lu, np.float32)
    n = a.shape[0]
    piv = None if ipiv is None else np.ascontiguousarray(ipiv, np.int32)
    info = lib().orc_sgetri(n, _ptr(a), max(n, 1), _ptr(piv))
    return a, info


def smatinv(a):
    """Direct inverse of an original matrix: returns (inverse, ipiv, info)."""
    a = _fortran(a, np.float32)
    n = a.shape[0]
    c = np.zeros((n, n), np.float32, order="F")
    piv = np.zeros(max(n, 1), np.int32)
    info = lib().orc_smatinv(n, _ptr(a), max(n, 1), _ptr(c), max(n, 1), _ptr(piv))
    return c, piv[:n], info


def cgetrf(a, pivot=True, margins=False):
    lu = _fortran(a, np.complex64)
    m, n = lu.shape
    k = min(m, n)
    ipiv = np.zeros(max(k, 1), np.int32) if pivot else None
    margin = np.ones(max(k, 1), np.float32) if margins else None
    info = lib().orc_cgetrf(m, n, _ptr(lu), max(m, 1), _ptr(ipiv), _ptr(margin))
    if margins:
        return lu, (ipiv[:k] if pivot else None), info, margin[:k]
    return lu, (ipiv[:k] if pivot else None), info


def zgetrf(a, pivot=True):
    lu = _fortran(a, np.complex128)
    m, n = lu.shape
    k = min(m, n)
    ipiv = np.zeros(max(k, 1), np.int32) if pivot else None
    info = lib().orc_zgetrf(m, n, _ptr(lu), max(m, 1), _ptr(ipiv))
    return lu, (ipiv[:k] if pivot else None), info


def cgetrs(lu, ipiv, b, trans=0):
    lu = _fortran(lu, np.complex64)
    x = _fortran(b, np.complex64)
    if x.ndim == 1:
        x = x.reshape(-1, 1, order="F")
    n = lu.shape[0]
    piv = None if ipiv is None else np.ascontiguousarray(ipiv, np.int32)
    lib().orc_cgetrs(int(trans), n, x.shape[1], _ptr(lu), max(n, 1), _ptr(piv), _ptr(x), max(n, 1))
    return x


def cgetri(lu, ipiv):
    a = _fortran(lu, np.complex64)
    n = a.shape[0]
    piv = None if ipiv is None else np.ascontiguousarray(ipiv, np.int32)
    info = lib().orc_cgetri(n, _ptr(a), max(n, 1), _ptr(piv))
    return a, info


def cmatinv(a):
    a = _fortran(a, np.complex64)
    n = a.shape[0]
    c = np.zeros((n, n), np.complex64, order="F")
    piv = np.zeros(max(n, 1), np.int32)
    info = lib().orc_cmatinv(n, _ptr(a), max(n, 1), _ptr(c), max(n, 1), _ptr(piv))
    return c, piv[:n], info


def cpivot_score(re, im):
    return int(lib().orc_cpivot_score(float(re), float(im)))


# ------------------------------------------------------------------------------------
# Error metrics (FP64, host)
# ------------------------------------------------------------------------------------


def apply_ipiv_rows(a, ipiv):
    """P*A: the row interchanges of ipiv applied in order to a copy of a."""
    p = np.array(a, copy=True)
    for k, v in enumerate(ipiv):
        r = int(v) - 1
        if r != k:
            p[[k, r]] = p[[r, k]]
    return p


def unpack_lu(lu):
    m, n = lu.shape
    k = min(m, n)
    low = np.tril(lu[:, :k], -1)
    low[np.arange(k), np.arange(k)] = 1
    up = np.triu(lu[:k, :])
    return low, up


def lu_backward_error(a, lu, ipiv):
    """||P A - L U||_1 / (n ||A||_1) in FP64."""
    a64 = np.asarray(a, np.complex128 if np.iscomplexobj(lu) else np.float64)
    low, up = unpack_lu(np.asarray(lu, a64.dtype))
    pa = apply_ipiv_rows(a64, ipiv) if ipiv is not None else a64
    n = max(a64.shape[1], 1)
    denom = np.linalg.norm(a64, 1) if a64.size else 1.0
    if denom == 0:
        denom = 1.0
    return float(np.linalg.norm(pa - low @ up, 1) / (n * denom))


def solve_backward_error(a, x, b):
    """max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), FP64."""
    a64 = np.asarray(a, np.complex128 if np.iscomplexobj(a) else np.float64)
    x64 = np.asarray(x, a64.dtype).reshape(a64.shape[1], -1)
    b64 = np.asarray(b, a64.dtype).reshape(a64.shape[0], -1)
    r = b64 - a64 @ x64
    an = np.linalg.norm(a64, np.inf)
    out = 0.0
    for j in range(x64.shape[1]):
        den = an * np.linalg.norm(x64[:, j], np.inf) + np.linalg.norm(b64[:, j], np.inf)
        out = max(out, float(np.linalg.norm(r[:, j], np.inf) / (den if den else 1.0)))
    return out


def inverse_residual(a, c):
    """||I - C A||_1 / (n ||A||_1 ||C||_1 eps32) (test/complex_solver/complex_solver_test.cpp:86-103)."""
    a64 = np.asarray(a, np.complex128 if np.iscomplexobj(a) else np.float64)
    c64 = np.asarray(c, a64.dtype)
    n = a64.shape[0]
    if n == 0:
        return 0.0
    eps = 2.0 ** -24
    num = np.linalg.norm(np.eye(n) - c64 @ a64, 1)
    den = n * np.linalg.norm(a64, 1) * np.linalg.norm(c64, 1) * eps
    return float(num / den)
