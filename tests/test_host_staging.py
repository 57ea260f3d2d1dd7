"""Fault injection on the host row-major getrs staging (ops-solver_amd/csrc/host_getrs.h),
the code behind aclsolverSgetrs / aclsolverCgetrs, compiled with g++ against a stub runtime
(tests/host/staged_getrs_driver.cpp) that fails one step per run. The contract follows the
reference's host getrs (src/sgetrs/sgetrs_host.cpp:131-195, src/utils/pivot_check.h:27-42)
and its cleanup rule (one owner per buffer, nothing freed while work may be in flight):
argument errors never touch the device, a failed step returns its status with info
untouched and the staging freed exactly once after the stream drained, a failed drain keeps
the staging."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INVALID, BAD_ALLOC, RT_FAILURE = 100000, 200000, 500000


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is needed for the host fault-injection driver")
    out = tmp_path_factory.mktemp("staging") / "driver"
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "ops-solver_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "staged_getrs_driver.cpp"), "-o", str(out)], check=True)
    return str(out)


def run(driver, failure, index=0):
    res = subprocess.run([driver, failure, str(index)], capture_output=True, text=True, check=True)
    line = next(l for l in res.stdout.splitlines() if l.startswith("RESULT "))
    f = line.split()[1:]
    status, allocs, frees, dup, unsafe, live, info = map(int, f[:7])
    x = np.array([float(v) for v in f[7:]]).reshape(3, 2)
    return status, allocs, frees, dup, unsafe, live, info, x


def expected_solution():
    lower = np.array([[1, 0, 0], [0.5, 1, 0], [0.25, 0.5, 1]])
    upper = np.array([[4, 2, 1], [0, 3, 2], [0, 0, 5]])
    perm = np.arange(3)
    for k, p in enumerate([2, 3, 3]):
        perm[[k, p - 1]] = perm[[p - 1, k]]
    a = np.empty((3, 3))
    a[perm] = lower @ upper  # P A = L U
    b = np.array([[1, 2], [3, 4], [5, 6]], float)
    return np.linalg.solve(a, b), b


def test_success_returns_solution_and_frees_once(driver):
    status, allocs, frees, dup, unsafe, live, info, x = run(driver, "none")
    want, _ = expected_solution()
    assert status == 0 and info == 0
    np.testing.assert_allclose(x, want, rtol=1e-5, atol=1e-6)
    assert (allocs, frees, dup, unsafe, live) == (1, 1, 0, 0, 0)


@pytest.mark.parametrize("failure", ["bad_n", "bad_ldb", "bad_trans"])
def test_argument_errors_touch_nothing(driver, failure):
    status, allocs, frees, dup, unsafe, live, info, x = run(driver, failure)
    _, b = expected_solution()
    assert status == INVALID and info == -99
    assert (allocs, frees, live) == (0, 0, 0)
    np.testing.assert_array_equal(x, b)


def test_bad_pivot_sets_info_minus_4(driver):
    status, allocs, frees, dup, unsafe, live, info, x = run(driver, "bad_ipiv")
    assert status == INVALID and info == -4
    assert (allocs, frees, live) == (0, 0, 0)


def test_allocation_failure(driver):
    status, allocs, frees, dup, unsafe, live, info, x = run(driver, "alloc", 1)
    _, b = expected_solution()
    assert status == BAD_ALLOC and info == -99
    assert (allocs, frees, live) == (1, 0, 0)
    np.testing.assert_array_equal(x, b)


@pytest.mark.parametrize("failure,index", [("clear", 0), ("copy", 1), ("copy", 2), ("copy", 3), ("launch", 0),
                                           ("copy", 4), ("copy", 5), ("wait", 0)])
def test_failed_step_frees_after_drain(driver, failure, index):
    status, allocs, frees, dup, unsafe, live, info, x = run(driver, failure, index)
    assert status == RT_FAILURE and info == -99
    assert (allocs, frees, dup, unsafe, live) == (1, 1, 0, 0, 0)
    if failure != "wait":
        _, b = expected_solution()
        np.testing.assert_array_equal(x, b)  # B is written back last, after a clean solve


def test_failed_drain_keeps_staging(driver):
    status, allocs, frees, dup, unsafe, live, info, x = run(driver, "sync")
    assert status == RT_FAILURE and info == -99
    assert (allocs, frees, dup, unsafe, live) == (1, 0, 0, 0, 1)
