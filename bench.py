#!/usr/bin/env python3
"""Headline benchmark of the MI355X LU solver library.

Metric: LU solve GFLOP/s at n=16384, BF16 MFMA factorization + GMRES-IR (FP64 residual)
to a backward error <= 1e-6, beside the plain FP32 getrf+getrs on the same GPU.

Workload (one step = one complete solve A x = b from an FP32 A already resident in HBM):
  A = Q1 diag(s) Q2^T, n = 16384, Q1/Q2 from Householder QR of seeded N(0,1) matrices,
  s log-spaced from 1 to 1e6 (cond 1e6); b = A x_true in FP64 (rounded to FP32),
  x_true ~ N(0,1). value = (2/3 n^3 + 2 n^2) / wall time of the step (nominal LAPACK
  count; refinement work is not credited).

Also reported: the FP32 solve (SSgesv) time and ratio, backward/forward errors, the
trailing-GEMM roofline (per-launch HIP events inside the timed region), the host LAPACK
baseline on a bounded sample, and the batched-inverse configuration (65536 x 32x32).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "ops-solver_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "LU solve GFLOP/s n=16384 BF16+IR vs FP32 at 1e-6 bwd err; batched solves/s"
PEAK_BF16 = 2.5e15  # dense MFMA, FLOP/s
PEAK_HBM = 8.0e12  # B/s


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_system(torch, n, seed, kind="qr"):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    if kind == "shifted":
        # N(0,1)/sqrt(n) + 2I: well conditioned, no QR; the LU runs the same launch shapes
        # (used for the counter passes, which slow every dispatch of the QR down).
        at = torch.randn(n, n, device="cuda", generator=g) / n ** 0.5
        at.diagonal().add_(2.0)
        x_true = torch.randn(n, device="cuda", generator=g, dtype=torch.float64)
        b = (at.T.double() @ x_true).float().contiguous()
        return at, b, x_true
    G1 = torch.randn(n, n, device="cuda", generator=g)
    Q1, _ = torch.linalg.qr(G1)
    del G1
    G2 = torch.randn(n, n, device="cuda", generator=g)
    Q2, _ = torch.linalg.qr(G2)
    del G2
    s = torch.logspace(0, 6, n, device="cuda", dtype=torch.float32)
    a = (Q1 * s) @ Q2.T  # row-major A
    del Q1, Q2
    at = a.T.contiguous()  # row-major A^T == column-major A
    del a
    x_true = torch.randn(n, device="cuda", generator=g, dtype=torch.float64)
    b64 = at.T.double() @ x_true
    b = b64.float().contiguous()
    return at, b, x_true


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "gemm_packed_bf16x6_n16384_traffic.md")


def measured_traffic(n):
    """HBM bytes per launch of the packed trailing-update GEMM at n=16384, from the committed
    FETCH_SIZE / WRITE_SIZE counter summary (scaled by the calibration copy); None otherwise."""
    if n != 16384 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def errors(torch, at, x, b, x_true):
    a64 = at.T.double()
    x64 = x.double()
    r = b.double() - a64 @ x64
    anorm = a64.abs().sum(dim=1).max()
    berr = (r.abs().max() / (anorm * x64.abs().max() + b.double().abs().max())).item()
    ferr = ((x64 - x_true).abs().max() / x_true.abs().max()).item()
    del a64
    return berr, ferr


def gesv(A, torch, h, name, n, at, b, x, work, lwork, dipiv, dinfo):
    niter = ctypes.c_int(0)
    f = getattr(A.lib(), f"aclsolver{name}gesv")
    st = f(h.h, n, 1, A.ptr(at), n, A.ptr(dipiv), A.ptr(b), n, A.ptr(x), n, A.ptr(work), lwork,
           ctypes.byref(niter), A.ptr(dinfo))
    if st != A.SUCCESS:
        raise RuntimeError(f"{name}gesv status {st}")
    return niter.value


def time_mode(A, torch, h, name, steps, warmup, n, at, b, x, work, lwork, dipiv, dinfo, profile=False):
    iters = 0
    for wi in range(warmup):
        t0 = time.perf_counter()
        it = gesv(A, torch, h, name, n, at, b, x, work, lwork, dipiv, dinfo)
        torch.cuda.synchronize()
        log(f"{name} warmup {wi + 1}/{warmup}: {(time.perf_counter() - t0) * 1e3:.1f} ms, {it} GMRES iterations")
    if profile:
        A.lib().aclsolverProfileBegin()
    t0 = time.perf_counter()
    for s in range(steps):
        iters = gesv(A, torch, h, name, n, at, b, x, work, lwork, dipiv, dinfo)
        log(f"{name} step {s + 1}/{steps} done ({iters} GMRES iterations)")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = None
    if profile:
        prof = {}
        for kind, key in ((0, "gemm"), (3, "leaf")):
            launches, work_f, ms = ctypes.c_int(0), ctypes.c_double(0), ctypes.c_double(0)
            A.lib().aclsolverProfileEnd(kind, ctypes.byref(launches), ctypes.byref(work_f), ctypes.byref(ms))
            prof[key] = (launches.value, work_f.value, ms.value)
    info = int(dinfo.item())
    return (t1 - t0) * 1e3 / steps, iters, info, prof


def _blas_threads():
    try:
        from threadpoolctl import threadpool_info

        return max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        return len(os.sched_getaffinity(0))


def _host_workers():
    # the host share of one GPU on the box (16 cores), or fewer where fewer are available
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline(n_cpu, reps=3):
    """Host LAPACK sgetrf+sgetrs (SciPy/OpenBLAS, the CPU path of the reference's tests)
    on one N(0,1)/sqrt(n) system of order n_cpu: median of reps runs after 1 warm-up, on
    every BLAS thread."""
    import numpy as np
    import scipy.linalg.lapack as la

    threads = _blas_threads()
    rng = np.random.default_rng(4)
    a = np.asfortranarray(rng.standard_normal((n_cpu, n_cpu)).astype(np.float32) / np.sqrt(n_cpu))
    b = rng.standard_normal((n_cpu, 1)).astype(np.float32)
    times = []
    for rep in range(reps + 1):
        t0 = time.perf_counter()
        lu, piv, info = la.sgetrf(a)
        x, info2 = la.sgetrs(lu, piv, b)
        t = time.perf_counter() - t0
        if rep:
            times.append(t)
    t = sorted(times)[len(times) // 2]
    flops = 2.0 / 3.0 * n_cpu ** 3 + 2.0 * n_cpu ** 2
    return {"value": round(flops / t / 1e9, 2), "unit": "GFLOP/s", "cores": int(threads), "kind": "reference",
            "ms": round(t * 1e3, 3),
            "sample": f"host LAPACK (SciPy OpenBLAS) sgetrf+sgetrs, one N(0,1)/sqrt(n) system of n={n_cpu}, "
                      f"median of {reps} runs ({t:.3f} s each) on {threads} threads"}


def cpu_complex_baseline(n=8192, nrhs=16, reps=2):
    """Host LAPACK cgetrf+cgetrs (SciPy/OpenBLAS) on the complex workload's distribution."""
    import numpy as np
    import scipy.linalg.lapack as la

    threads = _blas_threads()
    rng = np.random.default_rng(5)
    a = ((rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * np.sqrt(0.5 / n)).astype(np.complex64)
    a[np.diag_indices(n)] += 1.5
    a = np.asfortranarray(a)
    b = ((rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))).astype(np.complex64))
    times = []
    for rep in range(reps + 1):
        t0 = time.perf_counter()
        lu, piv, info = la.cgetrf(a)
        x, info2 = la.cgetrs(lu, piv, b)
        t = time.perf_counter() - t0
        if rep:
            times.append(t)
    t = sorted(times)[len(times) // 2]
    flops = 8.0 / 3.0 * n ** 3 + 8.0 * n * n * nrhs
    return {"value": round(flops / t / 1e9, 2), "unit": "GFLOP/s", "cores": int(threads), "kind": "reference",
            "ms": round(t * 1e3, 3),
            "sample": f"host LAPACK (SciPy OpenBLAS) cgetrf+cgetrs, n={n}, {nrhs} RHS, median of {reps} runs "
                      f"({t:.2f} s each) on {threads} threads"}


def _parallel_over_matrices(fn, items):
    """Runs fn(chunk) over chunks of items on a thread pool with single-threaded BLAS (the
    LAPACK calls release the GIL), as OpenMP over matrices would."""
    from concurrent.futures import ThreadPoolExecutor

    import contextlib

    workers = _host_workers()
    chunks = [items[i::workers] for i in range(workers)]
    try:
        from threadpoolctl import threadpool_limits

        ctx = threadpool_limits(limits=1, user_api="blas")
    except ImportError:
        ctx = contextlib.nullcontext()
    with ctx:
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=workers) as ex:
            list(ex.map(fn, chunks))
        return time.perf_counter() - t0, workers


def cpu_batched_inverse_baseline(sample=16384, n=32):
    """Host LAPACK inverses (numpy linalg.inv: sgesv per matrix) of `sample` matrices of the
    uniform-batch distribution, split over the host cores with single-threaded LAPACK."""
    import numpy as np

    rng = np.random.default_rng(32)
    a = (rng.standard_normal((sample, n, n)) / np.sqrt(n) + 2 * np.eye(n)).astype(np.float32)
    idx = list(range(0, sample, 256))

    def run(starts):
        for s0 in starts:
            np.linalg.inv(a[s0:s0 + 256])

    run(idx[:1])  # warm-up
    t, workers = _parallel_over_matrices(run, idx)
    return {"value": round(sample / t, 1), "unit": "matrices/s", "cores": workers, "kind": "reference",
            "sample": f"numpy linalg.inv (LAPACK sgesv) on {sample} of the 65536 32x32 matrices, "
                      f"{workers} threads over matrices, single-threaded LAPACK each ({t:.2f} s)"}


def cpu_uneven_baseline(sizes, nrhs=4, sample=2000):
    """Host LAPACK sgetrf+sgetrs (SciPy) on the first `sample` systems of the uneven batch,
    split over the host cores with single-threaded LAPACK."""
    import numpy as np
    import scipy.linalg.lapack as la

    rng = np.random.default_rng(20001)
    mats = []
    for ni in sizes[:sample]:
        ni = int(ni)
        a = np.asfortranarray((rng.standard_normal((ni, ni)) / np.sqrt(ni) + 2 * np.eye(ni)).astype(np.float32))
        b = np.asfortranarray(rng.standard_normal((ni, nrhs)).astype(np.float32))
        mats.append((a, b))

    def run(chunk):
        for a, b in chunk:
            lu, piv, info = la.sgetrf(a)
            la.sgetrs(lu, piv, b)

    run(mats[:4])
    t, workers = _parallel_over_matrices(run, mats)
    return {"value": round(sample / t, 1), "unit": "solves/s", "cores": workers, "kind": "reference",
            "sample": f"SciPy LAPACK sgetrf+sgetrs on the first {sample} of the 20000 systems (n_i in [8, 256], "
                      f"4 RHS), {workers} threads over systems, single-threaded LAPACK each ({t:.2f} s)"}


def small_dense(A, torch, h, n=512, steps=20):
    """The plain FP32 sgetrf + sgetrs at n = 512 (N(0,1)/sqrt(n), one RHS), drop-in entry
    points: sgetrf (with its finite scan) and the Device getrs; A restored outside the
    events each step."""
    import numpy as np

    L = A.lib()
    rng = np.random.default_rng(512)
    a0 = torch.from_numpy((rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32)).cuda()
    b0 = torch.from_numpy(rng.standard_normal(n).astype(np.float32)).cuda()
    a, b = a0.clone(), b0.clone()
    lw = ctypes.c_int(0)
    assert L.aclsolverSgetrf_bufferSize(h.h, n, n, A.ptr(a), n, ctypes.byref(lw)) == A.SUCCESS
    work = torch.empty(lw.value, device="cuda")
    ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    info = torch.empty(1, dtype=torch.int32, device="cuda")
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for s in range(steps + 2):
        a.copy_(a0)
        b.copy_(b0)
        torch.cuda.synchronize()
        ev0.record()
        assert L.aclsolverSgetrf(h.h, n, n, A.ptr(a), n, A.ptr(work), A.ptr(ipiv), A.ptr(info)) == A.SUCCESS
        assert L.aclsolverSgetrsDevice(h.h, 0, n, 1, A.ptr(a), n, A.ptr(ipiv), A.ptr(b), n, A.ptr(info)) == A.SUCCESS
        ev1.record()
        torch.cuda.synchronize()
        if s >= 2:
            times.append(ev0.elapsed_time(ev1))
    ms = sorted(times)[len(times) // 2]
    a64 = a0.T.double()
    x64 = b.double()
    r = b0.double() - a64 @ x64
    berr = (r.abs().max() / (a64.abs().sum(1).max() * x64.abs().max() + b0.double().abs().max())).item()
    flops = 2.0 / 3.0 * n ** 3 + 2.0 * n * n
    return {"workload": f"sgetrf+sgetrs FP32 n={n}, N(0,1)/sqrt(n), 1 RHS (drop-in sgetrf incl. finite scan)",
            "ms": round(ms, 4), "GFLOP_per_s": round(flops / ms / 1e6, 1), "backward_error": berr,
            "info": int(info.item())}


def batched_inverse(A, torch, h, steps=20):
    """65536 FP32 32x32 matrices, A_i = G_i/sqrt(32) + 2I: direct batched inverse."""
    n, batch = 32, 65536
    g = torch.Generator(device="cuda")
    g.manual_seed(32)
    a = torch.randn(batch, n, n, device="cuda", generator=g) / n ** 0.5 + 2 * torch.eye(n, device="cuda")
    a = a.contiguous()
    c = torch.empty_like(a)
    ipiv = torch.empty(batch, n, dtype=torch.int32, device="cuda")
    info = torch.empty(batch, dtype=torch.int32, device="cuda")
    L = A.lib()

    def run():
        st = L.aclsolverSmatinvStridedBatched(h.h, n, A.ptr(a), n, n * n, A.ptr(c), n, n * n, A.ptr(ipiv),
                                              A.ptr(info), batch)
        assert st == A.SUCCESS

    run()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps):
        run()
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    # The drop-in entry point: Device arrays of Device pointers (pointer-array validation
    # included in the timed calls).
    pa = (a.data_ptr() + torch.arange(batch, device="cuda", dtype=torch.int64) * n * n * 4)
    pc = (c.data_ptr() + torch.arange(batch, device="cuda", dtype=torch.int64) * n * n * 4)

    def run_ptr():
        st = L.aclsolverSmatinvBatched(h.h, n, A.ptr(pa), n, A.ptr(pc), n, A.ptr(info), batch)
        assert st == A.SUCCESS

    run_ptr()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run_ptr()
    torch.cuda.synchronize()
    ms_ptr = (time.perf_counter() - t0) * 1e3 / steps
    # check a sample against A C = I
    idx = torch.arange(0, batch, 4093, device="cuda")
    res = (torch.bmm(a[idx].transpose(1, 2).contiguous().double(), c[idx].transpose(1, 2).contiguous().double())
           - torch.eye(n, device="cuda", dtype=torch.float64)).abs().max().item()
    bytes_ = batch * (2 * n * n * 4 + n * 4 + 4)
    return {"workload": "smatinv 65536 x 32x32 FP32, A_i = G_i/sqrt(32) + 2I (output inverse + ipiv + info)",
            "ms": round(ms, 4), "matrices_per_s": round(batch / ms * 1e3, 1),
            "GB_per_s": round(bytes_ / ms / 1e6, 1), "hbm_frac": round(bytes_ / ms / 1e-3 / PEAK_HBM, 4),
            "max_abs_AC_minus_I_sample": res, "all_info_zero": bool((info == 0).all().item()),
            "dropin_pointer_array_ms": round(ms_ptr, 4),
            "dropin_note": "aclsolverSmatinvBatched (pointer arrays, host wall clock per call incl. validation)"}


def complex_solve(A, torch, h, n=8192, nrhs=16, steps=2):
    """cgetrf + cgetrs (device column-major) of one complex64 system of order n with nrhs
    right-hand sides, re/im i.i.d. N(0, 1/2)/sqrt(n) plus 1.5 on the diagonal; and
    cmatinv_batched on 4096 complex64 matrices of 64x64 with the same distribution."""
    L = A.lib()
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    a0 = torch.complex(torch.randn(n, n, device="cuda", generator=g), torch.randn(n, n, device="cuda", generator=g))
    a0 = a0 * (0.5 ** 0.5 / n ** 0.5) + 1.5 * torch.eye(n, device="cuda", dtype=torch.complex64)
    b0 = torch.complex(torch.randn(nrhs, n, device="cuda", generator=g), torch.randn(nrhs, n, device="cuda", generator=g))
    a = torch.empty_like(a0)  # rows of the row-major tensor = columns of the column-major matrix
    b = torch.empty_like(b0)
    lw = ctypes.c_int(0)
    assert L.aclsolverCgetrf_bufferSize(h.h, n, n, A.ptr(a), n, ctypes.byref(lw)) == A.SUCCESS
    work = torch.empty(max(lw.value, 1), dtype=torch.complex64, device="cuda")
    ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    info = torch.empty(1, dtype=torch.int32, device="cuda")
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for s in range(steps + 1):
        a.copy_(a0)
        b.copy_(b0)
        torch.cuda.synchronize()
        if s == 1:
            L.aclsolverProfileBegin()
        ev0.record()
        assert L.aclsolverCgetrf(h.h, n, n, A.ptr(a), n, A.ptr(work), A.ptr(ipiv), A.ptr(info)) == A.SUCCESS
        assert L.aclsolverCgetrsDevice(h.h, 0, n, nrhs, A.ptr(a), n, A.ptr(ipiv), A.ptr(b), n, A.ptr(info)) == A.SUCCESS
        ev1.record()
        torch.cuda.synchronize()
        if s:
            times.append(ev0.elapsed_time(ev1))
        log(f"complex n={n} step {s}: {ev0.elapsed_time(ev1):.2f} ms")
    ms = sum(times) / len(times)
    launches, gwork, gms = ctypes.c_int(0), ctypes.c_double(0), ctypes.c_double(0)
    L.aclsolverProfileEnd(4, ctypes.byref(launches), ctypes.byref(gwork), ctypes.byref(gms))
    # backward error in complex128 (column-major A: matrix = a0^T), and the forward error
    # against an FP64 solve of the same system (torch.linalg.solve in complex128: LU with
    # partial pivoting, zgetrf + zgetrs)
    a64 = a0.T.to(torch.complex128)
    x64 = b.T.to(torch.complex128)
    b64 = b0.T.to(torch.complex128)
    r = b64 - a64 @ x64
    berr = (r.abs().max() / (a64.abs().sum(1).max() * x64.abs().max() + b0.abs().max())).item()
    xref = torch.linalg.solve(a64, b64)
    ferr = ((x64 - xref).abs().max() / xref.abs().max()).item()
    del a64, x64, b64, r, xref
    flops = 8.0 / 3.0 * n ** 3 + 8.0 * n * n * nrhs
    peak_fp32 = 157.3e12
    gemm_rate = gwork.value / max(gms.value * 1e-3, 1e-12)
    out = {"workload": f"cgetrf+cgetrs complex64 n={n}, {nrhs} RHS, re/im N(0,1/2)/sqrt(n) + 1.5 I",
           "ms": round(ms, 3), "GFLOP_per_s": round(flops / ms / 1e6, 1), "backward_error": berr,
           "forward_error_vs_fp64_solve": ferr, "info": int(info.item()),
           "trailing_gemm": {"kernel": "cgemm_sub_kernel: complex trailing update from 4 real FP32 MFMA products "
                                       "(v_mfma_f32_32x32x2f32), FP32 accumulate",
                             "launches_per_solve": launches.value // max(steps, 1),
                             "ms_per_solve": round(gms.value / max(steps, 1), 3),
                             "achieved": round(gemm_rate / 1e12, 2), "peak": peak_fp32 / 1e12, "unit": "TFLOP/s",
                             "mfma_frac": round(gemm_rate / peak_fp32, 4)}}
    del a, a0, b, b0, work
    torch.cuda.empty_cache()
    # cmatinv_batched 4096 x 64x64
    nb, batch = 64, 4096
    m0 = torch.complex(torch.randn(batch, nb, nb, device="cuda", generator=g),
                       torch.randn(batch, nb, nb, device="cuda", generator=g))
    m0 = (m0 * (0.5 ** 0.5 / nb ** 0.5) + 1.5 * torch.eye(nb, device="cuda", dtype=torch.complex64)).contiguous()
    mi = torch.empty_like(m0)
    pa = (m0.data_ptr() + torch.arange(batch, device="cuda", dtype=torch.int64) * nb * nb * 8)
    pc = (mi.data_ptr() + torch.arange(batch, device="cuda", dtype=torch.int64) * nb * nb * 8)
    binfo = torch.empty(batch, dtype=torch.int32, device="cuda")

    def inv():
        assert L.aclsolverCmatinvBatched(h.h, nb, A.ptr(pa), nb, A.ptr(pc), nb, A.ptr(binfo), batch) == A.SUCCESS

    inv()
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(5):
        inv()
    ev1.record()
    torch.cuda.synchronize()
    ims = ev0.elapsed_time(ev1) / 5
    idx = torch.arange(0, batch, 511, device="cuda")
    am = m0[idx].transpose(1, 2).to(torch.complex128)
    cm = mi[idx].transpose(1, 2).to(torch.complex128)
    res = (torch.matmul(am, cm) - torch.eye(nb, device="cuda", dtype=torch.complex128)).abs().max().item()
    bytes_ = batch * (2 * nb * nb * 8 + 4)
    out["cmatinv_batched"] = {"workload": "4096 complex64 matrices of 64x64, same distribution (inverse + info)",
                              "ms": round(ims, 4), "matrices_per_s": round(batch / ims * 1e3, 1),
                              "GB_per_s": round(bytes_ / ims / 1e6, 1),
                              "hbm_frac": round(bytes_ / ims / 1e-3 / PEAK_HBM, 4),
                              "GFLOP_per_s": round(8.0 * nb ** 3 * batch / ims / 1e6, 1),
                              "max_abs_AC_minus_I_sample": res, "all_info_zero": bool((binfo == 0).all().item())}
    return out


def uneven_batched(A, torch, h, steps=5):
    """20000 FP32 systems, n_i ~ U{8..256} (seeded), A_i = N(0,1)/sqrt(n_i) + 2I, 4 RHS:
    getrf + getrs of every system through one plan execution: four size-class launches
    (plan built and autotuned once, untimed). Inputs are restored between steps outside the
    timed events."""
    import numpy as np

    count, nrhs = 20000, 4
    rng = np.random.default_rng(20000)
    sizes = rng.integers(8, 257, size=count).astype(np.int32)
    g = torch.Generator(device="cuda")
    g.manual_seed(256)
    offA = np.concatenate([[0], np.cumsum(sizes.astype(np.int64) ** 2)])
    offB = np.concatenate([[0], np.cumsum(sizes.astype(np.int64) * nrhs)])
    a0 = torch.randn(int(offA[-1]), device="cuda", generator=g)
    b0 = torch.randn(int(offB[-1]), device="cuda", generator=g)
    scale = torch.from_numpy(np.repeat(1.0 / np.sqrt(sizes), sizes.astype(np.int64) ** 2).astype(np.float32)).cuda()
    a0 *= scale
    diag = np.concatenate([offA[i] + np.arange(n) * (n + 1) for i, n in enumerate(sizes)])
    a0[torch.from_numpy(diag).cuda()] += 2.0
    a = a0.clone()
    b = b0.clone()
    pa = (a.data_ptr() + torch.from_numpy(offA[:-1] * 4).cuda()).to(torch.int64)
    pb = (b.data_ptr() + torch.from_numpy(offB[:-1] * 4).cuda()).to(torch.int64)
    ipiv = torch.empty(int(sizes.sum()), dtype=torch.int32, device="cuda")
    info = torch.empty(count, dtype=torch.int32, device="cuda")
    L = A.lib()
    plan = ctypes.c_void_p(None)
    assert L.aclsolverSgesvVbatched_createPlan(h.h, count, sizes.ctypes.data, nrhs, ctypes.byref(plan)) == A.SUCCESS
    ms_total = 0.0
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for step in range(steps + 1):
        a.copy_(a0)
        b.copy_(b0)
        torch.cuda.synchronize()
        ev0.record()
        assert L.aclsolverSgesvVbatched_execute(h.h, plan, A.ptr(pa), A.ptr(pb), A.ptr(ipiv), A.ptr(info)) == A.SUCCESS
        ev1.record()
        torch.cuda.synchronize()
        if step:
            ms_total += ev0.elapsed_time(ev1)
    ms = ms_total / steps
    sched = np.zeros(12, np.int32)
    assert L.aclsolverSgesvVbatched_getSchedule(plan, sched.ctypes.data) == A.SUCCESS
    L.aclsolverSgesvVbatched_destroyPlan(plan)
    cfg_names = {1: "C2x4", 2: "C2x8", 3: "C3x8", 4: "C3x16", 5: "C4x16"}
    n = sizes.astype(np.float64)
    bytes_ = float(np.sum((2 * n * n + 2 * n * nrhs) * 4 + n * 4 + 4))
    flops = float(np.sum(2.0 / 3.0 * n ** 3 + 2 * n * n * nrhs))
    # spot-check a few solutions: A x = b
    worst = 0.0
    for i in (0, 1, 777, 19999):
        ni = int(sizes[i])
        Ai = a0[offA[i]:offA[i + 1]].view(ni, ni).T.double()
        Bi = b0[offB[i]:offB[i + 1]].view(nrhs, ni).T.double()
        Xi = b[offB[i]:offB[i + 1]].view(nrhs, ni).T.double()
        r = (Bi - Ai @ Xi).abs().max() / (Ai.abs().sum(1).max() * Xi.abs().max() + Bi.abs().max())
        worst = max(worst, r.item())
    return {"workload": "20000 FP32 systems, n_i ~ U{8..256}, A_i = N(0,1)/sqrt(n_i) + 2I, 4 RHS; getrf+getrs "
                        "through one plan, one persistent launch: n<=64 one wave per system, 65..256 column-cyclic "
                        "wave groups (configuration per 16-wide size bin from the autotuner)",
            "schedule": {f"{65 + 16 * b}..{80 + 16 * b}": cfg_names.get(int(c), str(int(c))) for b, c in enumerate(sched)},
            "ms": round(ms, 4), "solves_per_s": round(count / ms * 1e3, 1),
            "GB_per_s": round(bytes_ / ms / 1e6, 1), "hbm_frac": round(bytes_ / (ms * 1e-3) / PEAK_HBM, 4),
            "GFLOP_per_s": round(flops / ms / 1e6, 1), "max_backward_error_sample": worst,
            "all_info_zero": bool((info == 0).all().item())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--cpu-n", type=int, default=16384)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--modes", default="SS,SH,SB",
                    help="gesv variants to time: SS, SH, SB (SB, the headline, always runs)")
    ap.add_argument("--restart", type=int, default=60)
    ap.add_argument("--tournament", type=int, default=0, help="SH/SB tall panels: 1 tournament pivoting, 0 partial")
    ap.add_argument("--matrix", choices=["qr", "shifted"], default="qr",
                    help="qr: Q1 diag(s) Q2^T with cond 1e6 (the workload); shifted: N(0,1)/sqrt(n) + 2I")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="counter passes: first run one plain 1 GiB device copy (the reference dispatch for the "
                         "HBM counters) and record no timing events")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU (block-cyclic / batched sharding) "
                                   "path is a design only, --gpus must be 1"}))
        return 2
    import torch

    import aclsolver as A

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n = args.n
    if args.copy_calibration:
        src = torch.ones(256 * 1024 * 1024, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)  # 1 GiB read + 1 GiB write
        torch.cuda.synchronize()
        del src, dst
        log("calibration copy done (1 GiB)")
    log(f"generating A ({args.matrix}), n={n}")
    at, b, x_true = make_system(torch, n, seed=16384, kind=args.matrix)
    torch.cuda.synchronize()
    h = A.Handle()
    A.lib().aclsolverSetRefinement(h.h, ctypes.c_double(1e-6), 20, args.restart)
    A.lib().aclsolverSetTournamentPivoting(h.h, args.tournament)
    lw = ctypes.c_size_t(0)
    assert A.lib().aclsolverSBgesv_bufferSize(h.h, n, 1, ctypes.byref(lw)) == A.SUCCESS
    work = torch.empty(lw.value, dtype=torch.uint8, device="cuda")
    x = torch.empty(n, device="cuda")
    dipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    dinfo = torch.empty(1, dtype=torch.int32, device="cuda")

    res = {}
    modes = [m for m in args.modes.split(",") if m]
    if "SB" not in modes:
        modes.append("SB")
    for name in modes:
        log(f"timing {name}gesv: warmup {args.warmup}, steps {args.steps}")
        ms, iters, info, prof = time_mode(A, torch, h, name, args.steps, args.warmup, n, at, b, x, work, lw.value,
                                          dipiv, dinfo, profile=(name == "SB" and not args.copy_calibration))
        berr, ferr = errors(torch, at, x, b, x_true)
        res[name] = {"ms": ms, "gmres_iterations": iters, "info": info, "backward_error": berr,
                     "forward_error": ferr, "profile": prof}
        if prof is not None:
            for key, (cnt, wk, tms) in prof.items():
                log(f"{name} {key}: {cnt} launches, {tms:.2f} ms summed")
        log(f"{name}: {ms:.2f} ms, iters {iters}, info {info}, berr {berr:.3e}, ferr {ferr:.3e}")

    flops = 2.0 / 3.0 * n ** 3 + 2.0 * n ** 2
    sb = res["SB"]
    value = flops / (sb["ms"] * 1e-3) / 1e9
    prof = sb["profile"] or {}
    launches, gflop, gms = prof.get("gemm", (0, 0.0, 0.0))
    achieved = gflop / (gms * 1e-3) if gms > 0 else 0.0
    leaves, leaf_bytes, leaf_ms = prof.get("leaf", (0, 0.0, 0.0))
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sb["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"dense solve n={n}, " + (
                       "A=Q1 diag(s) Q2^T (s log-spaced 1..1e6, cond 1e6)" if args.matrix == "qr" else
                       "A=N(0,1)/sqrt(n)+2I") + ", 1 RHS; BF16 MFMA LU (FP32 accumulate) + GMRES-IR with FP64 "
                       "residual to bwd err 1e-6",
                   "n": n, "nrhs": 1, "parallelism": "1 GPU"},
        "roofline": {"bound": "mfma",
                     "kernel": "gemm_packed_kernel<bf16,3>: trailing update, 3-part bf16 operand split "
                               "(6 MFMA products per algorithmic FMA), fp32 accumulate",
                     "achieved": round(achieved / 1e12, 2), "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_BF16, 4), "traffic": measured_traffic(n),
                     "launches": launches, "avg_launch_ms": round(gms / max(launches, 1), 4),
                     "avg_launch_gflop": round(gflop / max(launches, 1) / 1e9, 3)},
        "panel_leaf": {"bound": "latency (sequential columns); HBM shown for scale",
                       "kernel": ("tournament-pivoting panel leaves (tree + apply launches, 32 columns each)"
                                  if args.tournament else
                                  "partial-pivoting panel leaves (leaf_multi_kernel / leaf_getf2_kernel, one "
                                  "launch per leaf of the recursive panel)"),
                       "launches": leaves, "ms_total": round(leaf_ms, 3),
                       "avg_launch_ms": round(leaf_ms / max(leaves, 1), 4),
                       "achieved": round(leaf_bytes / max(leaf_ms * 1e-3, 1e-12) / 1e9, 1), "peak": PEAK_HBM / 1e9,
                       "unit": "GB/s",
                       "frac": round(leaf_bytes / max(leaf_ms * 1e-3, 1e-12) / PEAK_HBM, 5)},
        "backward_error": sb["backward_error"],
        "forward_error_vs_fp64_truth": sb["forward_error"],
        "gmres_iterations": sb["gmres_iterations"],
        "gmres_iterations_by_mode": {k: v["gmres_iterations"] for k, v in res.items()},
    }
    if "SS" in res:
        ss = res["SS"]
        out.update({"fp32_ms_per_step": round(ss["ms"], 3), "fp32_backward_error": ss["backward_error"],
                    "fp32_GFLOP_per_s": round(flops / (ss["ms"] * 1e-3) / 1e9, 1),
                    "speedup_vs_fp32": round(ss["ms"] / sb["ms"], 3)})
    if "SH" in res:
        sh = res["SH"]
        out["fp16_ir"] = {"ms": round(sh["ms"], 3), "GFLOP_per_s": round(flops / (sh["ms"] * 1e-3) / 1e9, 1),
                          "backward_error": sh["backward_error"], "forward_error": sh["forward_error"],
                          "gmres_iterations": sh["gmres_iterations"]}
        if "SS" in res:
            out["fp16_ir"]["speedup_vs_fp32"] = round(res["SS"]["ms"] / sh["ms"], 3)
    del at
    torch.cuda.empty_cache()
    if not args.skip_batched:
        log("plain sgetrf+sgetrs n=512")
        out["small_dense"] = small_dense(A, torch, h)
        log("batched inverse 65536 x 32x32")
        out["batched_inverse"] = batched_inverse(A, torch, h)
        log("uneven batched getrf+getrs, 20000 systems")
        out["uneven_batched"] = uneven_batched(A, torch, h)
        log("complex cgetrf+cgetrs n=8192, cmatinv 4096 x 64x64")
        out["complex"] = complex_solve(A, torch, h)
    if not args.skip_cpu:
        log(f"host LAPACK baseline n={args.cpu_n}")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
        if not args.skip_batched:
            import numpy as np

            log("host LAPACK baselines of the other workloads")
            out["small_dense"]["cpu_baseline"] = cpu_baseline(512, reps=51)
            out["batched_inverse"]["cpu_baseline"] = cpu_batched_inverse_baseline(sample=65536)
            sizes = np.random.default_rng(20000).integers(8, 257, size=20000)
            out["uneven_batched"]["cpu_baseline"] = cpu_uneven_baseline(sizes, sample=20000)
            out["complex"]["cpu_baseline"] = cpu_complex_baseline()
    h.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
