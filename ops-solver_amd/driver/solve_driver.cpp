// Command-line driver over the C ABI (no Python): builds a seeded system on the host, then
// times one of the dense LU paths on the GPU. Used for profiler runs, where a lean process
// keeps the counter passes short.
//
//   solve_driver <mode> <n> [steps] [--copy-calibration]
//     mode SB | SH | SS : aclsolver{SB,SH,SS}gesv, 1 RHS, refinement to backward error 1e-6
//     mode SGETRF       : aclsolverSgetrf + aclsolverSgetrsDevice, 1 RHS
//     mode CGETRF       : aclsolverCgetrf + aclsolverCgetrsDevice (complex64), 16 RHS
//     mode VBATCHED     : aclsolverSgesvVbatched plan + execute: n systems of orders
//                         U{8..256} (seed n), A_i = N(0,1)/sqrt(n_i) + 2I, 4 RHS
//     mode MATINV       : n matrices of 32 x 32, A_i = N(0,1)/sqrt(32) + 2I: the strided
//                         batched inverse (ipiv + info), then the drop-in pointer-array
//                         aclsolverSmatinvBatched
//   Real modes: A = N(0,1)/sqrt(n) + 2I (column-major, splitmix64 + Box-Muller, seed n),
//   b ~ N(0,1). CGETRF: re/im of A i.i.d. N(0,1/2)/sqrt(n) plus 1.5 on the diagonal, re/im
//   of B i.i.d. N(0,1/2).
//   --copy-calibration first copies 1 GiB device to device (one plain copy dispatch of known
//   size, to scale the FETCH_SIZE / WRITE_SIZE counters).
//
// Prints one line per step (milliseconds by HIP events on the handle's stream) and the final
// backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) computed on the host in
// double precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cann_ops_solver.h"

namespace
{
uint64_t Next(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double Normal(uint64_t* s)
{
    const double u1 = (static_cast<double>(Next(s) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    const double u2 = static_cast<double>(Next(s) >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

#define CHECK_HIP(x)                                                                         \
    do                                                                                       \
    {                                                                                        \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
        {                                                                                    \
            std::fprintf(stderr, "%s:%d: %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            std::exit(2);                                                                    \
        }                                                                                    \
    } while (0)

#define CHECK_ST(x)                                                            \
    do                                                                         \
    {                                                                          \
        aclsolverStatus_t s_ = (x);                                            \
        if (s_ != ACLSOLVER_STATUS_SUCCESS)                                    \
        {                                                                      \
            std::fprintf(stderr, "%s:%d: status %d\n", __FILE__, __LINE__, s_); \
            std::exit(3);                                                      \
        }                                                                      \
    } while (0)
// cgetrf + cgetrs of one complex64 system with 16 right-hand sides; prints per-step times and
// the backward error max|B - A X| / (||A||_inf max|X| + max|B|) of the first two right-hand
// sides in double precision.
int RunComplex(int n, int steps)
{
    constexpr int nrhs = 16;
    const size_t nn = static_cast<size_t>(n) * n;
    std::vector<aclFloatComplex> a(nn), b(static_cast<size_t>(n) * nrhs);
    uint64_t seed = static_cast<uint64_t>(n) + 5;
    const double scale = std::sqrt(0.5 / n);
    for (size_t i = 0; i < nn; ++i)
    {
        a[i].real = static_cast<float>(Normal(&seed) * scale);
        a[i].imag = static_cast<float>(Normal(&seed) * scale);
    }
    for (int i = 0; i < n; ++i) a[static_cast<size_t>(i) * n + i].real += 1.5f;
    for (auto& v : b)
    {
        v.real = static_cast<float>(Normal(&seed) * std::sqrt(0.5));
        v.imag = static_cast<float>(Normal(&seed) * std::sqrt(0.5));
    }
    aclsolverHandle_t h = nullptr;
    CHECK_ST(aclsolverCreate(&h));
    hipStream_t stream = nullptr;
    CHECK_HIP(hipStreamCreate(&stream));
    CHECK_ST(aclsolverSetStream(h, stream));
    aclFloatComplex *dA = nullptr, *dLU = nullptr, *dB = nullptr, *dX = nullptr, *work = nullptr;
    int *dIpiv = nullptr, *dInfo = nullptr;
    const size_t bb = b.size() * sizeof(aclFloatComplex);
    CHECK_HIP(hipMalloc(&dA, nn * sizeof(aclFloatComplex)));
    CHECK_HIP(hipMalloc(&dLU, nn * sizeof(aclFloatComplex)));
    CHECK_HIP(hipMalloc(&dB, bb));
    CHECK_HIP(hipMalloc(&dX, bb));
    CHECK_HIP(hipMalloc(&dIpiv, n * sizeof(int)));
    CHECK_HIP(hipMalloc(&dInfo, sizeof(int)));
    CHECK_HIP(hipMemcpy(dA, a.data(), nn * sizeof(aclFloatComplex), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(dB, b.data(), bb, hipMemcpyHostToDevice));
    int lw = 0;
    CHECK_ST(aclsolverCgetrf_bufferSize(h, n, n, dA, n, &lw));
    CHECK_HIP(hipMalloc(&work, std::max(lw, 1) * sizeof(aclFloatComplex)));
    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    for (int s = 0; s < steps; ++s)
    {
        CHECK_HIP(hipMemcpyAsync(dLU, dA, nn * sizeof(aclFloatComplex), hipMemcpyDeviceToDevice, stream));
        CHECK_HIP(hipMemcpyAsync(dX, dB, bb, hipMemcpyDeviceToDevice, stream));
        CHECK_HIP(hipEventRecord(e0, stream));
        CHECK_ST(aclsolverCgetrf(h, n, n, dLU, n, work, dIpiv, dInfo));
        CHECK_ST(aclsolverCgetrsDevice(h, 0, n, nrhs, dLU, n, dIpiv, dX, n, dInfo));
        CHECK_HIP(hipEventRecord(e1, stream));
        CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        std::printf("CGETRF n=%d step %d: %.3f ms\n", n, s + 1, ms);
        std::fflush(stdout);
    }
    int info = 0;
    std::vector<aclFloatComplex> x(b.size());
    CHECK_HIP(hipMemcpy(&info, dInfo, sizeof(int), hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(x.data(), dX, bb, hipMemcpyDeviceToHost));
    double rmax = 0.0, xmax = 0.0, bmax = 0.0, anorm = 0.0;
    std::vector<double> rowsum(n, 0.0);
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < n; ++i)
        {
            const aclFloatComplex v = a[static_cast<size_t>(c) * n + i];
            rowsum[i] += std::hypot(static_cast<double>(v.real), static_cast<double>(v.imag));
        }
    for (int i = 0; i < n; ++i) anorm = std::fmax(anorm, rowsum[i]);
    for (int q = 0; q < 2; ++q)  // the first two right-hand sides (host time)
    {
        std::vector<double> rr(n), ri(n);
        for (int i = 0; i < n; ++i)
        {
            rr[i] = b[static_cast<size_t>(q) * n + i].real;
            ri[i] = b[static_cast<size_t>(q) * n + i].imag;
        }
        for (int c = 0; c < n; ++c)
        {
            const double xr = x[static_cast<size_t>(q) * n + c].real, xi = x[static_cast<size_t>(q) * n + c].imag;
            xmax = std::fmax(xmax, std::hypot(xr, xi));
            for (int i = 0; i < n; ++i)
            {
                const aclFloatComplex v = a[static_cast<size_t>(c) * n + i];
                rr[i] -= v.real * xr - v.imag * xi;
                ri[i] -= v.real * xi + v.imag * xr;
            }
        }
        for (int i = 0; i < n; ++i)
        {
            rmax = std::fmax(rmax, std::hypot(rr[i], ri[i]));
            bmax = std::fmax(bmax, std::hypot(static_cast<double>(b[static_cast<size_t>(q) * n + i].real),
                                              static_cast<double>(b[static_cast<size_t>(q) * n + i].imag)));
        }
    }
    std::printf("CGETRF n=%d info %d backward error %.3e\n", n, info, rmax / (anorm * xmax + bmax));
    std::fflush(stdout);
    CHECK_HIP(hipEventDestroy(e0));
    CHECK_HIP(hipEventDestroy(e1));
    CHECK_HIP(hipFree(work));
    CHECK_HIP(hipFree(dA));
    CHECK_HIP(hipFree(dLU));
    CHECK_HIP(hipFree(dB));
    CHECK_HIP(hipFree(dX));
    CHECK_HIP(hipFree(dIpiv));
    CHECK_HIP(hipFree(dInfo));
    CHECK_ST(aclsolverDestroy(h));
    CHECK_HIP(hipStreamDestroy(stream));
    return info == 0 ? 0 : 4;
}
// Times `steps` executions of one uneven-batch plan over `count` systems (A and B restored
// before each, outside the events); prints the per-step times and the plan's schedule.
int RunVbatched(int count, int steps)
{
    constexpr int nrhs = 4;
    std::vector<int> ns(count);
    uint64_t seed = static_cast<uint64_t>(count);
    for (int i = 0; i < count; ++i) ns[i] = 8 + static_cast<int>(Next(&seed) % 249);
    std::vector<size_t> offA(count + 1, 0), offB(count + 1, 0);
    for (int i = 0; i < count; ++i)
    {
        offA[i + 1] = offA[i] + static_cast<size_t>(ns[i]) * ns[i];
        offB[i + 1] = offB[i] + static_cast<size_t>(ns[i]) * nrhs;
    }
    std::vector<float> a(offA[count]), b(offB[count]);
    for (int i = 0; i < count; ++i)
    {
        const double sc = 1.0 / std::sqrt(static_cast<double>(ns[i]));
        for (size_t e = offA[i]; e < offA[i + 1]; ++e) a[e] = static_cast<float>(Normal(&seed) * sc);
        for (int d = 0; d < ns[i]; ++d) a[offA[i] + static_cast<size_t>(d) * ns[i] + d] += 2.0f;
    }
    for (auto& v : b) v = static_cast<float>(Normal(&seed));
    aclsolverHandle_t h = nullptr;
    CHECK_ST(aclsolverCreate(&h));
    hipStream_t stream = nullptr;
    CHECK_HIP(hipStreamCreate(&stream));
    CHECK_ST(aclsolverSetStream(h, stream));
    float *dA0 = nullptr, *dB0 = nullptr, *dA = nullptr, *dB = nullptr;
    float **dAp = nullptr, **dBp = nullptr;
    int *dIpiv = nullptr, *dInfo = nullptr;
    CHECK_HIP(hipMalloc(&dA0, a.size() * sizeof(float)));
    CHECK_HIP(hipMalloc(&dB0, b.size() * sizeof(float)));
    CHECK_HIP(hipMalloc(&dA, a.size() * sizeof(float)));
    CHECK_HIP(hipMalloc(&dB, b.size() * sizeof(float)));
    CHECK_HIP(hipMalloc(&dAp, count * sizeof(float*)));
    CHECK_HIP(hipMalloc(&dBp, count * sizeof(float*)));
    CHECK_HIP(hipMalloc(&dIpiv, offA.size() * sizeof(int) + static_cast<size_t>(count) * 256 * sizeof(int)));
    CHECK_HIP(hipMalloc(&dInfo, count * sizeof(int)));
    CHECK_HIP(hipMemcpy(dA0, a.data(), a.size() * sizeof(float), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(dB0, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float*> pa(count), pb(count);
    for (int i = 0; i < count; ++i)
    {
        pa[i] = dA + offA[i];
        pb[i] = dB + offB[i];
    }
    CHECK_HIP(hipMemcpy(dAp, pa.data(), count * sizeof(float*), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(dBp, pb.data(), count * sizeof(float*), hipMemcpyHostToDevice));
    aclsolverVbatchedPlan_t plan = nullptr;
    CHECK_ST(aclsolverSgesvVbatched_createPlan(h, count, ns.data(), nrhs, &plan));
    int sched[12] = {};
    CHECK_ST(aclsolverSgesvVbatched_getSchedule(plan, sched));
    std::printf("VBATCHED schedule (bins 65..256 by 16):");
    for (int v : sched) std::printf(" %d", v);
    std::printf("\n");
    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    for (int s = 0; s < steps; ++s)
    {
        CHECK_HIP(hipMemcpyAsync(dA, dA0, a.size() * sizeof(float), hipMemcpyDeviceToDevice, stream));
        CHECK_HIP(hipMemcpyAsync(dB, dB0, b.size() * sizeof(float), hipMemcpyDeviceToDevice, stream));
        CHECK_HIP(hipEventRecord(e0, stream));
        CHECK_ST(aclsolverSgesvVbatched_execute(h, plan, dAp, dBp, dIpiv, dInfo));
        CHECK_HIP(hipEventRecord(e1, stream));
        CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        std::printf("VBATCHED count=%d step %d: %.3f ms\n", count, s + 1, ms);
        std::fflush(stdout);
    }
    std::vector<int> info(count);
    CHECK_HIP(hipMemcpy(info.data(), dInfo, count * sizeof(int), hipMemcpyDeviceToHost));
    int bad = 0;
    for (int v : info) bad += v != 0;
    std::printf("VBATCHED count=%d nonzero info %d\n", count, bad);
    std::fflush(stdout);
    CHECK_ST(aclsolverSgesvVbatched_destroyPlan(plan));
    CHECK_HIP(hipEventDestroy(e0));
    CHECK_HIP(hipEventDestroy(e1));
    for (void* p : {static_cast<void*>(dA0), static_cast<void*>(dB0), static_cast<void*>(dA), static_cast<void*>(dB),
                    static_cast<void*>(dAp), static_cast<void*>(dBp), static_cast<void*>(dIpiv),
                    static_cast<void*>(dInfo)})
        CHECK_HIP(hipFree(p));
    CHECK_ST(aclsolverDestroy(h));
    CHECK_HIP(hipStreamDestroy(stream));
    return bad == 0 ? 0 : 4;
}

// Times the strided batched inverse, then the drop-in pointer-array entry point (host wall
// clock per call, validation included), on `count` 32 x 32 matrices.
int RunMatinv(int count, int steps)
{
    constexpr int n = 32;
    const size_t per = static_cast<size_t>(n) * n, total = per * count;
    std::vector<float> a(total);
    uint64_t seed = static_cast<uint64_t>(count) + 32;
    const double sc = 1.0 / std::sqrt(static_cast<double>(n));
    for (size_t e = 0; e < total; ++e) a[e] = static_cast<float>(Normal(&seed) * sc);
    for (int i = 0; i < count; ++i)
        for (int d = 0; d < n; ++d) a[i * per + static_cast<size_t>(d) * n + d] += 2.0f;
    aclsolverHandle_t h = nullptr;
    CHECK_ST(aclsolverCreate(&h));
    hipStream_t stream = nullptr;
    CHECK_HIP(hipStreamCreate(&stream));
    CHECK_ST(aclsolverSetStream(h, stream));
    float *dA = nullptr, *dC = nullptr;
    float **dAp = nullptr, **dCp = nullptr;
    int *dIpiv = nullptr, *dInfo = nullptr;
    CHECK_HIP(hipMalloc(&dA, total * sizeof(float)));
    CHECK_HIP(hipMalloc(&dC, total * sizeof(float)));
    CHECK_HIP(hipMalloc(&dAp, count * sizeof(float*)));
    CHECK_HIP(hipMalloc(&dCp, count * sizeof(float*)));
    CHECK_HIP(hipMalloc(&dIpiv, static_cast<size_t>(count) * n * sizeof(int)));
    CHECK_HIP(hipMalloc(&dInfo, count * sizeof(int)));
    CHECK_HIP(hipMemcpy(dA, a.data(), total * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float*> pa(count), pc(count);
    for (int i = 0; i < count; ++i)
    {
        pa[i] = dA + i * per;
        pc[i] = dC + i * per;
    }
    CHECK_HIP(hipMemcpy(dAp, pa.data(), count * sizeof(float*), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(dCp, pc.data(), count * sizeof(float*), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    for (int s = 0; s < steps; ++s)
    {
        CHECK_HIP(hipEventRecord(e0, stream));
        CHECK_ST(aclsolverSmatinvStridedBatched(h, n, dA, n, static_cast<long long>(per), dC, n,
                                                static_cast<long long>(per), dIpiv, dInfo, count));
        CHECK_HIP(hipEventRecord(e1, stream));
        CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        std::printf("MATINV strided count=%d step %d: %.4f ms\n", count, s + 1, ms);
        std::fflush(stdout);
    }
    for (int s = 0; s < steps; ++s)
    {
        CHECK_HIP(hipStreamSynchronize(stream));
        const auto t0 = std::chrono::steady_clock::now();
        CHECK_ST(aclsolverSmatinvBatched(h, n, dAp, n, dCp, n, dInfo, count));
        CHECK_HIP(hipStreamSynchronize(stream));
        const auto t1 = std::chrono::steady_clock::now();
        std::printf("MATINV pointer-array count=%d step %d: %.4f ms wall\n", count, s + 1,
                    std::chrono::duration<double, std::milli>(t1 - t0).count());
        std::fflush(stdout);
    }
    std::vector<int> info(count);
    CHECK_HIP(hipMemcpy(info.data(), dInfo, count * sizeof(int), hipMemcpyDeviceToHost));
    int bad = 0;
    for (int v : info) bad += v != 0;
    std::printf("MATINV count=%d nonzero info %d\n", count, bad);
    std::fflush(stdout);
    CHECK_HIP(hipEventDestroy(e0));
    CHECK_HIP(hipEventDestroy(e1));
    for (void* p : {static_cast<void*>(dA), static_cast<void*>(dC), static_cast<void*>(dAp), static_cast<void*>(dCp),
                    static_cast<void*>(dIpiv), static_cast<void*>(dInfo)})
        CHECK_HIP(hipFree(p));
    CHECK_ST(aclsolverDestroy(h));
    CHECK_HIP(hipStreamDestroy(stream));
    return bad == 0 ? 0 : 4;
}
}  // namespace

int main(int argc, char** argv)
{
    if (argc < 3)
    {
        std::fprintf(stderr, "usage: %s SB|SH|SS|SGETRF|CGETRF|VBATCHED|MATINV n [steps] [--copy-calibration]\n",
                     argv[0]);
        return 1;
    }
    const std::string mode = argv[1];
    const int n = std::atoi(argv[2]);
    int steps = 1;
    bool calib = false;
    for (int i = 3; i < argc; ++i)
    {
        if (std::strcmp(argv[i], "--copy-calibration") == 0)
            calib = true;
        else
            steps = std::atoi(argv[i]);
    }
    if (n <= 0 || steps <= 0 ||
        (mode != "SB" && mode != "SH" && mode != "SS" && mode != "SGETRF" && mode != "CGETRF" && mode != "VBATCHED" &&
         mode != "MATINV"))
    {
        std::fprintf(stderr, "bad arguments\n");
        return 1;
    }

    if (calib)
    {
        const size_t bytes = size_t(1) << 30;
        void *src = nullptr, *dst = nullptr;
        CHECK_HIP(hipMalloc(&src, bytes));
        CHECK_HIP(hipMalloc(&dst, bytes));
        CHECK_HIP(hipMemset(src, 0, bytes));
        CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
        CHECK_HIP(hipDeviceSynchronize());
        CHECK_HIP(hipFree(src));
        CHECK_HIP(hipFree(dst));
        std::printf("calibration copy done (1 GiB)\n");
    }

    if (mode == "CGETRF") return RunComplex(n, steps);
    if (mode == "VBATCHED") return RunVbatched(n, steps);
    if (mode == "MATINV") return RunMatinv(n, steps);

    const size_t nn = static_cast<size_t>(n) * n;
    std::vector<float> a(nn), b(n);
    uint64_t seed = static_cast<uint64_t>(n);
    const double scale = 1.0 / std::sqrt(static_cast<double>(n));
    for (size_t i = 0; i < nn; ++i) a[i] = static_cast<float>(Normal(&seed) * scale);
    for (int i = 0; i < n; ++i) a[static_cast<size_t>(i) * n + i] += 2.0f;
    for (int i = 0; i < n; ++i) b[i] = static_cast<float>(Normal(&seed));

    aclsolverHandle_t h = nullptr;
    CHECK_ST(aclsolverCreate(&h));
    hipStream_t stream = nullptr;
    CHECK_HIP(hipStreamCreate(&stream));
    CHECK_ST(aclsolverSetStream(h, stream));

    float *dA = nullptr, *dLU = nullptr, *dB = nullptr, *dX = nullptr;
    int *dIpiv = nullptr, *dInfo = nullptr;
    CHECK_HIP(hipMalloc(&dA, nn * sizeof(float)));
    CHECK_HIP(hipMalloc(&dB, n * sizeof(float)));
    CHECK_HIP(hipMalloc(&dX, n * sizeof(float)));
    CHECK_HIP(hipMalloc(&dIpiv, n * sizeof(int)));
    CHECK_HIP(hipMalloc(&dInfo, sizeof(int)));
    CHECK_HIP(hipMemcpy(dA, a.data(), nn * sizeof(float), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(dB, b.data(), n * sizeof(float), hipMemcpyHostToDevice));

    void* work = nullptr;
    size_t lwork = 0;
    const bool gesv = mode != "SGETRF";
    if (gesv)
    {
        CHECK_ST(aclsolverSetRefinement(h, 1e-6, 20, 30));
        if (mode == "SB") CHECK_ST(aclsolverSBgesv_bufferSize(h, n, 1, &lwork));
        if (mode == "SH") CHECK_ST(aclsolverSHgesv_bufferSize(h, n, 1, &lwork));
        if (mode == "SS") CHECK_ST(aclsolverSSgesv_bufferSize(h, n, 1, &lwork));
    }
    else
    {
        int lw = 0;
        CHECK_ST(aclsolverSgetrf_bufferSize(h, n, n, dA, n, &lw));
        lwork = static_cast<size_t>(lw) * sizeof(float);
        CHECK_HIP(hipMalloc(&dLU, nn * sizeof(float)));
    }
    CHECK_HIP(hipMalloc(&work, lwork > 0 ? lwork : 4));

    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    int niter = 0;
    for (int s = 0; s < steps; ++s)
    {
        if (!gesv)
        {
            CHECK_HIP(hipMemcpyAsync(dLU, dA, nn * sizeof(float), hipMemcpyDeviceToDevice, stream));
            CHECK_HIP(hipMemcpyAsync(dX, dB, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
        }
        CHECK_HIP(hipEventRecord(e0, stream));
        if (mode == "SB")
            CHECK_ST(aclsolverSBgesv(h, n, 1, dA, n, dIpiv, dB, n, dX, n, work, lwork, &niter, dInfo));
        else if (mode == "SH")
            CHECK_ST(aclsolverSHgesv(h, n, 1, dA, n, dIpiv, dB, n, dX, n, work, lwork, &niter, dInfo));
        else if (mode == "SS")
            CHECK_ST(aclsolverSSgesv(h, n, 1, dA, n, dIpiv, dB, n, dX, n, work, lwork, &niter, dInfo));
        else
        {
            CHECK_ST(aclsolverSgetrf(h, n, n, dLU, n, static_cast<float*>(work), dIpiv, dInfo));
            CHECK_ST(aclsolverSgetrsDevice(h, 0, n, 1, dLU, n, dIpiv, dX, n, dInfo));
        }
        CHECK_HIP(hipEventRecord(e1, stream));
        CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        std::printf("%s n=%d step %d: %.3f ms, %d GMRES iterations\n", mode.c_str(), n, s + 1, ms, niter);
        std::fflush(stdout);
    }

    int info = 0;
    std::vector<float> x(n);
    CHECK_HIP(hipMemcpy(&info, dInfo, sizeof(int), hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(x.data(), dX, n * sizeof(float), hipMemcpyDeviceToHost));
    double rmax = 0.0, xmax = 0.0, bmax = 0.0, anorm = 0.0;
    std::vector<double> r(b.begin(), b.end()), rowsum(n, 0.0);
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < n; ++i)
        {
            const double v = a[static_cast<size_t>(c) * n + i];
            r[i] -= v * x[c];
            rowsum[i] += std::fabs(v);
        }
    for (int i = 0; i < n; ++i)
    {
        rmax = std::fmax(rmax, std::fabs(r[i]));
        xmax = std::fmax(xmax, std::fabs(static_cast<double>(x[i])));
        bmax = std::fmax(bmax, std::fabs(static_cast<double>(b[i])));
        anorm = std::fmax(anorm, rowsum[i]);
    }
    std::printf("%s n=%d info %d backward error %.3e\n", mode.c_str(), n, info, rmax / (anorm * xmax + bmax));
    std::fflush(stdout);

    CHECK_HIP(hipEventDestroy(e0));
    CHECK_HIP(hipEventDestroy(e1));
    CHECK_HIP(hipFree(work));
    if (dLU) CHECK_HIP(hipFree(dLU));
    CHECK_HIP(hipFree(dA));
    CHECK_HIP(hipFree(dB));
    CHECK_HIP(hipFree(dX));
    CHECK_HIP(hipFree(dIpiv));
    CHECK_HIP(hipFree(dInfo));
    CHECK_ST(aclsolverDestroy(h));
    CHECK_HIP(hipStreamDestroy(stream));
    return info == 0 ? 0 : 4;
}
