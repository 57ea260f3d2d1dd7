// Command-line driver over the C ABI (no Python): builds a seeded system on the host, then
// times one of the dense LU paths on the GPU. Used for profiler runs, where a lean process
// keeps the counter passes short.
//
//   solve_driver <mode> <n> [steps] [--copy-calibration]
//     mode SB | SH | SS : aclsolver{SB,SH,SS}gesv, 1 RHS, refinement to backward error 1e-6
//     mode SGETRF       : aclsolverSgetrf + aclsolverSgetrsDevice, 1 RHS
//   A = N(0,1)/sqrt(n) + 2I (column-major, splitmix64 + Box-Muller, seed n), b ~ N(0,1).
//   --copy-calibration first copies 1 GiB device to device (one plain copy dispatch of known
//   size, to scale the FETCH_SIZE / WRITE_SIZE counters).
//
// Prints one line per step (milliseconds by HIP events on the handle's stream) and the final
// backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) computed on the host in
// double precision.
#include <hip/hip_runtime.h>

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
}  // namespace

int main(int argc, char** argv)
{
    if (argc < 3)
    {
        std::fprintf(stderr, "usage: %s SB|SH|SS|SGETRF n [steps] [--copy-calibration]\n", argv[0]);
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
    if (n <= 0 || steps <= 0 || (mode != "SB" && mode != "SH" && mode != "SS" && mode != "SGETRF"))
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
