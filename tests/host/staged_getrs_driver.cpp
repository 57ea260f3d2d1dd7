// CPU driver for the host row-major getrs staging (ops-solver_amd/csrc/host_getrs.h) on a
// stub runtime that fails one chosen step: the k-th allocation, copy, the info-word clear,
// the device solve, the bounded-wait report or the stream synchronization. It prints one
// RESULT line: status, allocations, frees, duplicate frees, frees while work was in flight,
// live buffers, info, and the solution.
//
// The "device" is host memory: the solve applies the interchanges and the two triangular
// substitutions to the staged row-major copies, so the success case returns X.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "host_getrs.h"

using ops::hoststage::Staging;

struct StubRuntime
{
    std::string failure;
    int index = 0;
    int allocations = 0, frees = 0, copies = 0, duplicateFrees = 0, unsafeFrees = 0;
    bool inFlight = false;
    std::set<void*> live;
    int n = 0, nrhs = 0;
    long lda = 0, ldb = 0;

    int Malloc(void** p, size_t bytes)
    {
        if (++allocations == index && failure == "alloc") return 1;
        *p = std::malloc(bytes);
        if (*p == nullptr) return 1;
        live.insert(*p);
        return 0;
    }
    int Free(void* p)
    {
        ++frees;
        if (inFlight) ++unsafeFrees;
        if (live.erase(p) == 0)
        {
            ++duplicateFrees;
            return 1;
        }
        std::free(p);
        return 0;
    }
    int Copy(void* d, const void* s, size_t b)
    {
        if (++copies == index && failure == "copy") return 1;
        std::memcpy(d, s, b);
        inFlight = true;
        return 0;
    }
    int CopyToDevice(void* d, const void* h, size_t b) { return Copy(d, h, b); }
    int CopyToHost(void* h, const void* d, size_t b) { return Copy(h, d, b); }
    int ClearWord(int* p)
    {
        if (failure == "clear") return 1;
        *p = 0;
        inFlight = true;
        return 0;
    }
    int Solve(const Staging& s)
    {
        if (failure == "launch") return 1;
        inFlight = true;
        const float* a = static_cast<const float*>(s.aRow);  // row-major LU, leading dimension lda
        float* b = static_cast<float*>(s.bRow);              // row-major B, leading dimension ldb
        for (int k = 0; k < n; ++k)
        {
            const int p = s.piv[k] - 1;
            if (p != k)
                for (int c = 0; c < nrhs; ++c) std::swap(b[k * ldb + c], b[p * ldb + c]);
        }
        for (int c = 0; c < nrhs; ++c)
        {
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < i; ++k) b[i * ldb + c] -= a[i * lda + k] * b[k * ldb + c];
            for (int i = n - 1; i >= 0; --i)
            {
                for (int k = i + 1; k < n; ++k) b[i * ldb + c] -= a[i * lda + k] * b[k * ldb + c];
                b[i * ldb + c] /= a[i * lda + i];
            }
        }
        if (failure == "wait") *s.wait = 1;
        return 0;
    }
    int Synchronize()
    {
        if (failure == "sync") return 1;
        inFlight = false;
        return 0;
    }
    size_t WorkBytes(int, int) { return 64; }
};

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    StubRuntime rt;
    rt.failure = argv[1];
    rt.index = std::atoi(argv[2]);
    // Factors of a 3 x 3 system (row-major, lda 4): P A = L U with ipiv {2, 3, 3};
    // L = [1 0 0; 0.5 1 0; 0.25 0.5 1], U = [4 2 1; 0 3 2; 0 0 5].
    long n = 3, nrhs = 2, lda = 4, ldb = 3;
    std::vector<float> a = {4, 2, 1, -1, 0.5f, 3, 2, -1, 0.25f, 0.5f, 5, -1};
    std::vector<int> ipiv = {2, 3, 3};
    std::vector<float> b = {1, 2, -7, 3, 4, -7, 5, 6, -7};
    long nn = n, rr = nrhs;
    unsigned long long trans = 0;
    if (rt.failure == "bad_n") nn = 0;
    if (rt.failure == "bad_ldb") ldb = 1;
    if (rt.failure == "bad_trans") trans = 2;
    if (rt.failure == "bad_ipiv") ipiv[1] = 4;
    rt.n = static_cast<int>(n);
    rt.nrhs = static_cast<int>(nrhs);
    rt.lda = lda;
    rt.ldb = ldb;
    const ops::hoststage::Limits lim{16384, 16383, 16383, 16384, 16383, 1};
    int info = -99;
    const int status = ops::hoststage::StagedGetrs(rt, "stub getrs", lim, sizeof(float), trans, nn, rr, a.data(), lda,
                                                   ipiv.data(), b.data(), ldb, &info);
    std::printf("RESULT %d %d %d %d %d %zu %d", status, rt.allocations, rt.frees, rt.duplicateFrees, rt.unsafeFrees,
                rt.live.size(), info);
    for (long i = 0; i < n; ++i)
        for (long c = 0; c < nrhs; ++c) std::printf(" %.6g", b[i * 3 + c]);
    std::printf("\n");
    for (void* p : rt.live) std::free(p);  // teardown of the stub's own host memory
    return 0;
}
