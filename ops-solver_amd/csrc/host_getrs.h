// Host row-major getrs (aclsolverSgetrs / aclsolverCgetrs): argument checks, staging to the
// device and back, and the single-owner cleanup of the staging buffer, written against a
// small runtime interface. The library instantiates it with the HIP runtime; the CPU tests
// instantiate the same code with a stub runtime that fails the k-th allocation, copy, launch
// or synchronization, and check status, info and buffer ownership after each failure.
//
// Contract (reference src/sgetrs/sgetrs_host.cpp:131-195, src/cgetrs/cgetrs_host.cpp:151-231,
// src/utils/pivot_check.h:27-42):
//   * bad sizes, leading dimensions, trans or null buffers: INVALID_PARAM, info untouched,
//     no device work;
//   * an ipiv entry outside [1, n]: info = -4, INVALID_PARAM, no device work;
//   * allocation failure: BAD_ALLOC, info untouched;
//   * copy or launch failure: RT_FAILURE after the stream has drained, staging freed,
//     info untouched;
//   * synchronization failure: RT_FAILURE, staging kept (work may still be in flight);
//   * success: B holds X, info = 0.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>

namespace ops
{
namespace hoststage
{
// Status values of the getrs ABI (aclError).
constexpr int kOk = 0;
constexpr int kInvalidParam = 100000;
constexpr int kBadAlloc = 200000;
constexpr int kRuntimeFailure = 500000;

struct Limits
{
    int64_t maxOrder;     // n and nrhs
    int64_t maxLdaPad;    // lda - n
    int64_t maxLdbPad;    // ldb - nrhs
    int64_t wideRhs;      // nrhs above this ...
    int64_t wideLdbPad;   // ... allows at most this ldb - nrhs
    uint64_t maxTrans;
};

// Device layout of one staged solve: the row-major copies of A and B as given (lda, ldb),
// their column-major transposes, the pivots, a wait-report word and the solver workspace.
struct Staging
{
    void* base = nullptr;
    size_t bytes = 0;
    void* aRow = nullptr;
    void* aCol = nullptr;
    void* bRow = nullptr;
    void* bCol = nullptr;
    int* piv = nullptr;
    int* wait = nullptr;
    void* work = nullptr;
};

inline Staging Layout(size_t elem, int64_t n, int64_t nrhs, int64_t lda, int64_t ldb, size_t workBytes)
{
    Staging s;
    const size_t aRow = static_cast<size_t>(n) * lda, bRow = static_cast<size_t>(n) * ldb;
    const size_t aCol = static_cast<size_t>(n) * n, bCol = static_cast<size_t>(n) * nrhs;
    size_t off = 0;
    auto take = [&](size_t b) {
        const size_t at = off;
        off = (off + b + 255) / 256 * 256;
        return at;
    };
    const size_t oARow = take(aRow * elem), oACol = take(aCol * elem), oBRow = take(bRow * elem),
                 oBCol = take(bCol * elem), oPiv = take(static_cast<size_t>(n) * 4), oWait = take(4),
                 oWork = take(workBytes);
    s.bytes = off;
    s.aRow = reinterpret_cast<void*>(oARow);
    s.aCol = reinterpret_cast<void*>(oACol);
    s.bRow = reinterpret_cast<void*>(oBRow);
    s.bCol = reinterpret_cast<void*>(oBCol);
    s.piv = reinterpret_cast<int*>(oPiv);
    s.wait = reinterpret_cast<int*>(oWait);
    s.work = reinterpret_cast<void*>(oWork);
    return s;
}

template <typename P>
inline P* At(void* base, P* offset)
{
    return reinterpret_cast<P*>(static_cast<char*>(base) + reinterpret_cast<uintptr_t>(offset));
}

// RT provides (all return 0 on success; copies and launches are enqueued on one stream):
//   int Malloc(void** p, size_t bytes);  int Free(void* p);
//   int CopyToDevice(void* dst, const void* src, size_t bytes);
//   int CopyToHost(void* dst, const void* src, size_t bytes);
//   int ClearWord(int* p);                       // *p = 0 on the device
//   int Solve(const Staging& s);                 // transposes in, solve, transpose out
//   int Synchronize();
//   size_t WorkBytes(int n, int nrhs);           // device solve workspace
template <typename RT>
int StagedGetrs(RT& rt, const char* name, const Limits& lim, size_t elem, uint64_t trans, int64_t n, int64_t nrhs,
                void* A, int64_t lda, int32_t* ipiv, void* B, int64_t ldb, int32_t* info)
{
    if (!(n > 0 && n <= lim.maxOrder && nrhs > 0 && nrhs <= lim.maxOrder && lda >= n && ldb >= nrhs &&
          A != nullptr && B != nullptr && ipiv != nullptr && info != nullptr && lda - n <= lim.maxLdaPad &&
          ldb - nrhs <= lim.maxLdbPad && (nrhs <= lim.wideRhs || ldb - nrhs <= lim.wideLdbPad)))
    {
        std::fprintf(stderr, "[ERROR] %s invalid param: n, nrhs, lda, ldb or a null buffer.\n", name);
        return kInvalidParam;
    }
    if (trans > lim.maxTrans)
    {
        std::fprintf(stderr, "[ERROR] %s invalid param: trans must be at most %llu.\n", name,
                     static_cast<unsigned long long>(lim.maxTrans));
        return kInvalidParam;
    }
    for (int64_t i = 0; i < n; ++i)
        if (ipiv[i] < 1 || ipiv[i] > n)
        {
            *info = -4;
            std::fprintf(stderr, "[ERROR] %s invalid ipiv[%lld] = %d, expected in [1, %lld].\n", name,
                         static_cast<long long>(i), ipiv[i], static_cast<long long>(n));
            return kInvalidParam;
        }
    Staging s = Layout(elem, n, nrhs, lda, ldb, rt.WorkBytes(static_cast<int>(n), static_cast<int>(nrhs)));
    void* base = nullptr;
    if (rt.Malloc(&base, s.bytes) != 0 || base == nullptr) return kBadAlloc;
    s.base = base;
    s.aRow = At(base, s.aRow);
    s.aCol = At(base, s.aCol);
    s.bRow = At(base, s.bRow);
    s.bCol = At(base, s.bCol);
    s.piv = At(base, s.piv);
    s.wait = At(base, s.wait);
    s.work = At(base, s.work);
    const size_t aBytes = static_cast<size_t>(n) * lda * elem, bBytes = static_cast<size_t>(n) * ldb * elem;
    int waitWord = 0;
    int e = rt.ClearWord(s.wait);
    if (e == 0) e = rt.CopyToDevice(s.aRow, A, aBytes);
    if (e == 0) e = rt.CopyToDevice(s.bRow, B, bBytes);
    if (e == 0) e = rt.CopyToDevice(s.piv, ipiv, static_cast<size_t>(n) * sizeof(int32_t));
    if (e == 0) e = rt.Solve(s);
    if (e == 0) e = rt.CopyToHost(&waitWord, s.wait, sizeof(int));
    if (e == 0) e = rt.CopyToHost(B, s.bRow, bBytes);  // last: no earlier failure touches B
    // Whatever was enqueued must finish before the staging can be released; if the stream
    // cannot be drained the buffer stays allocated (a later free could race the device).
    if (rt.Synchronize() != 0) return kRuntimeFailure;
    (void)rt.Free(base);
    if (e != 0) return kRuntimeFailure;
    if (waitWord != 0)
    {
        std::fprintf(stderr, "[ERROR] %s: a bounded wait of the triangular solve gave up.\n", name);
        return kRuntimeFailure;
    }
    *info = 0;
    return kOk;
}
}  // namespace hoststage
}  // namespace ops
