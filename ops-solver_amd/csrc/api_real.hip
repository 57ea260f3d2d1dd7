// Real LU entry points: argument contract, validation and dispatch.
// Reference: src/sgetrf/sgetrf_host.cpp:24-124, src/sgetri/sgetri_host.cpp:12-120,
// src/sgetri_batched/sgetri_batched_host.cpp:41-113, src/smatinv_batched/smatinv_batched_host.cpp:11-61,
// src/sgetrs/sgetrs_host.cpp:131-195, src/utils/real_lu_common.h:18-182, src/utils/pivot_check.h:27-42.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "host_getrs.h"
#include "internal.h"
#include "lu_kernels.h"

namespace ops
{
hipError_t MatinvBatched(int n, const float* const* Aarr, const float* Abase, int lda, int64_t strideA,
                         float* const* Carr, float* Cbase, int ldc, int64_t strideC, int* ipiv, int* info, int batch,
                         hipStream_t stream);
hipError_t GetriSmallBatched(int n, const float* const* Aarr, int lda, const int* piv, float* const* Carr, int ldc,
                             int* info, int batch, hipStream_t stream);
hipError_t TransposeRM(int rows, int cols, const float* src, int64_t lds, float* dst, int64_t ldd, hipStream_t stream);

namespace
{
constexpr int kMaxBatch = 10000000;  // src/utils/real_lu_common.h:18

int FactorRowLimit(aclsolverHandle_t h) { return H(h)->strict ? 8192 : kEngineMaxOrder; }
int GetriLimit(aclsolverHandle_t h) { return H(h)->strict ? 4096 : kEngineMaxOrder; }
int MatinvLimit(aclsolverHandle_t h) { return H(h)->strict ? 32 : 64; }

bool Overlap(const void* a, size_t aBytes, const void* b, size_t bBytes)
{
    if (aBytes == 0 || bBytes == 0) return false;
    auto x = reinterpret_cast<uintptr_t>(a), y = reinterpret_cast<uintptr_t>(b);
    return x <= y ? y - x < aBytes : x - y < bBytes;
}

// Capacity: the allocation holding p must have at least bytes from p on.
// One queried device allocation [base, end): pointers inside it need no further query.
struct RangeCache
{
    uintptr_t base = 0, end = 0;
};

bool Fits(const void* p, size_t bytes, RangeCache* cache = nullptr)
{
    const uintptr_t q = reinterpret_cast<uintptr_t>(p);
    if (cache != nullptr && cache->end != 0 && q >= cache->base && q < cache->end && bytes <= cache->end - q)
        return true;
    if (!IsDevice(p)) return false;
    void* base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess)
    {
        (void)hipGetLastError();
        return true;  // range unknown: accept (pool allocations)
    }
    auto off = reinterpret_cast<uintptr_t>(p) - reinterpret_cast<uintptr_t>(base);
    if (cache != nullptr)
    {
        cache->base = reinterpret_cast<uintptr_t>(base);
        cache->end = cache->base + size;
    }
    return off <= size && bytes <= size - off;
}

__global__ void first_zero_diag_kernel(int n, const float* a, int lda, int* info)
{
    __shared__ int s_first;
    if (threadIdx.x == 0) s_first = INT_MAX;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x)
        if (a[static_cast<size_t>(k) * lda + k] == 0.0f) atomicMin(&s_first, k);
    __syncthreads();
    if (threadIdx.x == 0) *info = s_first == INT_MAX ? 0 : s_first + 1;
}

__global__ void set_identity_kernel(int n, float* x, int ldx)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i < n) x[static_cast<size_t>(c) * ldx + i] = (i == c) ? 1.0f : 0.0f;
}

__global__ void copy_matrix_kernel(int rows, int cols, const float* src, int lds, float* dst, int ldd,
                                   const int* skipIfInfo)
{
    if (skipIfInfo != nullptr && *skipIfInfo != 0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i < rows && c < cols) dst[static_cast<size_t>(c) * ldd + i] = src[static_cast<size_t>(c) * lds + i];
}

// Copies pointer arrays (and pivots) to the host and checks each matrix view, as
// real_lu::ArrayValidation does (src/utils/real_lu_common.h:196-428): every pointer must be
// Device memory with room for its matrix, outputs must not overlap inputs, each other,
// the pivots or info. Returns 0 or the (1-based, handle excluded) argument index.
int CheckBatchArrays(aclsolverHandle_t h, int n, const void* aArr, int lda, const int* piv, const void* cArr, int ldc,
                     int* info, int batch, int aArg, int pivArg, int cArg, bool* contiguous, const float** firstA,
                     float** firstC)
{
    hipStream_t s = Stream(h);
    std::vector<uintptr_t> in(batch), out(batch);
    std::vector<int> pv(piv ? static_cast<size_t>(batch) * n : 0);
    if (hipMemcpyAsync(in.data(), aArr, sizeof(uintptr_t) * batch, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(out.data(), cArr, sizeof(uintptr_t) * batch, hipMemcpyDeviceToHost, s) != hipSuccess ||
        (piv && hipMemcpyAsync(pv.data(), piv, sizeof(int) * pv.size(), hipMemcpyDeviceToHost, s) != hipSuccess) ||
        hipStreamSynchronize(s) != hipSuccess)
        return -1;
    const size_t aBytes = static_cast<size_t>(lda) * n * sizeof(float);
    const size_t cBytes = static_cast<size_t>(ldc) * n * sizeof(float);
    *firstA = reinterpret_cast<const float*>(in[0]);
    *firstC = reinterpret_cast<float*>(out[0]);
    // Both arrays stepping through memory one whole matrix at a time (the usual layout):
    // the per-matrix views are then valid exactly when the two unions are, so one check of
    // each union replaces the per-matrix checks and the sort. Any failure falls through to
    // the per-matrix checks, which name the argument.
    bool seq = in[0] % alignof(float) == 0 && out[0] % alignof(float) == 0;
    for (int b = 1; seq && b < batch; ++b) seq = in[b] == in[0] + b * aBytes && out[b] == out[0] + b * cBytes;
    if (seq)
    {
        const size_t aAll = aBytes * batch, cAll = cBytes * batch;
        const void* pa = *firstA;
        const void* pc = *firstC;
        bool ok = Fits(pa, aAll) && Fits(pc, cAll) && !Overlap(pc, cAll, pa, aAll) &&
                  !Overlap(pc, cAll, info, sizeof(int) * batch) &&
                  !(piv && Overlap(pc, cAll, piv, sizeof(int) * pv.size())) &&
                  !Overlap(pc, cAll, aArr, sizeof(uintptr_t) * batch) && !Overlap(pc, cAll, cArr, sizeof(uintptr_t) * batch);
        for (size_t i = 0; ok && i < pv.size(); ++i)
            ok = pv[i] >= static_cast<int>(i % n) + 1 && pv[i] <= n;
        if (ok)
        {
            *contiguous = true;
            return 0;
        }
    }
    *contiguous = true;
    struct Span
    {
        uintptr_t b, e;
        int role;
    };
    std::vector<Span> spans;
    spans.reserve(static_cast<size_t>(batch) * 2);
    // Matrices of a batch usually share a few allocations: one query per allocation.
    RangeCache inCache, outCache;
    for (int b = 0; b < batch; ++b)
    {
        const void* pa = reinterpret_cast<const void*>(in[b]);
        const void* pc = reinterpret_cast<const void*>(out[b]);
        if (in[b] % alignof(float) != 0 || !Fits(pa, aBytes, &inCache)) return aArg;
        if (out[b] % alignof(float) != 0 || !Fits(pc, cBytes, &outCache)) return cArg;
        if (in[b] != in[0] + b * aBytes || out[b] != out[0] + b * cBytes) *contiguous = false;
        if (Overlap(pc, cBytes, info, sizeof(int) * batch)) return cArg;
        if (piv && Overlap(pc, cBytes, piv, sizeof(int) * pv.size())) return pivArg;
        if (Overlap(pc, cBytes, aArr, sizeof(uintptr_t) * batch) || Overlap(pc, cBytes, cArr, sizeof(uintptr_t) * batch))
            return cArg;
        spans.push_back({in[b], in[b] + aBytes, 0});
        spans.push_back({out[b], out[b] + cBytes, 1});
        if (piv)
            for (int k = 0; k < n; ++k)
            {
                int v = pv[static_cast<size_t>(b) * n + k];
                if (v < k + 1 || v > n) return pivArg;
            }
    }
    std::sort(spans.begin(), spans.end(),
              [](const Span& x, const Span& y) { return x.b < y.b || (x.b == y.b && x.role < y.role); });
    uintptr_t readEnd = 0, writeEnd = 0;
    for (const auto& sp : spans)
    {
        if (sp.role == 0)
        {
            if (sp.b < writeEnd) return cArg;
            readEnd = std::max(readEnd, sp.e);
        }
        else
        {
            if (sp.b < readEnd || sp.b < writeEnd) return cArg;
            writeEnd = sp.e;
        }
    }
    return 0;
}

}  // namespace
}  // namespace ops

using namespace ops;

// ---------------------------------------------------------------------------------------
// sgetrf
// ---------------------------------------------------------------------------------------
OPS_API aclsolverStatus_t aclsolverSgetrf_bufferSize(aclsolverHandle_t handle, int m, int n, float* A, int lda,
                                                     int* Lwork)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    if (m < 0 || m > FactorRowLimit(handle) || n < 0 || lda < std::max(1, m) || Lwork == nullptr ||
        IsDevice(Lwork) || (A != nullptr && !IsDevice(A)))
        return ACLSOLVER_STATUS_INVALID_VALUE;
    *Lwork = (m == 0 || n == 0) ? 0 : static_cast<int>((RealGetrfWorkspaceBytes(m, n) + 3) / 4);
    return ACLSOLVER_STATUS_SUCCESS;
}

namespace
{
aclsolverStatus_t Sgetrf(aclsolverHandle_t handle, int m, int n, float* A, int lda, float* Workspace, int* devIpiv,
                         int* devInfo, bool tournament)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    // FactorArgument, src/sgetrf/sgetrf_host.cpp:46-73
    int arg = 0;
    if (m < 0 || m > FactorRowLimit(handle))
        arg = 1;
    else if (n < 0 || static_cast<int64_t>(m) * n > (int64_t{1} << 40))
        arg = 2;
    else if (m > 0 && n > 0 && !IsDevice(A))
        arg = 3;
    else if (lda < std::max(1, m))
        arg = 4;
    else if (m > 0 && n > 0 && !IsDevice(Workspace))
        arg = 5;
    else if (devIpiv != nullptr && !IsDevice(devIpiv))
        arg = 6;
    else if (!IsDevice(devInfo))
        arg = 7;
    if (arg) return Invalid(handle, devInfo, 1, arg);
    hipStream_t s = Stream(handle);
    if (m == 0 || n == 0)
    {
        FillInfo(devInfo, 1, 0, s);
        return Launched();
    }
    // CheckSingle, src/utils/real_lu_common.h:61-89: alignment, capacity, disjoint views.
    const size_t aBytes = static_cast<size_t>(lda) * n * sizeof(float);
    const size_t wBytes = RealGetrfWorkspaceBytes(m, n);
    const size_t pBytes = static_cast<size_t>(std::min(m, n)) * sizeof(int);
    if (reinterpret_cast<uintptr_t>(devInfo) % alignof(int) != 0) return ACLSOLVER_STATUS_INVALID_VALUE;
    if (reinterpret_cast<uintptr_t>(A) % alignof(float) != 0 || !Fits(A, aBytes)) return Invalid(handle, devInfo, 1, 3);
    if (reinterpret_cast<uintptr_t>(Workspace) % alignof(float) != 0 || !Fits(Workspace, wBytes))
        return Invalid(handle, devInfo, 1, 5);
    if (devIpiv && (reinterpret_cast<uintptr_t>(devIpiv) % alignof(int) != 0 || !Fits(devIpiv, pBytes)))
        return Invalid(handle, devInfo, 1, 6);
    if (Overlap(A, aBytes, Workspace, wBytes)) return Invalid(handle, devInfo, 1, 5);
    if (devIpiv && Overlap(A, aBytes, devIpiv, pBytes)) return Invalid(handle, devInfo, 1, 6);
    if (devIpiv && Overlap(Workspace, wBytes, devIpiv, pBytes)) return Invalid(handle, devInfo, 1, 6);
    if (Overlap(A, aBytes, devInfo, sizeof(int)) || Overlap(Workspace, wBytes, devInfo, sizeof(int)) ||
        (devIpiv && Overlap(devIpiv, pBytes, devInfo, sizeof(int))))
        return ACLSOLVER_STATUS_INVALID_VALUE;
    // CheckFinite, src/sgetrf/sgetrf_host.cpp:118: non-finite input -> devInfo = -3.
    bool finite = true;
    if (!AllFinite(handle, A, m, n, lda, 0, 1, &finite)) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    if (!finite) return Invalid(handle, devInfo, 1, 3);
    if (RealGetrfPerm(m, n, A, lda, devIpiv, devInfo, Workspace, Precision::FP32, nullptr, s, nullptr,
                      HandleLookAhead(handle), tournament) != hipSuccess)
        return ACLSOLVER_STATUS_EXECUTION_FAILED;
    return Launched();
}
}  // namespace

OPS_API aclsolverStatus_t aclsolverSgetrf(aclsolverHandle_t handle, int m, int n, float* A, int lda, float* Workspace,
                                          int* devIpiv, int* devInfo)
{
    return Sgetrf(handle, m, n, A, lda, Workspace, devIpiv, devInfo, false);
}

// EXTENSION: the same factorization with tournament pivoting on the tall panels.
OPS_API aclsolverStatus_t aclsolverSgetrfTournament(aclsolverHandle_t handle, int m, int n, float* A, int lda,
                                                    float* Workspace, int* devIpiv, int* devInfo)
{
    return Sgetrf(handle, m, n, A, lda, Workspace, devIpiv, devInfo, true);
}

// ---------------------------------------------------------------------------------------
// Device column-major getrs (EXTENSION)
// ---------------------------------------------------------------------------------------
OPS_API aclsolverStatus_t aclsolverSgetrsDevice(aclsolverHandle_t handle, int trans, int n, int nrhs, const float* dA,
                                                int lda, const int* devIpiv, float* dB, int ldb, int* devInfo)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    int arg = 0;
    if (trans < 0 || trans > 1)
        arg = 1;
    else if (n < 0 || n > kEngineMaxOrder)
        arg = 2;
    else if (nrhs < 0)
        arg = 3;
    else if (n > 0 && !IsDevice(dA))
        arg = 4;
    else if (lda < std::max(1, n))
        arg = 5;
    else if (devIpiv != nullptr && !IsDevice(devIpiv))
        arg = 6;
    else if (n > 0 && nrhs > 0 && !IsDevice(dB))
        arg = 7;
    else if (ldb < std::max(1, n))
        arg = 8;
    else if (!IsDevice(devInfo))
        arg = 9;
    if (arg) return Invalid(handle, devInfo, 1, arg);
    hipStream_t s = Stream(handle);
    FillInfo(devInfo, 1, 0, s);
    if (n == 0 || nrhs == 0) return Launched();
    void* work = HandleScratch(handle, RealGetrsWorkspaceBytes(n, nrhs));
    if (work == nullptr) return ACLSOLVER_STATUS_ALLOC_FAILED;
    if (RealGetrs(trans, n, nrhs, dA, lda, devIpiv, dB, ldb, work, s, devInfo) != hipSuccess)
        return ACLSOLVER_STATUS_EXECUTION_FAILED;
    return Launched();
}

// ---------------------------------------------------------------------------------------
// Host-staged row-major sgetrs (src/sgetrs/sgetrs_host.cpp:131-195)
// ---------------------------------------------------------------------------------------
namespace
{
// The HIP side of the staged host getrs (host_getrs.h): copies and launches on the handle's
// stream, the real solve between the two transposes.
struct RealStagingRuntime
{
    hipStream_t s;
    int trans;
    int n, nrhs;
    int64_t lda, ldb;
    int Malloc(void** p, size_t bytes)
    {
        if (hipMalloc(p, bytes) == hipSuccess) return 0;
        (void)hipGetLastError();
        return 1;
    }
    int Free(void* p) { return hipFree(p) != hipSuccess; }
    int CopyToDevice(void* d, const void* h, size_t b) { return hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, s) != hipSuccess; }
    int CopyToHost(void* h, const void* d, size_t b) { return hipMemcpyAsync(h, d, b, hipMemcpyDeviceToHost, s) != hipSuccess; }
    int ClearWord(int* p)
    {
        FillInfo(p, 1, 0, s);
        return hipGetLastError() != hipSuccess;
    }
    int Synchronize() { return hipStreamSynchronize(s) != hipSuccess; }
    size_t WorkBytes(int nn, int rr) { return RealGetrsWorkspaceBytes(nn, rr); }
    int Solve(const hoststage::Staging& st)
    {
        float* aRow = static_cast<float*>(st.aRow);
        float* aCol = static_cast<float*>(st.aCol);
        float* bRow = static_cast<float*>(st.bRow);
        float* bCol = static_cast<float*>(st.bCol);
        // Row-major (r, c) at X[r*ld + c] is column-major X^T: transpose into column-major.
        hipError_t e = TransposeRM(n, n, aRow, lda, aCol, n, s);
        if (e == hipSuccess) e = TransposeRM(n, nrhs, bRow, ldb, bCol, n, s);
        if (e == hipSuccess) e = RealGetrs(trans, n, nrhs, aCol, n, st.piv, bCol, n, st.work, s, st.wait);
        // back: column-major (n x nrhs, ld n) -> row-major with ldb (the transpose of an
        // nrhs x n row-major view).
        if (e == hipSuccess) e = TransposeRM(nrhs, n, bCol, n, bRow, ldb, s);
        return e != hipSuccess;
    }
};
}  // namespace

OPS_API aclError aclsolverSgetrs(aclsolverHandle_t handle, const uint64_t trans, const int64_t n, const int64_t nrhs,
                                 float* A, const int64_t lda, int32_t* ipiv, float* B, const int64_t ldb,
                                 int32_t* info)
{
    const int64_t lim = (handle != nullptr && H(handle)->strict) ? 8192 : kEngineMaxOrder;
    const hoststage::Limits limits{lim, 16383, 16383, lim, 16383, 1};
    RealStagingRuntime rt{handle != nullptr ? Stream(handle) : nullptr, static_cast<int>(trans),
                          static_cast<int>(n), static_cast<int>(nrhs), lda, ldb};
    return hoststage::StagedGetrs(rt, "aclsolverSgetrs", limits, sizeof(float), trans, n, nrhs, A, lda, ipiv, B, ldb,
                                  info);
}

// ---------------------------------------------------------------------------------------
// sgetri (in place, from LU): X = A^-1 solved from the factors on the identity.
// ---------------------------------------------------------------------------------------
static size_t GetriWorkBytes(int n)
{
    return static_cast<size_t>(n) * n * sizeof(float) + RealGetrsWorkspaceBytes(n, n) + 256;
}

OPS_API aclsolverStatus_t aclsolverSgetri_bufferSize(aclsolverHandle_t handle, int n, float* A, int lda, int* Lwork)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    if (n < 0 || n > GetriLimit(handle) || lda < std::max(1, n) || Lwork == nullptr || IsDevice(Lwork) ||
        (A != nullptr && !IsDevice(A)))
        return ACLSOLVER_STATUS_INVALID_VALUE;
    *Lwork = n == 0 ? 0 : static_cast<int>((GetriWorkBytes(n) + 3) / 4);
    return ACLSOLVER_STATUS_SUCCESS;
}

OPS_API aclsolverStatus_t aclsolverSgetri(aclsolverHandle_t handle, int n, float* A, int lda, int* devIpiv,
                                          float* Workspace, int Lwork, int* devInfo)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    int arg = 0;
    if (n < 0 || n > GetriLimit(handle))
        arg = 1;
    else if (n > 0 && !IsDevice(A))
        arg = 2;
    else if (lda < std::max(1, n))
        arg = 3;
    else if (devIpiv != nullptr && !IsDevice(devIpiv))
        arg = 4;
    else if (n > 0 && !IsDevice(Workspace))
        arg = 5;
    else if (n > 0 && static_cast<size_t>(Lwork) * 4 < GetriWorkBytes(n))
        arg = 6;
    else if (!IsDevice(devInfo))
        arg = 7;
    if (arg) return Invalid(handle, devInfo, 1, arg);
    hipStream_t s = Stream(handle);
    if (n == 0)
    {
        FillInfo(devInfo, 1, 0, s);
        return Launched();
    }
    if (devIpiv)
    {
        std::vector<int> hp(n);
        if (hipMemcpyAsync(hp.data(), devIpiv, n * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return ACLSOLVER_STATUS_EXECUTION_FAILED;
        for (int k = 0; k < n; ++k)
            if (hp[k] < k + 1 || hp[k] > n) return Invalid(handle, devInfo, 1, 4);
    }
    bool finite = true;
    if (!AllFinite(handle, A, n, n, lda, 0, 1, &finite)) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    if (!finite) return Invalid(handle, devInfo, 1, 2);
    float* X = Workspace;
    void* work = reinterpret_cast<char*>(Workspace) + static_cast<size_t>(n) * n * sizeof(float);
    work = reinterpret_cast<char*>(work) + (256 - reinterpret_cast<uintptr_t>(work) % 256) % 256;
    hipLaunchKernelGGL(first_zero_diag_kernel, dim3(1), dim3(256), 0, s, n, A, lda, devInfo);
    hipLaunchKernelGGL(set_identity_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, X, n);
    if (RealGetrs(0, n, n, A, lda, devIpiv, X, n, work, s) != hipSuccess) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    hipLaunchKernelGGL(copy_matrix_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, n, X, n, A, lda, devInfo);
    return Launched();
}

// ---------------------------------------------------------------------------------------
// Batched inverses
// ---------------------------------------------------------------------------------------
OPS_API aclsolverStatus_t aclsolverSgetriBatched(aclsolverHandle_t handle, int n, const float* const Aarray[], int lda,
                                                 const int* PivotArray, float* const Carray[], int ldc,
                                                 int* infoArray, int batchSize)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    // BatchedArgument (factor = false), src/utils/real_lu_common.h:40-60
    int arg = 0;
    if (n < 0 || n > GetriLimit(handle))
        arg = 1;
    else if (n > 0 && batchSize > 0 && !IsDevice(Aarray))
        arg = 2;
    else if (lda < std::max(1, n))
        arg = 3;
    else if (PivotArray != nullptr && !IsDevice(PivotArray))
        arg = 4;
    else if (n > 0 && batchSize > 0 && !IsDevice(Carray))
        arg = 5;
    else if (ldc < std::max(1, n))
        arg = 6;
    else if (batchSize > 0 && !IsDevice(infoArray))
        arg = 7;
    else if (batchSize < 0 || batchSize > kMaxBatch)
        arg = 8;
    if (arg) return Invalid(handle, infoArray, batchSize < 0 ? 1 : (batchSize > 0 ? batchSize : 0), arg);
    hipStream_t s = Stream(handle);
    if (batchSize == 0) return ACLSOLVER_STATUS_SUCCESS;
    if (n == 0)
    {
        FillInfo(infoArray, batchSize, 0, s);
        return Launched();
    }
    bool contiguous = true;
    const float* a0 = nullptr;
    float* c0 = nullptr;
    int bad = CheckBatchArrays(handle, n, Aarray, lda, PivotArray, Carray, ldc, infoArray, batchSize, 2, 4, 5,
                               &contiguous, &a0, &c0);
    if (bad < 0) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    if (bad) return Invalid(handle, infoArray, batchSize, bad);
    bool finite = true;
    if (!AllFiniteArrays(handle, Aarray, n, n, lda, batchSize, &finite)) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    if (!finite) return Invalid(handle, infoArray, batchSize, 2);
    if (n <= 64)
    {
        if (GetriSmallBatched(n, Aarray, lda, PivotArray, Carray, ldc, infoArray, batchSize, s) != hipSuccess)
            return ACLSOLVER_STATUS_EXECUTION_FAILED;
        return Launched();
    }
    if (n <= 512)
    {
        // Blocked batched inverse, in chunks whose diagonal-block inverses fit 64 MiB.
        const size_t per = GetriBlockedBatchedWorkspaceBytes(n, 1);
        const int chunk = static_cast<int>(std::max<size_t>(1, std::min<size_t>(65535, (64u << 20) / per)));
        float* work = static_cast<float*>(HandleScratch(handle, per * std::min(chunk, batchSize)));
        if (work == nullptr) return ACLSOLVER_STATUS_ALLOC_FAILED;
        for (int b0 = 0; b0 < batchSize; b0 += chunk)
        {
            const int nb = std::min(chunk, batchSize - b0);
            if (GetriBlockedBatched(n, Aarray + b0, lda, PivotArray ? PivotArray + static_cast<size_t>(b0) * n : nullptr,
                                    Carray + b0, ldc, infoArray + b0, nb, work, s) != hipSuccess)
                return ACLSOLVER_STATUS_EXECUTION_FAILED;
        }
        return Launched();
    }
    // Larger orders: per matrix, solve the factors against the identity straight into C.
    std::vector<uintptr_t> in(batchSize), out(batchSize);
    if (hipMemcpyAsync(in.data(), Aarray, sizeof(uintptr_t) * batchSize, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(out.data(), Carray, sizeof(uintptr_t) * batchSize, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return ACLSOLVER_STATUS_EXECUTION_FAILED;
    void* work = HandleScratch(handle, RealGetrsWorkspaceBytes(n, n));
    if (work == nullptr) return ACLSOLVER_STATUS_ALLOC_FAILED;
    for (int b = 0; b < batchSize; ++b)
    {
        const float* a = reinterpret_cast<const float*>(in[b]);
        float* c = reinterpret_cast<float*>(out[b]);
        hipLaunchKernelGGL(first_zero_diag_kernel, dim3(1), dim3(256), 0, s, n, a, lda, infoArray + b);
        hipLaunchKernelGGL(set_identity_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, c, ldc);
        if (RealGetrs(0, n, n, a, lda, PivotArray ? PivotArray + static_cast<size_t>(b) * n : nullptr, c, ldc, work,
                      s) != hipSuccess)
            return ACLSOLVER_STATUS_EXECUTION_FAILED;
    }
    return Launched();
}

OPS_API aclsolverStatus_t aclsolverSmatinvBatched(aclsolverHandle_t handle, int n, const float* const A[], int lda,
                                                  float* const Ainv[], int lda_inv, int* info, int batchSize)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    // BatchedArgument (factor = true): n=1, A=2, lda=3, Ainv=4, lda_inv=5, info=6, batch=7.
    int arg = 0;
    if (n < 0 || n > MatinvLimit(handle))
        arg = 1;
    else if (n > 0 && batchSize > 0 && !IsDevice(A))
        arg = 2;
    else if (lda < std::max(1, n))
        arg = 3;
    else if (n > 0 && batchSize > 0 && !IsDevice(Ainv))
        arg = 4;
    else if (lda_inv < std::max(1, n))
        arg = 5;
    else if (batchSize > 0 && !IsDevice(info))
        arg = 6;
    else if (batchSize < 0 || batchSize > kMaxBatch)
        arg = 7;
    if (arg) return Invalid(handle, info, batchSize < 0 ? 1 : (batchSize > 0 ? batchSize : 0), arg);
    hipStream_t s = Stream(handle);
    if (batchSize == 0) return ACLSOLVER_STATUS_SUCCESS;
    if (n == 0)
    {
        FillInfo(info, batchSize, 0, s);
        return Launched();
    }
    bool contiguous = true;
    const float* a0 = nullptr;
    float* c0 = nullptr;
    int bad = CheckBatchArrays(handle, n, A, lda, nullptr, Ainv, lda_inv, info, batchSize, 2, 4, 4, &contiguous, &a0,
                               &c0);
    if (bad < 0) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    if (bad) return Invalid(handle, info, batchSize, bad);
    bool finite = true;
    if (!AllFiniteArrays(handle, A, n, n, lda, batchSize, &finite)) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    if (!finite) return Invalid(handle, info, batchSize, 2);
    hipError_t e;
    if (contiguous)
        e = MatinvBatched(n, nullptr, a0, lda, static_cast<int64_t>(lda) * n, nullptr, c0, lda_inv,
                          static_cast<int64_t>(lda_inv) * n, nullptr, info, batchSize, s);
    else
        e = MatinvBatched(n, A, nullptr, lda, 0, Ainv, nullptr, lda_inv, 0, nullptr, info, batchSize, s);
    if (e != hipSuccess) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    return Launched();
}

// EXTENSION for measurement and the uneven-batch path: contiguous strided batched direct
// inverse with per-matrix pivots, no validation pass (the caller owns the contract).
OPS_API aclsolverStatus_t aclsolverSmatinvStridedBatched(aclsolverHandle_t handle, int n, const float* A, int lda,
                                                         long long strideA, float* Ainv, int lda_inv,
                                                         long long strideInv, int* ipiv, int* info, int batchSize)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    if (n < 0 || n > 64 || lda < std::max(1, n) || lda_inv < std::max(1, n) || batchSize < 0 ||
        strideA < static_cast<long long>(lda) * n || strideInv < static_cast<long long>(lda_inv) * n)
        return ACLSOLVER_STATUS_INVALID_VALUE;
    if (batchSize == 0 || n == 0) return ACLSOLVER_STATUS_SUCCESS;
    if (MatinvBatched(n, nullptr, A, lda, strideA, nullptr, Ainv, lda_inv, strideInv, ipiv, info, batchSize,
                      Stream(handle)) != hipSuccess)
        return ACLSOLVER_STATUS_EXECUTION_FAILED;
    return Launched();
}
