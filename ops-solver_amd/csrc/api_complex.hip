// Complex LU entry points: argument contract, validation and dispatch.
// Reference: src/cgetrf/cgetrf_host.cpp:20-82, src/cgetri/cgetri_host.cpp:76-115,
// src/cgetri_batched/cgetri_batched_host.cpp:34-79, src/cmatinv_batched/cmatinv_batched_host.cpp:34-75,
// src/cgetrs/cgetrs_host.cpp:95-231, src/utils/complex_solver_host.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "host_getrs.h"
#include "internal.h"

namespace ops
{
hipError_t CMatinvBatched(int n, const float2* const* Aarr, const float2* Abase, int lda, int64_t strideA,
                          float2* const* Carr, float2* Cbase, int ldc, int64_t strideC, int* info, int batch,
                          hipStream_t stream);
hipError_t CGetriSmallBatched(int n, const float2* const* Aarr, int lda, const int* piv, float2* const* Carr, int ldc,
                              int* info, int batch, hipStream_t stream);

namespace
{
constexpr int kMaxBatchMatinv = 10000000;  // cmatinv_batched_host.cpp batch <= 1e7
constexpr int kMaxBatchGetri = 1000000;    // cgetri_batched batch <= 1e6

int RowLimit(aclsolverHandle_t h) { return H(h)->strict ? 4096 : 16384; }
int ColLimit(aclsolverHandle_t h) { return H(h)->strict ? 8192 : 16384; }
int MatinvLimit(aclsolverHandle_t h) { return H(h)->strict ? 32 : 64; }

__global__ void c_first_zero_diag_kernel(int n, const float2* a, int lda, int* info)
{
    __shared__ int s_first;
    if (threadIdx.x == 0) s_first = INT_MAX;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x)
    {
        const float2 v = a[static_cast<size_t>(k) * lda + k];
        if (v.x == 0.0f && v.y == 0.0f) atomicMin(&s_first, k);
    }
    __syncthreads();
    if (threadIdx.x == 0) *info = s_first == INT_MAX ? 0 : s_first + 1;
}

__global__ void c_identity_kernel(int n, float2* x, int ldx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i < n) x[static_cast<size_t>(c) * ldx + i] = make_float2(i == c ? 1.0f : 0.0f, 0.0f);
}

__global__ void c_copy_kernel(int rows, int cols, const float2* src, int lds, float2* dst, int ldd, const int* skipIf)
{
    if (skipIf != nullptr && *skipIf != 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i < rows && c < cols) dst[static_cast<size_t>(c) * ldd + i] = src[static_cast<size_t>(c) * lds + i];
}

__global__ __launch_bounds__(256) void c_transpose_rm_kernel(int rows, int cols, const float2* __restrict__ src,
                                                             int64_t lds, float2* __restrict__ dst, int64_t ldd)
{
    __shared__ float2 tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int q = ty; q < 32; q += 8)
    {
        const int r = r0 + q, c = c0 + tx;
        if (r < rows && c < cols) tile[q][tx] = src[r * lds + c];
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8)
    {
        const int c = c0 + q, r = r0 + tx;
        if (r < rows && c < cols) dst[c * ldd + r] = tile[tx][q];
    }
}

hipError_t CTransposeRM(int rows, int cols, const float2* src, int64_t lds, float2* dst, int64_t ldd, hipStream_t s)
{
    hipLaunchKernelGGL(c_transpose_rm_kernel, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(256), 0, s, rows, cols,
                       src, lds, dst, ldd);
    return hipGetLastError();
}

// Device-pointer check that remembers the last queried allocation [base, end), so the
// matrices of a batch that share an allocation need one query, not one each.
struct DeviceRange
{
    uintptr_t base = 0, end = 0;
    bool Holds(const void* p)
    {
        const uintptr_t q = reinterpret_cast<uintptr_t>(p);
        if (end != 0 && q >= base && q < end) return true;
        if (!IsDevice(p)) return false;
        void* b = nullptr;
        size_t size = 0;
        if (hipMemGetAddressRange(&b, &size, const_cast<void*>(p)) == hipSuccess)
        {
            base = reinterpret_cast<uintptr_t>(b);
            end = base + size;
        }
        else
            (void)hipGetLastError();
        return true;
    }
};

bool Overlap(const void* a, size_t aBytes, const void* b, size_t bBytes)
{
    if (aBytes == 0 || bBytes == 0) return false;
    auto x = reinterpret_cast<uintptr_t>(a), y = reinterpret_cast<uintptr_t>(b);
    return x <= y ? y - x < aBytes : x - y < bBytes;
}

// Pointer arrays to host + per-matrix checks (ComplexSolver::ValidateMatrices): Device
// pointers, outputs disjoint from inputs, each other and info. Returns 0 or the argument.
int CheckArrays(aclsolverHandle_t h, int n, const void* aArr, int lda, const void* cArr, int ldc, const int* info,
                int batch, const int* piv, int aArg, int cArg, int pivArg, std::vector<uintptr_t>* inOut,
                std::vector<uintptr_t>* outOut)
{
    hipStream_t s = Stream(h);
    std::vector<uintptr_t> in(batch), out(batch);
    std::vector<int> pv(piv ? static_cast<size_t>(batch) * n : 0);
    if (hipMemcpyAsync(in.data(), aArr, sizeof(uintptr_t) * batch, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(out.data(), cArr, sizeof(uintptr_t) * batch, hipMemcpyDeviceToHost, s) != hipSuccess ||
        (piv && hipMemcpyAsync(pv.data(), piv, sizeof(int) * pv.size(), hipMemcpyDeviceToHost, s) != hipSuccess) ||
        hipStreamSynchronize(s) != hipSuccess)
        return -1;
    const size_t aBytes = static_cast<size_t>(lda) * n * sizeof(float2);
    const size_t cBytes = static_cast<size_t>(ldc) * n * sizeof(float2);
    // Both arrays stepping through one allocation each, one whole matrix at a time: the
    // per-matrix checks reduce to checks of the two unions. Any failure falls through to the
    // per-matrix checks, which name the argument.
    bool seq = in[0] % alignof(float) == 0 && out[0] % alignof(float) == 0;
    for (int b = 1; seq && b < batch; ++b) seq = in[b] == in[0] + b * aBytes && out[b] == out[0] + b * cBytes;
    if (seq)
    {
        DeviceRange ra, rc;
        const size_t aAll = aBytes * batch, cAll = cBytes * batch;
        const void* pa = reinterpret_cast<const void*>(in[0]);
        const void* pc = reinterpret_cast<const void*>(out[0]);
        bool ok = ra.Holds(pa) && ra.end != 0 && in[batch - 1] < ra.end && rc.Holds(pc) && rc.end != 0 &&
                  out[batch - 1] < rc.end && !Overlap(pc, cAll, pa, aAll) && !Overlap(pc, cAll, info, sizeof(int) * batch);
        for (size_t i = 0; ok && i < pv.size(); ++i)
            ok = pv[i] >= static_cast<int>(i % n) + 1 && pv[i] <= n;
        if (ok)
        {
            if (inOut) *inOut = in;
            if (outOut) *outOut = out;
            return 0;
        }
    }
    struct Span
    {
        uintptr_t b, e;
        int role;
    };
    std::vector<Span> spans;
    DeviceRange inRange, outRange;
    for (int b = 0; b < batch; ++b)
    {
        if (in[b] % alignof(float) || !inRange.Holds(reinterpret_cast<void*>(in[b]))) return aArg;
        if (out[b] % alignof(float) || !outRange.Holds(reinterpret_cast<void*>(out[b]))) return cArg;
        if (Overlap(reinterpret_cast<void*>(out[b]), cBytes, info, sizeof(int) * batch)) return cArg;
        spans.push_back({in[b], in[b] + aBytes, 0});
        spans.push_back({out[b], out[b] + cBytes, 1});
        if (piv)
            for (int k = 0; k < n; ++k)
            {
                const int v = pv[static_cast<size_t>(b) * n + k];
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
    if (inOut) *inOut = in;
    if (outOut) *outOut = out;
    return 0;
}

// Factor-and-solve inverse of one matrix into C (used beyond the register kernels' orders).
aclsolverStatus_t InverseByFactor(aclsolverHandle_t h, int n, const float2* a, int lda, const int* pivIn, bool factor,
                                  float2* c, int ldc, int* info)
{
    hipStream_t s = Stream(h);
    const size_t luBytes = static_cast<size_t>(n) * n * sizeof(float2);
    const size_t need = luBytes + static_cast<size_t>(n) * sizeof(int) + ComplexGetrfWorkspaceBytes(n, n) +
                        ComplexGetrsWorkspaceBytes(n, n) + 1024;
    char* base = static_cast<char*>(HandleScratch(h, need));
    if (base == nullptr) return ACLSOLVER_STATUS_ALLOC_FAILED;
    float2* lu = reinterpret_cast<float2*>(base);
    int* piv = reinterpret_cast<int*>(base + luBytes);
    char* gw = base + luBytes + ((static_cast<size_t>(n) * sizeof(int) + 255) / 256) * 256;
    char* sw = gw + ((ComplexGetrfWorkspaceBytes(n, n) + 255) / 256) * 256;
    hipLaunchKernelGGL(c_copy_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, n, a, lda, lu, n, nullptr);
    const int* p = pivIn;
    if (factor)
    {
        if (ComplexGetrf(n, n, lu, n, piv, info, gw, s) != hipSuccess) return ACLSOLVER_STATUS_EXECUTION_FAILED;
        p = piv;
    }
    else
        hipLaunchKernelGGL(c_first_zero_diag_kernel, dim3(1), dim3(256), 0, s, n, lu, n, info);
    hipLaunchKernelGGL(c_identity_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, c, ldc);
    if (ComplexGetrs(0, n, n, lu, n, p, c, ldc, sw, s) != hipSuccess) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    return Launched();
}
}  // namespace
}  // namespace ops

using namespace ops;

OPS_API aclsolverStatus_t aclsolverCgetrf_bufferSize(aclsolverHandle_t handle, int m, int n, aclFloatComplex* A,
                                                     int lda, int* Lwork)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    if (m < 0 || m > RowLimit(handle) || n < 0 || n > ColLimit(handle) || lda < std::max(1, m) ||
        Lwork == nullptr || (m > 0 && n > 0 && !IsDevice(A)))
        return ACLSOLVER_STATUS_INVALID_VALUE;
    *Lwork = static_cast<int>((ComplexGetrfWorkspaceBytes(m, n) + sizeof(float2) - 1) / sizeof(float2));
    return ACLSOLVER_STATUS_SUCCESS;
}

OPS_API aclsolverStatus_t aclsolverCgetrf(aclsolverHandle_t handle, int m, int n, aclFloatComplex* A, int lda,
                                          aclFloatComplex* Workspace, int* devIpiv, int* devInfo)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    int arg = 0;
    if (m < 0 || m > RowLimit(handle))
        arg = 1;
    else if (n < 0 || n > ColLimit(handle))
        arg = 2;
    else if (lda < std::max(1, m))
        arg = 4;
    if (arg) return Invalid(handle, devInfo, 1, arg);
    if (!IsDevice(devInfo) || reinterpret_cast<uintptr_t>(devInfo) % alignof(int)) return ACLSOLVER_STATUS_INVALID_VALUE;
    if (m > 0 && n > 0)
    {
        if (!IsDevice(A)) return Invalid(handle, devInfo, 1, 3);
        if (devIpiv != nullptr && !IsDevice(devIpiv)) return Invalid(handle, devInfo, 1, 6);
        if (!IsDevice(Workspace)) return Invalid(handle, devInfo, 1, 5);
    }
    hipStream_t s = Stream(handle);
    FillInfo(devInfo, 1, 0, s);
    if (m == 0 || n == 0) return Launched();
    if (ComplexGetrf(m, n, reinterpret_cast<float2*>(A), lda, devIpiv, devInfo, Workspace, s,
                     HandleLookAhead(handle)) != hipSuccess)
        return ACLSOLVER_STATUS_EXECUTION_FAILED;
    return Launched();
}

OPS_API aclsolverStatus_t aclsolverCgetrsDevice(aclsolverHandle_t handle, int trans, int n, int nrhs,
                                                const aclFloatComplex* dA, int lda, const int* devIpiv,
                                                aclFloatComplex* dB, int ldb, int* devInfo)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    int arg = 0;
    if (trans < 0 || trans > 2)
        arg = 1;
    else if (n < 0 || n > 16384)
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
    void* work = HandleScratch(handle, ComplexGetrsWorkspaceBytes(n, nrhs));
    if (work == nullptr) return ACLSOLVER_STATUS_ALLOC_FAILED;
    if (ComplexGetrs(trans, n, nrhs, reinterpret_cast<const float2*>(dA), lda, devIpiv, reinterpret_cast<float2*>(dB),
                     ldb, work, s) != hipSuccess)
        return ACLSOLVER_STATUS_EXECUTION_FAILED;
    return Launched();
}

// Host-staged row-major cgetrs (src/cgetrs/cgetrs_host.cpp:151-231).
namespace
{
// The HIP side of the staged host getrs (host_getrs.h) for complex64.
struct ComplexStagingRuntime
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
    size_t WorkBytes(int nn, int rr) { return ComplexGetrsWorkspaceBytes(nn, rr); }
    int Solve(const hoststage::Staging& st)
    {
        float2* aRow = static_cast<float2*>(st.aRow);
        float2* aCol = static_cast<float2*>(st.aCol);
        float2* bRow = static_cast<float2*>(st.bRow);
        float2* bCol = static_cast<float2*>(st.bCol);
        hipError_t e = CTransposeRM(n, n, aRow, lda, aCol, n, s);
        if (e == hipSuccess) e = CTransposeRM(n, nrhs, bRow, ldb, bCol, n, s);
        if (e == hipSuccess) e = ComplexGetrs(trans, n, nrhs, aCol, n, st.piv, bCol, n, st.work, s);
        if (e == hipSuccess) e = CTransposeRM(nrhs, n, bCol, n, bRow, ldb, s);
        return e != hipSuccess;
    }
};
}  // namespace

OPS_API aclError aclsolverCgetrs(aclsolverHandle_t handle, const uint64_t trans, const int64_t n, const int64_t nrhs,
                                 aclFloatComplex* A, const int64_t lda, int32_t* ipiv, aclFloatComplex* B,
                                 const int64_t ldb, int32_t* info)
{
    const int64_t lim = (handle != nullptr && H(handle)->strict) ? 8192 : 16384;
    const hoststage::Limits limits{lim, 8191, 8191, 4096, 4095, 2};
    ComplexStagingRuntime rt{handle != nullptr ? Stream(handle) : nullptr, static_cast<int>(trans),
                             static_cast<int>(n), static_cast<int>(nrhs), lda, ldb};
    return hoststage::StagedGetrs(rt, "aclsolverCgetrs", limits, sizeof(float2), trans, n, nrhs, A, lda, ipiv, B, ldb,
                                  info);
}

static size_t CGetriWorkBytes(int n)
{
    return static_cast<size_t>(n) * n * sizeof(float2) + ComplexGetrsWorkspaceBytes(n, n) + 256;
}

OPS_API aclsolverStatus_t aclsolverCgetri_bufferSize(aclsolverHandle_t handle, int n, aclFloatComplex* A, int lda,
                                                     int* Lwork)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    const int lim = H(handle)->strict ? 4096 : 16384;
    if (n < 0 || n > lim || lda < std::max(1, n) || Lwork == nullptr || (n > 0 && A != nullptr && !IsDevice(A)))
        return ACLSOLVER_STATUS_INVALID_VALUE;
    *Lwork = n == 0 ? 0 : static_cast<int>((CGetriWorkBytes(n) + sizeof(float2) - 1) / sizeof(float2));
    return ACLSOLVER_STATUS_SUCCESS;
}

OPS_API aclsolverStatus_t aclsolverCgetri(aclsolverHandle_t handle, int n, aclFloatComplex* A, int lda, int* devIpiv,
                                          aclFloatComplex* Workspace, int Lwork, int* devInfo)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    const int lim = H(handle)->strict ? 4096 : 16384;
    int arg = 0;
    if (n < 0 || n > lim)
        arg = 1;
    else if (n > 0 && !IsDevice(A))
        arg = 2;
    else if (lda < std::max(1, n))
        arg = 3;
    else if (devIpiv != nullptr && !IsDevice(devIpiv))
        arg = 4;
    else if (n > 0 && !IsDevice(Workspace))
        arg = 5;
    else if (n > 0 && static_cast<size_t>(Lwork) * sizeof(float2) < CGetriWorkBytes(n))
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
    float2* X = reinterpret_cast<float2*>(Workspace);
    char* w = reinterpret_cast<char*>(X + static_cast<size_t>(n) * n);
    w += (256 - reinterpret_cast<uintptr_t>(w) % 256) % 256;
    float2* a = reinterpret_cast<float2*>(A);
    hipLaunchKernelGGL(c_first_zero_diag_kernel, dim3(1), dim3(256), 0, s, n, a, lda, devInfo);
    hipLaunchKernelGGL(c_identity_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, X, n);
    if (ComplexGetrs(0, n, n, a, lda, devIpiv, X, n, w, s) != hipSuccess) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    hipLaunchKernelGGL(c_copy_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, n, X, n, a, lda, devInfo);
    return Launched();
}

OPS_API aclsolverStatus_t aclsolverCgetriBatched(aclsolverHandle_t handle, int n,
                                                 const aclFloatComplex* const Aarray[], int lda, const int* PivotArray,
                                                 aclFloatComplex* const Carray[], int ldc, int* infoArray,
                                                 int batchSize)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    const int lim = H(handle)->strict ? 4096 : 16384;
    int arg = 0;
    if (n < 0 || n > lim)
        arg = 1;
    else if (lda < std::max(1, n))
        arg = 3;
    else if (ldc < std::max(1, n))
        arg = 6;
    else if (batchSize < 0 || batchSize > kMaxBatchGetri)
        arg = 8;
    if (arg) return Invalid(handle, infoArray, batchSize < 0 ? 1 : (batchSize > 0 ? batchSize : 0), arg);
    if (batchSize == 0) return ACLSOLVER_STATUS_SUCCESS;
    if (!IsDevice(infoArray)) return ACLSOLVER_STATUS_INVALID_VALUE;
    hipStream_t s = Stream(handle);
    if (n == 0)
    {
        FillInfo(infoArray, batchSize, 0, s);
        return Launched();
    }
    if (!IsDevice(Aarray)) return Invalid(handle, infoArray, batchSize, 2);
    if (PivotArray != nullptr && !IsDevice(PivotArray)) return Invalid(handle, infoArray, batchSize, 4);
    if (!IsDevice(Carray)) return Invalid(handle, infoArray, batchSize, 5);
    std::vector<uintptr_t> in, out;
    int bad = CheckArrays(handle, n, Aarray, lda, Carray, ldc, infoArray, batchSize, PivotArray, 2, 5, 4, &in, &out);
    if (bad < 0) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    if (bad) return Invalid(handle, infoArray, batchSize, bad);
    const auto* A2 = reinterpret_cast<const float2* const*>(Aarray);
    auto* C2 = reinterpret_cast<float2* const*>(Carray);
    if (n <= 64)
    {
        if (CGetriSmallBatched(n, A2, lda, PivotArray, C2, ldc, infoArray, batchSize, s) != hipSuccess)
            return ACLSOLVER_STATUS_EXECUTION_FAILED;
        return Launched();
    }
    for (int b = 0; b < batchSize; ++b)
    {
        auto st = InverseByFactor(handle, n, reinterpret_cast<const float2*>(in[b]), lda,
                                  PivotArray ? PivotArray + static_cast<size_t>(b) * n : nullptr, false,
                                  reinterpret_cast<float2*>(out[b]), ldc, infoArray + b);
        if (st != ACLSOLVER_STATUS_SUCCESS) return st;
    }
    return Launched();
}

OPS_API aclsolverStatus_t aclsolverCmatinvBatched(aclsolverHandle_t handle, int n, const aclFloatComplex* const A[],
                                                  int lda, aclFloatComplex* const Ainv[], int lda_inv, int* info,
                                                  int batchSize)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    int arg = 0;
    if (n < 0 || n > MatinvLimit(handle))
        arg = 1;
    else if (lda < std::max(1, n))
        arg = 3;
    else if (lda_inv < std::max(1, n))
        arg = 5;
    else if (batchSize < 0 || batchSize > kMaxBatchMatinv)
        arg = 7;
    if (arg) return Invalid(handle, info, batchSize < 0 ? 1 : (batchSize > 0 ? batchSize : 0), arg);
    if (batchSize == 0) return ACLSOLVER_STATUS_SUCCESS;
    if (!IsDevice(info)) return ACLSOLVER_STATUS_INVALID_VALUE;
    hipStream_t s = Stream(handle);
    if (n == 0)
    {
        FillInfo(info, batchSize, 0, s);
        return Launched();
    }
    if (!IsDevice(A)) return Invalid(handle, info, batchSize, 2);
    if (!IsDevice(Ainv)) return Invalid(handle, info, batchSize, 4);
    std::vector<uintptr_t> in, out;
    int bad = CheckArrays(handle, n, A, lda, Ainv, lda_inv, info, batchSize, nullptr, 2, 4, 4, &in, &out);
    if (bad < 0) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    if (bad) return Invalid(handle, info, batchSize, bad);
    const auto* A2 = reinterpret_cast<const float2* const*>(A);
    auto* C2 = reinterpret_cast<float2* const*>(Ainv);
    if (n <= 64)
    {
        if (CMatinvBatched(n, A2, nullptr, lda, 0, C2, nullptr, lda_inv, 0, info, batchSize, s) != hipSuccess)
            return ACLSOLVER_STATUS_EXECUTION_FAILED;
        return Launched();
    }
    for (int b = 0; b < batchSize; ++b)
    {
        auto st = InverseByFactor(handle, n, reinterpret_cast<const float2*>(in[b]), lda, nullptr, true,
                                  reinterpret_cast<float2*>(out[b]), lda_inv, info + b);
        if (st != ACLSOLVER_STATUS_SUCCESS) return st;
    }
    return Launched();
}
