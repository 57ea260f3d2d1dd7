// Handle lifetime, stream binding, pointer classification, info fill and the finite scan.
// Reference: src/utils/aclsolver_auxiliary.cpp:167-262 (Create/Destroy/SetStream/GetStream),
// src/utils/real_lu_runtime.h:43-78 (Device(), Invalid(), Launched()),
// src/utils/real_lu_common.h:90-132 + src/sgetrf/sgetrf_kernel.cpp:98-152 (finite check).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

#include <new>
#include <vector>

#include "internal.h"

namespace ops
{

namespace
{
unsigned g_spinLimit = ACLSOLVER_DEFAULT_SPIN_LIMIT;

__global__ void fold_wait_error_kernel(const int* err, int* info)
{
    if (*err != 0) *info = kInfoWaitTimeout;
}
}  // namespace

unsigned SpinLimit() { return g_spinLimit; }

void FoldWaitError(const int* err, int* info, hipStream_t stream)
{
    if (err == nullptr || info == nullptr) return;
    hipLaunchKernelGGL(fold_wait_error_kernel, dim3(1), dim3(1), 0, stream, err, info);
}

bool IsDevice(const void* p)
{
    if (p == nullptr) return false;
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, p) != hipSuccess)
    {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

bool IsHostOnly(const void* p)
{
    if (p == nullptr) return false;
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, p) != hipSuccess)
    {
        (void)hipGetLastError();
        return true;  // unregistered host memory
    }
    return attr.type == hipMemoryTypeHost || attr.type == hipMemoryTypeUnregistered;
}

void* HandleScratch(aclsolverHandle_t h, size_t bytes)
{
    Handle* s = H(h);
    if (bytes == 0) bytes = 256;
    if (s->scratch != nullptr && s->scratchBytes >= bytes) return s->scratch;
    if (s->scratch != nullptr)
    {
        // Earlier work on the stream may still use the old block.
        if (hipStreamSynchronize(s->stream) != hipSuccess) return nullptr;
        (void)hipFree(s->scratch);
        s->scratch = nullptr;
        s->scratchBytes = 0;
    }
    size_t grow = bytes + bytes / 4;
    if (hipMalloc(&s->scratch, grow) != hipSuccess)
    {
        (void)hipGetLastError();
        s->scratch = nullptr;
        return nullptr;
    }
    s->scratchBytes = grow;
    return s->scratch;
}

__global__ void fill_info_kernel(int* info, int count, int value)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) info[i] = value;
}

void FillInfo(int* info, int count, int value, hipStream_t stream)
{
    if (count <= 0 || info == nullptr) return;
    int blocks = (count + 255) / 256;
    hipLaunchKernelGGL(fill_info_kernel, dim3(blocks), dim3(256), 0, stream, info, count, value);
}

aclsolverStatus_t Invalid(aclsolverHandle_t h, int* info, int count, int argument)
{
    if (count > 0 && info != nullptr && reinterpret_cast<uintptr_t>(info) % alignof(int) == 0 && IsDevice(info))
        FillInfo(info, count, -argument, Stream(h));
    return ACLSOLVER_STATUS_INVALID_VALUE;
}

aclsolverStatus_t Launched()
{
    return hipGetLastError() == hipSuccess ? ACLSOLVER_STATUS_SUCCESS : ACLSOLVER_STATUS_EXECUTION_FAILED;
}

// flags[0] != 0 when any scanned element is NaN or Inf. One pass over the matrix.
__global__ void finite_kernel(const float* a, int rows, int cols, int lda, int64_t batchStride, int batch, int* flag)
{
    int64_t total = static_cast<int64_t>(rows) * cols * batch;
    int bad = 0;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
    {
        int64_t b = e / (static_cast<int64_t>(rows) * cols);
        int64_t rc = e - b * rows * cols;
        int c = static_cast<int>(rc / rows), r = static_cast<int>(rc - static_cast<int64_t>(c) * rows);
        float v = a[b * batchStride + static_cast<int64_t>(c) * lda + r];
        bad |= !(v - v == 0.0f);
    }
    if (__any(bad) && (threadIdx.x % warpSize) == 0) atomicOr(flag, 1);
}

__global__ void finite_arrays_kernel(const float* const* arrays, int rows, int cols, int lda, int batch, int* flag)
{
    int bad = 0;
    for (int b = blockIdx.y; b < batch; b += gridDim.y)
    {
        const float* a = arrays[b];
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < rows * cols; e += gridDim.x * blockDim.x)
        {
            int c = e / rows, r = e - c * rows;
            float v = a[static_cast<int64_t>(c) * lda + r];
            bad |= !(v - v == 0.0f);
        }
    }
    if (__any(bad) && (threadIdx.x % warpSize) == 0) atomicOr(flag, 1);
}

// Small matrices stored contiguously (lda == rows): each workgroup checks 4 matrices at a
// time with their loads in flight together, 16-byte loads where the matrix allows them.
__global__ __launch_bounds__(256) void finite_arrays_small_kernel(const float* const* arrays, int per, int batch,
                                                                  int* flag)
{
    constexpr int U = 4;
    int bad = 0;
    for (int b0 = blockIdx.x * U; b0 < batch; b0 += gridDim.x * U)
    {
        const float* p[U];
#pragma unroll
        for (int u = 0; u < U; ++u) p[u] = b0 + u < batch ? arrays[b0 + u] : nullptr;
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            if (p[u] == nullptr) continue;  // uniform
            if (per % 4 == 0 && (reinterpret_cast<uintptr_t>(p[u]) & 15) == 0)
            {
                const float4* q = reinterpret_cast<const float4*>(p[u]);
                for (int e = threadIdx.x; e < per / 4; e += blockDim.x)
                {
                    const float4 v = q[e];
                    bad |= !(v.x - v.x == 0.0f) | !(v.y - v.y == 0.0f) | !(v.z - v.z == 0.0f) | !(v.w - v.w == 0.0f);
                }
            }
            else
                for (int e = threadIdx.x; e < per; e += blockDim.x)
                {
                    const float v = p[u][e];
                    bad |= !(v - v == 0.0f);
                }
        }
    }
    if (__any(bad) && (threadIdx.x % warpSize) == 0) atomicOr(flag, 1);
}

static bool ReadFlag(int* dflag, int* hflag, hipStream_t stream, bool* ok)
{
    *hflag = 0;
    bool good = hipMemcpyAsync(hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, stream) == hipSuccess &&
                hipStreamSynchronize(stream) == hipSuccess;
    *ok = *hflag == 0;
    return good;
}

bool AllFinite(aclsolverHandle_t h, const float* a, int rows, int cols, int lda, int64_t batchStride, int batch,
               bool* ok)
{
    *ok = true;
    if (rows <= 0 || cols <= 0 || batch <= 0) return true;
    hipStream_t stream = Stream(h);
    int* dflag = H(h)->dflag;
    FillInfo(dflag, 1, 0, stream);
    int64_t total = static_cast<int64_t>(rows) * cols * batch;
    int blocks = static_cast<int>(std::min<int64_t>((total + 255) / 256, 4096));
    hipLaunchKernelGGL(finite_kernel, dim3(blocks), dim3(256), 0, stream, a, rows, cols, lda, batchStride, batch,
                       dflag);
    return ReadFlag(dflag, H(h)->hflag, stream, ok);
}

bool AllFiniteArrays(aclsolverHandle_t h, const float* const* arrays, int rows, int cols, int lda, int batch,
                     bool* ok)
{
    *ok = true;
    if (rows <= 0 || cols <= 0 || batch <= 0) return true;
    hipStream_t stream = Stream(h);
    int* dflag = H(h)->dflag;
    FillInfo(dflag, 1, 0, stream);
    if (lda == rows && rows * cols <= 4096)
    {
        const int blocks = std::min((batch + 3) / 4, 65536);
        hipLaunchKernelGGL(finite_arrays_small_kernel, dim3(blocks), dim3(256), 0, stream, arrays, rows * cols, batch,
                           dflag);
        return ReadFlag(dflag, H(h)->hflag, stream, ok);
    }
    int per = (rows * cols + 255) / 256;
    dim3 grid(std::min(per, 64), std::min(batch, 4096));
    hipLaunchKernelGGL(finite_arrays_kernel, grid, dim3(256), 0, stream, arrays, rows, cols, lda, batch, dflag);
    return ReadFlag(dflag, H(h)->hflag, stream, ok);
}

}  // namespace ops

using namespace ops;

// Reference src/utils/aclsolver_auxiliary.cpp:167-197: *handle must be null on entry and
// a current device (context) is required.
OPS_API aclsolverStatus_t aclsolverCreate(aclsolverHandle_t* handle)
{
    if (handle == nullptr) return ACLSOLVER_STATUS_HANDLE_IS_NULLPTR;
    if (*handle != nullptr) return ACLSOLVER_STATUS_INVALID_VALUE;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess)
    {
        (void)hipGetLastError();
        return ACLSOLVER_STATUS_NOT_INITIALIZED;
    }
    Handle* h = new (std::nothrow) Handle();
    if (h == nullptr) return ACLSOLVER_STATUS_ALLOC_FAILED;
    h->device = device;
    if (hipMalloc(&h->dflag, 256) != hipSuccess || hipHostMalloc(&h->hflag, 256, hipHostMallocDefault) != hipSuccess)
    {
        (void)hipGetLastError();
        if (h->dflag) (void)hipFree(h->dflag);
        delete h;
        return ACLSOLVER_STATUS_ALLOC_FAILED;
    }
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, device) == hipSuccess)
    {
        h->numCUs = prop.multiProcessorCount;
        // gcnArchName is e.g. "gfx950:sramecc+:xnack-". OPS_SOLVER_ARCH_OVERRIDE replaces the
        // reported name (testing the mismatch path on a gfx950 device).
        const char* over = std::getenv("OPS_SOLVER_ARCH_OVERRIDE");
        const char* arch = (over != nullptr && over[0] != 0) ? over : prop.gcnArchName;
        h->archOk = std::strncmp(arch, "gfx950", 6) == 0 && (arch[6] == 0 || arch[6] == ':');
    }
    // Device address of the no-allocation entry point's scratch array, looked up once per
    // device while it is current (failure here only disables that entry point).
    if (h->archOk) (void)NoAllocScratch(device);
    // Placement of the tall panel leaves' workgroups (one probe launch per device).
    if (h->archOk) LeafSpreadProbe(device);
    *handle = h;
    return ACLSOLVER_STATUS_SUCCESS;
}

// Reference :199-232: drains the stream before releasing handle-owned scratch.
namespace ops
{
const LookAhead* HandleLookAhead(aclsolverHandle_t handle)
{
    Handle* h = H(handle);
    if (h->laReady) return h->la.side ? &h->la : nullptr;
    h->laReady = true;
    // OPS_SOLVER_LOOKAHEAD=0 keeps every factorization on the caller's stream (tuning aid).
    const char* env = std::getenv("OPS_SOLVER_LOOKAHEAD");
    if (env != nullptr && std::atoi(env) == 0) return nullptr;
    LookAhead la;
    // OPS_SOLVER_PANEL_CUS=<k> (0 < k < CUs): reserve about k CUs for the panel path and give
    // the trailing updates the rest (two CU-masked streams, default priority). The reserved
    // set, (c / 8 + c) % step == 0, takes the same share of every XCD whether CU numbers run
    // across the XCDs (c % 8) or through them (c / 32).
    const char* pcu = std::getenv("OPS_SOLVER_PANEL_CUS");
    const int panelCUs = pcu != nullptr ? std::atoi(pcu) : 0;
    bool ok = true;
    if (panelCUs > 0 && panelCUs < h->numCUs)
    {
        const int words = (h->numCUs + 31) / 32, step = h->numCUs / panelCUs;
        std::vector<uint32_t> pmask(words, 0u), smask(words, 0u);
        for (int c = 0; c < h->numCUs; ++c)
        {
            const bool reserved = (c / 8 + c) % step == 0;
            (reserved ? pmask : smask)[c / 32] |= 1u << (c % 32);
        }
        ok = hipExtStreamCreateWithCUMask(&la.panelStream, static_cast<uint32_t>(words), pmask.data()) == hipSuccess &&
             hipExtStreamCreateWithCUMask(&la.side, static_cast<uint32_t>(words), smask.data()) == hipSuccess &&
             hipEventCreateWithFlags(&la.join, hipEventDisableTiming) == hipSuccess;
    }
    else
        ok = hipStreamCreateWithFlags(&la.side, hipStreamNonBlocking) == hipSuccess;
    hipEvent_t* evs[5] = {&la.fork, &la.panel[0], &la.panel[1], &la.bulk[0], &la.bulk[1]};
    for (hipEvent_t* e : evs)
        ok = ok && hipEventCreateWithFlags(e, hipEventDisableTiming) == hipSuccess;
    if (!ok)
    {
        (void)hipGetLastError();
        for (hipEvent_t* e : evs)
            if (*e) (void)hipEventDestroy(*e);
        if (la.join) (void)hipEventDestroy(la.join);
        if (la.side) (void)hipStreamDestroy(la.side);
        if (la.panelStream) (void)hipStreamDestroy(la.panelStream);
        return nullptr;
    }
    h->la = la;
    return &h->la;
}
}  // namespace ops

OPS_API aclsolverStatus_t aclsolverDestroy(aclsolverHandle_t handle)
{
    if (handle == nullptr) return ACLSOLVER_STATUS_HANDLE_IS_NULLPTR;
    Handle* h = H(handle);
    if (h->la.side != nullptr)
    {
        if (hipStreamSynchronize(h->la.side) != hipSuccess) return ACLSOLVER_STATUS_EXECUTION_FAILED;
        if (h->la.panelStream != nullptr && hipStreamSynchronize(h->la.panelStream) != hipSuccess)
            return ACLSOLVER_STATUS_EXECUTION_FAILED;
        hipEvent_t evs[6] = {h->la.fork, h->la.panel[0], h->la.panel[1], h->la.bulk[0], h->la.bulk[1], h->la.join};
        for (hipEvent_t e : evs)
            if (e) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(h->la.side);
        if (h->la.panelStream) (void)hipStreamDestroy(h->la.panelStream);
        h->la = LookAhead{};
    }
    if (h->scratch != nullptr)
    {
        if (hipStreamSynchronize(h->stream) != hipSuccess) return ACLSOLVER_STATUS_EXECUTION_FAILED;
        if (hipFree(h->scratch) != hipSuccess) return ACLSOLVER_STATUS_EXECUTION_FAILED;
        h->scratch = nullptr;
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    if (h->dflag) (void)hipFree(h->dflag);
    if (h->hflag) (void)hipHostFree(h->hflag);
    h->dflag = nullptr;
    h->hflag = nullptr;
    h->magic = 0;
    delete h;
    return ACLSOLVER_STATUS_SUCCESS;
}

OPS_API aclsolverStatus_t aclsolverSetStream(aclsolverHandle_t handle, aclrtStream stream)
{
    if (handle == nullptr) return ACLSOLVER_STATUS_HANDLE_IS_NULLPTR;
    H(handle)->stream = static_cast<hipStream_t>(stream);
    return ACLSOLVER_STATUS_SUCCESS;
}

OPS_API aclsolverStatus_t aclsolverGetStream(aclsolverHandle_t handle, aclrtStream* stream)
{
    if (handle == nullptr) return ACLSOLVER_STATUS_HANDLE_IS_NULLPTR;
    if (stream == nullptr) return ACLSOLVER_STATUS_INVALID_VALUE;
    *stream = H(handle)->stream;
    return ACLSOLVER_STATUS_SUCCESS;
}

OPS_API aclsolverStatus_t aclsolverSetStrictLimits(aclsolverHandle_t handle, int strict)
{
    if (handle == nullptr) return ACLSOLVER_STATUS_HANDLE_IS_NULLPTR;
    H(handle)->strict = strict != 0;
    return ACLSOLVER_STATUS_SUCCESS;
}

OPS_API aclsolverStatus_t aclsolverSetRefinement(aclsolverHandle_t handle, double tolerance, int maxOuter,
                                                 int restart)
{
    if (handle == nullptr) return ACLSOLVER_STATUS_HANDLE_IS_NULLPTR;
    if (!(tolerance > 0) || maxOuter < 1 || restart < 1 || restart > 64) return ACLSOLVER_STATUS_INVALID_VALUE;
    H(handle)->irTolerance = tolerance;
    H(handle)->irMaxOuter = maxOuter;
    H(handle)->irRestart = restart;
    return ACLSOLVER_STATUS_SUCCESS;
}

OPS_API aclsolverStatus_t aclsolverDebugSetSpinLimit(unsigned int polls)
{
    g_spinLimit = polls;
    return ACLSOLVER_STATUS_SUCCESS;
}
