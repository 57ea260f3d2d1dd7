// Operator adapters (protocol version 2, include/solver_adapter_v2.h) for sgetrf, sgetri,
// sgetri_batched and smatinv_batched: one small library per operator, selected at compile
// time by OPS_ADAPTER_OP (0 sgetrf, 1 sgetri, 2 sgetri_batched, 3 smatinv_batched).
//
// The protocol (reference test/sgetrf/README.md:236-245, test/sgetrf/solver_adapter.h:25-69):
// the harness owns handle and stream; transport buffers are compact row-major host arrays;
// create validates the case and allocates; reset restores inputs and fills outputs, info
// and pivots with sentinels; execute calls the public operator once; readback (after the
// harness synchronized the stream) returns outputs, info and pivots; destroy releases a
// null, partial or complete state. Here the layout conversions run on the GPU: the transport
// is copied to the device once and a kernel writes the operator's padded column-major
// matrices (padding rows hold NaN, so an operator that read them would show it), and back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <limits>
#include <new>
#include <vector>

#include "cann_ops_solver.h"
#include "solver_adapter_v2.h"

#ifndef OPS_ADAPTER_OP
#error "OPS_ADAPTER_OP selects the operator: 0 sgetrf, 1 sgetri, 2 sgetri_batched, 3 smatinv_batched"
#endif

namespace
{
enum class Op
{
    Getrf = 0,
    Getri = 1,
    GetriBatched = 2,
    MatinvBatched = 3
};
constexpr Op kOp = static_cast<Op>(OPS_ADAPTER_OP);
constexpr bool kBatched = kOp == Op::GetriBatched || kOp == Op::MatinvBatched;
constexpr bool kFactor = kOp == Op::Getrf;
constexpr bool kInputLu = kOp == Op::Getri || kOp == Op::GetriBatched;
constexpr bool kInPlace = !kBatched;

enum Rc
{
    kOk = 0,
    kInvalid = 1,
    kAllocFailed = 2,
    kRuntime = 3,
    kQueryFailed = 4
};

// Largest orders this library's operators accept (extended limits).
constexpr int64_t kMaxRows = 16384;
constexpr int64_t kMaxOrder = kOp == Op::MatinvBatched ? 64 : 16384;

struct DeviceBuffer
{
    void* p = nullptr;
    size_t bytes = 0;
    bool Allocate(size_t b)
    {
        bytes = b;
        if (b == 0) return true;
        if (hipMalloc(&p, b) == hipSuccess) return true;
        (void)hipGetLastError();
        p = nullptr;
        return false;
    }
    bool Free()
    {
        if (p != nullptr && hipFree(p) != hipSuccess) return false;
        p = nullptr;
        bytes = 0;
        return true;
    }
};

struct State
{
    aclsolverHandle_t handle = nullptr;
    hipStream_t stream = nullptr;
    int m = 0, n = 0, lda = 0, ldc = 0, batch = 0, lwork = 0;
    size_t cells = 0;           // m * n per matrix
    size_t transportBytes = 0;  // batch * m * n floats
    size_t inPivots = 0, outPivots = 0, infoCount = 0;
    bool ready = false, reset = false, pending = false;
    DeviceBuffer a, c, work, piv, info, aArr, cArr, stage;
};

bool Mul(size_t x, size_t y, size_t* r)
{
    if (y != 0 && x > std::numeric_limits<size_t>::max() / y) return false;
    *r = x * y;
    return true;
}
bool FitsInt(int64_t v) { return v >= 0 && v <= INT_MAX; }

bool Validate(const SolverAdapterCaseV2* c, State* s)
{
    if (c == nullptr || c->base.version != SOLVER_ADAPTER_VERSION_V2 || c->base.struct_size != sizeof(*c)) return false;
    const SolverAdapterCase& b = c->base;
    if (!FitsInt(b.n) || !FitsInt(b.batch) || !FitsInt(b.lda) || b.b_bytes != 0 || c->in_tau_count != 0 ||
        c->out_tau_count != 0 || (!kBatched && b.batch != 1))
        return false;
    s->n = static_cast<int>(b.n);
    s->batch = static_cast<int>(b.batch);
    s->lda = static_cast<int>(b.lda);
    if (kFactor)
    {
        if (!FitsInt(c->m) || c->m > kMaxRows) return false;
        s->m = static_cast<int>(c->m);
    }
    else
    {
        if ((c->m != 0 && c->m != b.n) || b.n > kMaxOrder) return false;
        s->m = s->n;
    }
    if (s->lda < std::max(1, s->m)) return false;
    s->ldc = s->lda;
    if (kBatched)
    {
        if (!FitsInt(c->ldc) || c->ldc < std::max(1, s->n)) return false;
        s->ldc = static_cast<int>(c->ldc);
    }
    size_t total = 0, pivots = 0;
    if (!Mul(static_cast<size_t>(s->m), static_cast<size_t>(s->n), &s->cells) ||
        !Mul(s->cells, static_cast<size_t>(s->batch), &total) || !Mul(total, sizeof(float), &s->transportBytes) ||
        !Mul(static_cast<size_t>(std::min(s->m, s->n)), static_cast<size_t>(s->batch), &pivots))
        return false;
    s->infoCount = kBatched ? static_cast<size_t>(s->batch) : 1;
    if (b.info_count != s->infoCount || b.a_bytes != s->transportBytes || b.out_bytes != s->transportBytes)
        return false;
    // Pivot counts: 0 selects the operator's no-pivot form; otherwise exactly min(m,n)*batch.
    if (kFactor && !(c->in_ipiv_count == 0 && (c->out_ipiv_count == 0 || c->out_ipiv_count == pivots))) return false;
    if (kInputLu && !(c->out_ipiv_count == 0 && (c->in_ipiv_count == 0 || c->in_ipiv_count == pivots))) return false;
    if (kOp == Op::MatinvBatched && (c->in_ipiv_count != 0 || c->out_ipiv_count != 0)) return false;
    s->inPivots = c->in_ipiv_count;
    s->outPivots = c->out_ipiv_count;
    return true;
}

aclsolverStatus_t QueryWorkspace(State* s)
{
    if constexpr (kFactor) return aclsolverSgetrf_bufferSize(s->handle, s->m, s->n, nullptr, s->lda, &s->lwork);
    if constexpr (kOp == Op::Getri) return aclsolverSgetri_bufferSize(s->handle, s->n, nullptr, s->lda, &s->lwork);
    s->lwork = 0;
    return ACLSOLVER_STATUS_SUCCESS;
}

int Allocate(State* s)
{
    if (QueryWorkspace(s) != ACLSOLVER_STATUS_SUCCESS || s->lwork < 0) return kQueryFailed;
    size_t aCells = 0, cCells = 0, aBytes = 0, cBytes = 0;
    if (!Mul(static_cast<size_t>(s->lda) * s->n, s->batch, &aCells) || !Mul(aCells, sizeof(float), &aBytes) ||
        !Mul(static_cast<size_t>(s->ldc) * s->n, s->batch, &cCells) || !Mul(cCells, sizeof(float), &cBytes))
        return kAllocFailed;
    const size_t pivBytes = std::max(s->inPivots, s->outPivots) * sizeof(int32_t);
    if (!s->a.Allocate(aBytes) || !s->c.Allocate(kInPlace ? 0 : cBytes) ||
        !s->work.Allocate(static_cast<size_t>(s->lwork) * sizeof(float)) || !s->piv.Allocate(pivBytes) ||
        !s->info.Allocate(s->infoCount * sizeof(int32_t)) || !s->stage.Allocate(s->transportBytes))
        return kAllocFailed;
    if (kBatched &&
        (!s->aArr.Allocate(static_cast<size_t>(s->batch) * sizeof(void*)) ||
         !s->cArr.Allocate(static_cast<size_t>(s->batch) * sizeof(void*))))
        return kAllocFailed;
    s->ready = true;
    return kOk;
}

// dst (batch x padded column-major, leading dimension ld) <- src (batch x compact row-major
// m x n); padding rows get `pad`. One thread per destination element.
__global__ void to_column_major_kernel(const float* __restrict__ src, float* __restrict__ dst, int m, int n, int ld,
                                       long long total, float pad)
{
    for (long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<long long>(gridDim.x) * blockDim.x)
    {
        const long long per = static_cast<long long>(ld) * n;
        const long long b = e / per, rem = e - b * per;
        const int c = static_cast<int>(rem / ld), r = static_cast<int>(rem - static_cast<long long>(c) * ld);
        dst[e] = r < m ? src[b * m * static_cast<long long>(n) + static_cast<long long>(r) * n + c] : pad;
    }
}

// dst (compact row-major) <- src (padded column-major). One thread per transport element.
__global__ void to_row_major_kernel(const float* __restrict__ src, float* __restrict__ dst, int m, int n, int ld,
                                    long long total)
{
    for (long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<long long>(gridDim.x) * blockDim.x)
    {
        const long long per = static_cast<long long>(m) * n;
        const long long b = e / per, rem = e - b * per;
        const int r = static_cast<int>(rem / n), c = static_cast<int>(rem - static_cast<long long>(r) * n);
        dst[e] = src[b * static_cast<long long>(ld) * n + static_cast<long long>(c) * ld + r];
    }
}

dim3 Grid(long long total)
{
    const long long blocks = (total + 255) / 256;
    return dim3(static_cast<unsigned>(std::min<long long>(std::max<long long>(blocks, 1), 1 << 20)));
}

bool SyncPending(State* s)
{
    if (!s->pending) return true;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return false;
    s->pending = false;
    return true;
}

bool Reset(State* s, const void* a, const int32_t* ipiv, size_t ipivCount)
{
    hipStream_t st = s->stream;
    const long long aTotal = static_cast<long long>(s->lda) * s->n * s->batch;
    if (s->transportBytes != 0)
    {
        if (hipMemcpyAsync(s->stage.p, a, s->transportBytes, hipMemcpyHostToDevice, st) != hipSuccess) return false;
        hipLaunchKernelGGL(to_column_major_kernel, Grid(aTotal), dim3(256), 0, st, static_cast<const float*>(s->stage.p),
                           static_cast<float*>(s->a.p), s->m, s->n, s->lda, aTotal, __builtin_nanf(""));
        if (hipGetLastError() != hipSuccess) return false;
    }
    // Sentinels expose outputs the operator did not write.
    if ((s->c.bytes && hipMemsetAsync(s->c.p, 0xA5, s->c.bytes, st) != hipSuccess) ||
        (s->work.bytes && hipMemsetAsync(s->work.p, 0, s->work.bytes, st) != hipSuccess) ||
        (s->info.bytes && hipMemsetAsync(s->info.p, 0xA5, s->info.bytes, st) != hipSuccess) ||
        (s->piv.bytes && hipMemsetAsync(s->piv.p, 0xA5, s->piv.bytes, st) != hipSuccess) ||
        (ipivCount && hipMemcpyAsync(s->piv.p, ipiv, ipivCount * sizeof(int32_t), hipMemcpyHostToDevice, st) !=
                          hipSuccess))
        return false;
    if (kBatched && s->batch > 0)
    {
        std::vector<const void*> pa(static_cast<size_t>(s->batch)), pc(static_cast<size_t>(s->batch));
        for (int b = 0; b < s->batch; ++b)
        {
            pa[b] = static_cast<const char*>(s->a.p) + static_cast<size_t>(b) * s->lda * s->n * sizeof(float);
            pc[b] = static_cast<const char*>(s->c.p) + static_cast<size_t>(b) * s->ldc * s->n * sizeof(float);
        }
        if (hipMemcpyAsync(s->aArr.p, pa.data(), s->aArr.bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(s->cArr.p, pc.data(), s->cArr.bytes, hipMemcpyHostToDevice, st) != hipSuccess)
            return false;
    }
    return hipStreamSynchronize(st) == hipSuccess;
}

aclsolverStatus_t Execute(State* s)
{
    float* a = static_cast<float*>(s->a.p);
    int* piv = static_cast<int*>(s->piv.p);
    int* info = static_cast<int*>(s->info.p);
    if constexpr (kOp == Op::Getrf)
        return aclsolverSgetrf(s->handle, s->m, s->n, a, s->lda, static_cast<float*>(s->work.p),
                               s->outPivots ? piv : nullptr, info);
    if constexpr (kOp == Op::Getri)
        return aclsolverSgetri(s->handle, s->n, a, s->lda, s->inPivots ? piv : nullptr,
                               static_cast<float*>(s->work.p), s->lwork, info);
    if constexpr (kOp == Op::GetriBatched)
        return aclsolverSgetriBatched(s->handle, s->n, static_cast<const float* const*>(s->aArr.p), s->lda,
                                      s->inPivots ? piv : nullptr, static_cast<float* const*>(s->cArr.p), s->ldc,
                                      info, s->batch);
    return aclsolverSmatinvBatched(s->handle, s->n, static_cast<const float* const*>(s->aArr.p), s->lda,
                                   static_cast<float* const*>(s->cArr.p), s->ldc, info, s->batch);
}

bool Readback(State* s, void* out, int32_t* info, int32_t* ipiv)
{
    hipStream_t st = s->stream;
    const int ld = kInPlace ? s->lda : s->ldc;
    const float* src = static_cast<const float*>(kInPlace ? s->a.p : s->c.p);
    const long long total = static_cast<long long>(s->cells) * s->batch;
    if (total > 0)
    {
        hipLaunchKernelGGL(to_row_major_kernel, Grid(total), dim3(256), 0, st, src, static_cast<float*>(s->stage.p),
                           s->m, s->n, ld, total);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, s->stage.p, s->transportBytes, hipMemcpyDeviceToHost, st) != hipSuccess)
            return false;
    }
    if ((s->infoCount &&
         hipMemcpyAsync(info, s->info.p, s->infoCount * sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess) ||
        (s->outPivots &&
         hipMemcpyAsync(ipiv, s->piv.p, s->outPivots * sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess))
        return false;
    return hipStreamSynchronize(st) == hipSuccess;
}

const char* OpName()
{
    return kOp == Op::Getrf ? "sgetrf" : kOp == Op::Getri ? "sgetri" : kOp == Op::GetriBatched ? "sgetribatched"
                                                                                                : "smatinvbatched";
}
const void* OpSymbol()
{
    if constexpr (kOp == Op::Getrf) return reinterpret_cast<const void*>(&aclsolverSgetrf);
    if constexpr (kOp == Op::Getri) return reinterpret_cast<const void*>(&aclsolverSgetri);
    if constexpr (kOp == Op::GetriBatched) return reinterpret_cast<const void*>(&aclsolverSgetriBatched);
    return reinterpret_cast<const void*>(&aclsolverSmatinvBatched);
}
}  // namespace

#define ADAPTER_API extern "C" __attribute__((visibility("default")))

ADAPTER_API const SolverAdapterDesc* solver_adapter_descriptor(void)
{
    static const SolverAdapterDesc d{SOLVER_ADAPTER_VERSION_V2, sizeof(SolverAdapterDesc), OpName(), "float32",
                                     "device", "column_major", kBatched ? "array" : "scalar", OpSymbol()};
    return &d;
}

ADAPTER_API int solver_adapter_create_v2(const SolverAdapterCaseV2* c, void* handle, void* stream, void** state)
{
    if (state == nullptr) return kInvalid;
    *state = nullptr;
    if (handle == nullptr) return kInvalid;
    State* s = new (std::nothrow) State;
    if (s == nullptr) return kAllocFailed;
    *state = s;  // published first: destroy releases whatever follows
    s->handle = static_cast<aclsolverHandle_t>(handle);
    s->stream = static_cast<hipStream_t>(stream);
    if (!Validate(c, s)) return kInvalid;
    aclrtStream bound = nullptr;
    if (aclsolverGetStream(s->handle, &bound) != ACLSOLVER_STATUS_SUCCESS || bound != stream) return kInvalid;
    return Allocate(s);
}

ADAPTER_API int solver_adapter_reset_v2(void* state, const void* a, size_t a_bytes, const void* /*b*/, size_t b_bytes,
                                        const int32_t* ipiv, size_t ipiv_count, const void* /*tau*/, size_t tau_bytes)
{
    State* s = static_cast<State*>(state);
    if (s == nullptr || !s->ready || a_bytes != s->transportBytes || (a_bytes != 0 && a == nullptr) || b_bytes != 0 ||
        ipiv_count != s->inPivots || (ipiv_count != 0 && ipiv == nullptr) || tau_bytes != 0)
        return kInvalid;
    s->reset = false;
    if (!SyncPending(s)) return kRuntime;
    if (!Reset(s, a, ipiv, ipiv_count)) return kRuntime;
    s->reset = true;
    return kOk;
}

ADAPTER_API int solver_adapter_execute(void* state, int* dut_status)
{
    State* s = static_cast<State*>(state);
    if (s == nullptr || !s->ready || !s->reset || dut_status == nullptr) return kInvalid;
    s->pending = true;  // the call may enqueue work even when it reports an error
    *dut_status = static_cast<int>(Execute(s));
    return kOk;
}

ADAPTER_API int solver_adapter_readback_v2(void* state, void* out, size_t out_bytes, int32_t* info, size_t info_count,
                                           int32_t* ipiv, size_t ipiv_count, void* /*tau*/, size_t tau_bytes)
{
    State* s = static_cast<State*>(state);
    if (s == nullptr || !s->ready || !s->reset || out_bytes != s->transportBytes || (out_bytes != 0 && out == nullptr) ||
        info_count != s->infoCount || (info_count != 0 && info == nullptr) || ipiv_count != s->outPivots ||
        (ipiv_count != 0 && ipiv == nullptr) || tau_bytes != 0)
        return kInvalid;
    return Readback(s, out, info, ipiv) ? kOk : kRuntime;
}

ADAPTER_API int solver_adapter_destroy(void* state)
{
    State* s = static_cast<State*>(state);
    if (s == nullptr) return kOk;
    if (!SyncPending(s)) return kRuntime;  // keep everything for a retry: work may be in flight
    bool freed = true;
    DeviceBuffer* bufs[] = {&s->stage, &s->cArr, &s->aArr, &s->info, &s->piv, &s->work, &s->c, &s->a};
    for (DeviceBuffer* b : bufs) freed = b->Free() && freed;
    if (!freed)
    {
        s->ready = false;
        return kRuntime;
    }
    delete s;
    return kOk;
}
