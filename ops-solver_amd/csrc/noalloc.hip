// Device column-major getrs that allocates nothing (EXTENSION, no reference counterpart).
//
// aclsolverSgetrsDevice stages the right-hand sides, the row permutation and the chained
// triangular solve's publication words in handle scratch, which it grows with hipMalloc.
// Callers that may not allocate (code under stream capture, tight serving loops) call
// aclsolverSgetrsDeviceNoAlloc instead: the same kernels, with that staging in a fixed
// file-scope device array of kNoAllocScratchBytes. Each GPU holds its own copy of the array
// for the life of the process; its device address is looked up once per device with
// hipGetSymbolAddress (at handle creation, or at the first call on a device) and only that
// address is ever handed to kernels. Every call on one GPU shares the array: calls that
// overlap on one GPU (two streams, two threads) race on it, and each call re-initialises
// what must start at zero (the chained solve's publication words, by a fill kernel).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "internal.h"

namespace ops
{
namespace
{
constexpr size_t kNoAllocScratchBytes = ACLSOLVER_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_mutex;
void* g_ptr[kMaxDevices] = {};
}  // namespace

// Looks up (once per device) the device address of this device's copy of `scratch`; the
// device must be current. Returns nullptr when the lookup fails.
void* NoAllocScratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_ptr[device] == nullptr)
    {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess)
        {
            (void)hipGetLastError();
            return nullptr;
        }
        g_ptr[device] = p;
    }
    return g_ptr[device];
}

}  // namespace ops

using namespace ops;

OPS_API aclsolverStatus_t aclsolverSgetrsDeviceNoAlloc(aclsolverHandle_t handle, int trans, int n, int nrhs,
                                                       const float* dA, int lda, const int* devIpiv, float* dB,
                                                       int ldb, int* devInfo)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    int arg = 0;
    if (trans < 0 || trans > 1)
        arg = 1;
    else if (n < 0 || n > kEngineMaxOrder || RealGetrsWorkspaceBytes(n, 1) > kNoAllocScratchBytes)
        arg = 2;
    else if (nrhs < 0 || RealGetrsWorkspaceBytes(n, nrhs) > kNoAllocScratchBytes)
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
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess)
    {
        (void)hipGetLastError();
        return ACLSOLVER_STATUS_NOT_INITIALIZED;
    }
    void* work = NoAllocScratch(device);
    if (work == nullptr) return ACLSOLVER_STATUS_EXECUTION_FAILED;
    hipStream_t s = Stream(handle);
    FillInfo(devInfo, 1, 0, s);
    if (n == 0 || nrhs == 0) return Launched();
    if (RealGetrs(trans, n, nrhs, dA, lda, devIpiv, dB, ldb, work, s, devInfo) != hipSuccess)
        return ACLSOLVER_STATUS_EXECUTION_FAILED;
    return Launched();
}
