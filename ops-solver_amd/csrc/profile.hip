// Optional in-library kernel timing: when enabled, selected launches are bracketed by HIP
// events on the stream they are issued to, so a caller can read the average duration and
// the algorithmic work of, e.g., every trailing-update GEMM of a factorization.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "internal.h"

namespace ops
{
namespace
{
struct Record
{
    hipEvent_t start, stop;
    double work;
    int kernel;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Record> g_records;
std::vector<hipEvent_t> g_pool;

hipEvent_t Take()
{
    if (!g_pool.empty())
    {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

bool ProfilingOn() { return g_on; }

ProfileScope::ProfileScope(ProfileKernel k, double work, hipStream_t stream) : kernel_(k), work_(work), stream_(stream)
{
    if (!g_on) return;
    std::lock_guard<std::mutex> lock(g_mu);
    start_ = Take();
    if (start_) (void)hipEventRecord(start_, stream_);
}

ProfileScope::~ProfileScope()
{
    if (!g_on || start_ == nullptr) return;
    std::lock_guard<std::mutex> lock(g_mu);
    hipEvent_t stop = Take();
    if (stop == nullptr) return;
    (void)hipEventRecord(stop, stream_);
    g_records.push_back({start_, stop, work_, static_cast<int>(kernel_)});
}

}  // namespace ops

using namespace ops;

// EXTENSION (measurement): start collecting per-launch timings of the instrumented kernels.
OPS_API aclsolverStatus_t aclsolverProfileBegin(void)
{
    std::lock_guard<std::mutex> lock(g_mu);
    for (auto& r : g_records)
    {
        g_pool.push_back(r.start);
        g_pool.push_back(r.stop);
    }
    g_records.clear();
    g_on = true;
    return ACLSOLVER_STATUS_SUCCESS;
}

// EXTENSION (measurement): stop collecting; waits for the recorded events and reports, for
// kernel class `kernel` (0 = trailing-update GEMM of the LU), the number of launches, the
// summed algorithmic work (flops) and the summed duration in milliseconds.
OPS_API aclsolverStatus_t aclsolverProfileEnd(int kernel, int* launches, double* work, double* ms)
{
    std::lock_guard<std::mutex> lock(g_mu);
    g_on = false;
    int count = 0;
    double w = 0.0, t = 0.0;
    for (auto& r : g_records)
    {
        if (r.kernel != kernel) continue;
        if (hipEventSynchronize(r.stop) != hipSuccess) return ACLSOLVER_STATUS_EXECUTION_FAILED;
        float e = 0.0f;
        if (hipEventElapsedTime(&e, r.start, r.stop) != hipSuccess) return ACLSOLVER_STATUS_EXECUTION_FAILED;
        ++count;
        w += r.work;
        t += e;
    }
    if (launches) *launches = count;
    if (work) *work = w;
    if (ms) *ms = t;
    return ACLSOLVER_STATUS_SUCCESS;
}
