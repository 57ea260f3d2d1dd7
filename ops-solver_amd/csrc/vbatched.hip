// Uneven batched LU solve (getrf + getrs per matrix) in one persistent launch (EXTENSION;
// the reference has no batched getrf/getrs).
//
// Plan: matrices are binned by order into 64 / 128 / 256 row classes. A work item is a
// group of same-class matrices that fills one 1024-thread workgroup (16 x n<=64 with one
// thread per row, 4 x n<=128 with two, 1 x n<=256 with four), so every thread keeps 64
// matrix values in VGPRs; items are queued largest class first. A persistent grid pulls items from
// a device counter until the queue is empty, so every workgroup reaches the exit.
//
// Per matrix: every row lives in one thread's registers (row-per-thread, position-tracked
// partial pivoting exactly as the register leaf of getrf.hip: largest |a|, lowest position
// on ties, reciprocal scaling, fused rank-1 updates), the right-hand sides ride along
// with their rows, then a back substitution on the final positions. The column index k is
// a runtime value but wave-uniform, so each static-column operation sits behind a scalar
// branch instead of dynamic register indexing. Outputs: packed LU (rows stored at their
// final positions), 1-based ipiv, X, info (first zero pivot).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <climits>
#include <new>
#include <vector>

#include "internal.h"

namespace ops
{
namespace
{
constexpr int kNrhsMax = 4;  // right-hand sides per matrix carried in registers
constexpr int kWG = 1024;

struct Item
{
    int bin;     // 0: n<=64, 1: n<=128, 2: n<=256
    int first;   // index into the plan's matrix order
    int count;   // matrices in this item
};

// One matrix processed by G = NMAX * TPR threads. Thread t owns row r = t % NMAX and, when
// TPR = 2, the half h = t / NMAX of its columns (interleaved: column j = TPR * jj + h), so
// each thread keeps NMAX / TPR values in registers. The thread holding column k of a row
// computes that row's multiplier and shares it through LDS with the row's other thread.
template <int NMAX, int TPR>
__device__ void solve_group(int n, int nrhs, float* __restrict__ A, float* __restrict__ B, int* __restrict__ ipiv,
                            int* __restrict__ info, int t, bool live, float* s_prow, float* s_rhs, float* s_best,
                            int* s_bpos, float* s_l)
{
    constexpr int G = NMAX * TPR;
    constexpr int C = NMAX / TPR;
    constexpr bool WAVE = G == 64;
    constexpr int NWG = G / 64;
    auto sync = [&]()
    {
        if (WAVE)
        {
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
        else
            __syncthreads();
    };
    // NMAX is a multiple of 64, so the column half h is uniform across a wave: keeping it
    // in a scalar register turns every column test below into a scalar branch.
    const int r = t % NMAX, h = __builtin_amdgcn_readfirstlane(t / NMAX);
    const bool rowLive = live && r < n;
    const bool rhsOwner = h == 0;
    float a[C];
    float b[kNrhsMax];
#pragma unroll
    for (int jj = 0; jj < C; ++jj)
    {
        const int j = TPR * jj + h;
        a[jj] = (rowLive && j < n) ? A[static_cast<size_t>(j) * n + r] : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < kNrhsMax; ++q) b[q] = (rowLive && rhsOwner && q < nrhs) ? B[static_cast<size_t>(q) * n + r] : 0.0f;
    int pos = rowLive ? r : INT_MAX;
    float cur = a[0];  // this thread's value in column k (valid when it owns column k)
    int zero = 0;
    const int lane = threadIdx.x & 63;
    const int wig = t >> 6;
    const int steps = live ? n : 0;
    for (int k = 0; k < NMAX; ++k)
    {
        if (WAVE && k >= steps) break;  // wave-private groups stop independently
        const bool active = k < steps;
        const bool ownK = (k % TPR) == h;
        float best = (active && ownK && pos >= k && pos != INT_MAX) ? fabsf(cur) : -1.0f;
        int bp = (best >= 0.0f) ? pos : INT_MAX;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
        {
            const float ob = __shfl_xor(best, off);
            const int op = __shfl_xor(bp, off);
            if (ob > best || (ob == best && op < bp))
            {
                best = ob;
                bp = op;
            }
        }
        if (!WAVE)
        {
            if (lane == 0)
            {
                s_best[wig] = best;
                s_bpos[wig] = bp;
            }
            sync();
            best = s_best[0];
            bp = s_bpos[0];
#pragma unroll
            for (int w = 1; w < NWG; ++w)
            {
                const float ob = s_best[w];
                const int op = s_bpos[w];
                if (ob > best || (ob == best && op < bp))
                {
                    best = ob;
                    bp = op;
                }
            }
        }
        const int bpos = (bp == INT_MAX) ? k : bp;
        if (active)
        {
            if (pos == bpos)
            {
#pragma unroll
                for (int jj = 0; jj < C; ++jj) s_prow[TPR * jj + h] = a[jj];
                if (rhsOwner)
                {
#pragma unroll
                    for (int q = 0; q < kNrhsMax; ++q) s_rhs[q] = b[q];
                }
            }
            if (t == 0) ipiv[k] = bpos + 1;
            if (pos == bpos)
                pos = k;
            else if (pos == k)
                pos = bpos;
        }
        sync();
        float l = 0.0f;
        if (active)
        {
            const float d = s_prow[k];
            if (d == 0.0f && zero == 0) zero = k + 1;
            const bool below = pos > k && pos != INT_MAX && d != 0.0f;
            const float q = fabsf(d) >= FLT_MIN ? cur * (1.0f / d) : cur / d;
            if (ownK)
            {
                l = below ? q : 0.0f;
                if (TPR > 1) s_l[r] = l;
#pragma unroll
                for (int jj = 0; jj < C; ++jj)
                    if (TPR * jj + h == k && below) a[jj] = q;
            }
        }
        if (TPR > 1)
        {
            sync();
            if (active && !ownK) l = s_l[r];
        }
        if (active)
        {
            float next = 0.0f;
#pragma unroll
            for (int jj = 0; jj < C; ++jj)
            {
                const int j = TPR * jj + h;
                if (j > k)
                {
                    a[jj] = fmaf(-l, s_prow[j], a[jj]);
                    if (j == k + 1) next = a[jj];
                }
            }
            if (rhsOwner)
            {
#pragma unroll
                for (int qq = 0; qq < kNrhsMax; ++qq) b[qq] = fmaf(-l, s_rhs[qq], b[qq]);
            }
            cur = next;
        }
        sync();
    }
    // Back substitution over positions: the row at position p holds U(p, p:n).
    for (int k = NMAX - 1; k >= 0; --k)
    {
        if (WAVE && k >= steps) continue;
        const bool active = k < steps;
        const bool ownK = (k % TPR) == h;
        float ukk = 0.0f;
        if (active && ownK && pos != INT_MAX)
        {
#pragma unroll
            for (int jj = 0; jj < C; ++jj)
                if (TPR * jj + h == k) ukk = a[jj];
            s_l[r] = ukk;  // U(pos, k) of this row
        }
        sync();
        if (active && rhsOwner && pos == k)
        {
            const float dkk = s_l[r];
#pragma unroll
            for (int q = 0; q < kNrhsMax; ++q)
            {
                b[q] = b[q] / dkk;
                s_rhs[q] = b[q];
            }
        }
        sync();
        if (active && rhsOwner && pos < k)
        {
            const float u = s_l[r];
#pragma unroll
            for (int q = 0; q < kNrhsMax; ++q) b[q] = fmaf(-u, s_rhs[q], b[q]);
        }
        sync();
    }
    if (rowLive)
    {
#pragma unroll
        for (int jj = 0; jj < C; ++jj)
        {
            const int j = TPR * jj + h;
            if (j < n) A[static_cast<size_t>(j) * n + pos] = a[jj];
        }
        if (rhsOwner)
        {
#pragma unroll
            for (int q = 0; q < kNrhsMax; ++q)
                if (q < nrhs) B[static_cast<size_t>(q) * n + pos] = b[q];
        }
    }
    if (live && t == 0) *info = zero;
}

__global__ void reset_counter_kernel(int* counter) { *counter = 0; }

__global__ __launch_bounds__(kWG) void vbatched_kernel(const Item* __restrict__ items, int numItems,
                                                       const int* __restrict__ order, const int* __restrict__ nArr,
                                                       const long long* __restrict__ pivOff, int nrhs,
                                                       float* const* __restrict__ Aarr, float* const* __restrict__ Barr,
                                                       int* __restrict__ ipiv, int* __restrict__ info,
                                                       int* __restrict__ counter)
{
    // Per-group LDS slices: group g of a class with NMAX rows uses [g * NMAX, (g+1) * NMAX).
    __shared__ float s_prow[1024];
    __shared__ float s_l[1024];
    __shared__ float s_rhs[16][kNrhsMax];
    __shared__ float s_best[16][16];
    __shared__ int s_bpos[16][16];
    __shared__ int s_item;
    const int tid = threadIdx.x;
    for (;;)
    {
        if (tid == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= numItems) break;  // every workgroup reaches this exit once the queue drains
        const Item item = items[it];
        if (item.bin == 0)  // 16 matrices, one wave64 each, wave-level synchronization only
        {
            const int g = tid >> 6, t = tid & 63;
            const bool live = g < item.count;
            const int mi = live ? order[item.first + g] : 0;
            solve_group<64, 1>(live ? nArr[mi] : 0, nrhs, live ? Aarr[mi] : nullptr, live ? Barr[mi] : nullptr,
                               ipiv + (live ? pivOff[mi] : 0), info + mi, t, live, s_prow + g * 64, s_rhs[g],
                               s_best[g], s_bpos[g], s_l + g * 64);
        }
        else if (item.bin == 1)  // 4 matrices, 256 threads each (two threads per row)
        {
            const int g = tid >> 8, t = tid & 255;
            const bool live = g < item.count;
            const int mi = live ? order[item.first + g] : 0;
            solve_group<128, 2>(live ? nArr[mi] : 0, nrhs, live ? Aarr[mi] : nullptr, live ? Barr[mi] : nullptr,
                                ipiv + (live ? pivOff[mi] : 0), info + mi, t, live, s_prow + g * 128, s_rhs[g],
                                s_best[g], s_bpos[g], s_l + g * 128);
        }
        else  // 1 matrix, four threads per row
        {
            const int mi = order[item.first];
            solve_group<256, 4>(nArr[mi], nrhs, Aarr[mi], Barr[mi], ipiv + pivOff[mi], info + mi, tid, true, s_prow,
                                s_rhs[0], s_best[0], s_bpos[0], s_l);
        }
        __syncthreads();
    }
}

}  // namespace

struct VbatchedPlan
{
    int device = 0;
    int batch = 0;
    int nrhs = 0;
    int numItems = 0;
    int grid = 0;
    Item* dItems = nullptr;
    int* dOrder = nullptr;
    int* dN = nullptr;
    long long* dPivOff = nullptr;
    int* dCounter = nullptr;
    double flops = 0.0;
    double bytes = 0.0;
};

static void FreePlan(VbatchedPlan* p)
{
    if (p->dItems) (void)hipFree(p->dItems);
    if (p->dOrder) (void)hipFree(p->dOrder);
    if (p->dN) (void)hipFree(p->dN);
    if (p->dPivOff) (void)hipFree(p->dPivOff);
    if (p->dCounter) (void)hipFree(p->dCounter);
    delete p;
}

hipError_t VbatchedLaunch(VbatchedPlan* p, float* const* dA, float* const* dB, int* dIpiv, int* dInfo, hipStream_t s)
{
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, p->dCounter);
    hipLaunchKernelGGL(vbatched_kernel, dim3(p->grid), dim3(kWG), 0, s, p->dItems, p->numItems, p->dOrder, p->dN,
                       p->dPivOff, p->nrhs, dA, dB, dIpiv, dInfo, p->dCounter);
    return hipGetLastError();
}

// Autotuner: times each candidate persistent-grid size once on synthetic diagonally
// dominant systems of the plan's exact sizes and keeps the fastest.
static hipError_t Autotune(VbatchedPlan* p, const std::vector<int>& ns, hipStream_t s, int numCUs)
{
    const int batch = p->batch;
    std::vector<size_t> offA(batch + 1, 0), offB(batch + 1, 0);
    for (int i = 0; i < batch; ++i)
    {
        offA[i + 1] = offA[i] + static_cast<size_t>(ns[i]) * ns[i];
        offB[i + 1] = offB[i] + static_cast<size_t>(ns[i]) * p->nrhs;
    }
    std::vector<float> hA(offA[batch]), hB(offB[batch], 1.0f);
    for (int i = 0; i < batch; ++i)
        for (int c = 0; c < ns[i]; ++c)
            for (int r = 0; r < ns[i]; ++r)
                hA[offA[i] + static_cast<size_t>(c) * ns[i] + r] =
                    (r == c ? 2.0f : 0.0f) + 0.25f * static_cast<float>(((r * 7 + c * 13 + i) % 17) - 8) / 8.0f /
                                                 static_cast<float>(ns[i]);
    float *dA = nullptr, *dB = nullptr;
    float **dAp = nullptr, **dBp = nullptr;
    int *dPiv = nullptr, *dInfo = nullptr;
    size_t pivTotal = 0;
    for (int v : ns) pivTotal += v;
    hipError_t e = hipMalloc(&dA, hA.size() * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dB, hB.size() * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dAp, batch * sizeof(float*));
    if (e == hipSuccess) e = hipMalloc(&dBp, batch * sizeof(float*));
    if (e == hipSuccess) e = hipMalloc(&dPiv, pivTotal * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&dInfo, batch * sizeof(int));
    std::vector<float*> pa(batch), pb(batch);
    for (int i = 0; i < batch && e == hipSuccess; ++i)
    {
        pa[i] = dA + offA[i];
        pb[i] = dB + offB[i];
    }
    if (e == hipSuccess) e = hipMemcpy(dAp, pa.data(), batch * sizeof(float*), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dBp, pb.data(), batch * sizeof(float*), hipMemcpyHostToDevice);
    double bestT = 1e30;
    int bestGrid = numCUs;
    for (int mult : {1, 2, 3, 4})
    {
        if (e != hipSuccess) break;
        p->grid = std::max(1, std::min(p->numItems, numCUs * mult));
        double t = 0.0;
        for (int rep = 0; rep < 2 && e == hipSuccess; ++rep)
        {
            e = hipMemcpyAsync(dA, hA.data(), hA.size() * sizeof(float), hipMemcpyHostToDevice, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            const auto t0 = std::chrono::steady_clock::now();
            if (e == hipSuccess) e = VbatchedLaunch(p, dAp, dBp, dPiv, dInfo, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            const auto t1 = std::chrono::steady_clock::now();
            t = std::chrono::duration<double>(t1 - t0).count();  // keep the second (warm) run
        }
        if (e == hipSuccess && t < bestT)
        {
            bestT = t;
            bestGrid = p->grid;
        }
    }
    p->grid = bestGrid;
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dAp);
    (void)hipFree(dBp);
    (void)hipFree(dPiv);
    (void)hipFree(dInfo);
    return e;
}

}  // namespace ops

using namespace ops;

OPS_API aclsolverStatus_t aclsolverSgesvVbatched_createPlan(aclsolverHandle_t handle, int batchSize,
                                                            const int* nArray, int nrhs,
                                                            aclsolverVbatchedPlan_t* plan)
{
    if (handle == nullptr) return ACLSOLVER_STATUS_HANDLE_IS_NULLPTR;
    if (plan == nullptr || batchSize < 0 || (batchSize > 0 && nArray == nullptr) || nrhs < 0 || nrhs > kNrhsMax)
        return ACLSOLVER_STATUS_INVALID_VALUE;
    std::vector<int> ns(nArray, nArray + batchSize);
    for (int v : ns)
        if (v < 1 || v > 256) return ACLSOLVER_STATUS_INVALID_VALUE;
    VbatchedPlan* p = new (std::nothrow) VbatchedPlan();
    if (p == nullptr) return ACLSOLVER_STATUS_ALLOC_FAILED;
    p->batch = batchSize;
    p->nrhs = nrhs;
    (void)hipGetDevice(&p->device);
    // bins, largest first; within a bin, larger orders first (longest-processing-time order)
    std::vector<int> order(batchSize);
    for (int i = 0; i < batchSize; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ns[x] > ns[y]; });
    std::vector<Item> items;
    for (int i = 0; i < batchSize;)
    {
        const int n = ns[order[i]];
        const int bin = n <= 64 ? 0 : (n <= 128 ? 1 : 2);
        const int per = bin == 0 ? 16 : (bin == 1 ? 4 : 1);
        int cnt = 0;
        while (cnt < per && i + cnt < batchSize)
        {
            const int m = ns[order[i + cnt]];
            if ((m <= 64 ? 0 : (m <= 128 ? 1 : 2)) != bin) break;
            ++cnt;
        }
        items.push_back({bin, i, cnt});
        i += cnt;
    }
    std::vector<long long> pivOff(batchSize + 1, 0);
    for (int i = 0; i < batchSize; ++i)
    {
        pivOff[i + 1] = pivOff[i] + ns[i];
        const double n = ns[i];
        p->flops += 2.0 / 3.0 * n * n * n + 2.0 * n * n * nrhs;
        p->bytes += (2.0 * n * n + 2.0 * n * nrhs) * 4.0 + n * 4.0 + 4.0;
    }
    p->numItems = static_cast<int>(items.size());
    hipStream_t s = Stream(handle);
    hipError_t e = hipSuccess;
    if (batchSize > 0)
    {
        e = hipMalloc(&p->dItems, std::max<size_t>(1, items.size()) * sizeof(Item));
        if (e == hipSuccess) e = hipMalloc(&p->dOrder, batchSize * sizeof(int));
        if (e == hipSuccess) e = hipMalloc(&p->dN, batchSize * sizeof(int));
        if (e == hipSuccess) e = hipMalloc(&p->dPivOff, batchSize * sizeof(long long));
        if (e == hipSuccess) e = hipMalloc(&p->dCounter, 256);
        if (e == hipSuccess)
            e = hipMemcpy(p->dItems, items.data(), items.size() * sizeof(Item), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(p->dOrder, order.data(), batchSize * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(p->dN, ns.data(), batchSize * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(p->dPivOff, pivOff.data(), batchSize * sizeof(long long), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = Autotune(p, ns, s, H(handle)->numCUs);
    }
    if (e != hipSuccess)
    {
        (void)hipGetLastError();
        FreePlan(p);
        return ACLSOLVER_STATUS_ALLOC_FAILED;
    }
    *plan = p;
    return ACLSOLVER_STATUS_SUCCESS;
}

OPS_API aclsolverStatus_t aclsolverSgesvVbatched_execute(aclsolverHandle_t handle, aclsolverVbatchedPlan_t plan,
                                                         float* const dAarray[], float* const dBarray[], int* dIpiv,
                                                         int* dInfo)
{
    if (handle == nullptr) return ACLSOLVER_STATUS_HANDLE_IS_NULLPTR;
    if (plan == nullptr) return ACLSOLVER_STATUS_INVALID_VALUE;
    VbatchedPlan* p = static_cast<VbatchedPlan*>(plan);
    if (p->batch == 0) return ACLSOLVER_STATUS_SUCCESS;
    if (dAarray == nullptr || dBarray == nullptr || dIpiv == nullptr || dInfo == nullptr)
        return ACLSOLVER_STATUS_INVALID_VALUE;
    ProfileScope prof(ProfileKernel::Vbatched, p->bytes, Stream(handle));
    if (VbatchedLaunch(p, dAarray, dBarray, dIpiv, dInfo, Stream(handle)) != hipSuccess)
        return ACLSOLVER_STATUS_EXECUTION_FAILED;
    return ACLSOLVER_STATUS_SUCCESS;
}

OPS_API aclsolverStatus_t aclsolverSgesvVbatched_destroyPlan(aclsolverVbatchedPlan_t plan)
{
    if (plan == nullptr) return ACLSOLVER_STATUS_INVALID_VALUE;
    FreePlan(static_cast<VbatchedPlan*>(plan));
    return ACLSOLVER_STATUS_SUCCESS;
}

OPS_API aclsolverStatus_t aclsolverSgesvVbatched(aclsolverHandle_t handle, int batchSize, const int* nArray, int nrhs,
                                                 float* const dAarray[], float* const dBarray[], int* dIpiv,
                                                 int* dInfo)
{
    aclsolverVbatchedPlan_t plan = nullptr;
    aclsolverStatus_t st = aclsolverSgesvVbatched_createPlan(handle, batchSize, nArray, nrhs, &plan);
    if (st != ACLSOLVER_STATUS_SUCCESS) return st;
    st = aclsolverSgesvVbatched_execute(handle, plan, dAarray, dBarray, dIpiv, dInfo);
    if (hipStreamSynchronize(Stream(handle)) != hipSuccess) st = ACLSOLVER_STATUS_EXECUTION_FAILED;
    aclsolverSgesvVbatched_destroyPlan(plan);
    return st;
}
