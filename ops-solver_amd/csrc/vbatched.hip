// Uneven batched LU solve (getrf + getrs per matrix) in one persistent launch (EXTENSION;
// the reference has no batched getrf/getrs).
//
// Plan: matrices are binned by order into 64 / 128 / 256 row classes; a work item fills one
// 512-thread workgroup (8 matrices of n <= 64 with one wave each, 4 of n <= 128 with two
// waves each, or 1 of n <= 256 with two threads per row), and every thread keeps at most
// 128 matrix values in VGPRs. Items are queued largest class first, larger orders first
// (longest-processing-time order). A persistent grid pulls items from a device counter
// until the queue is empty, so every workgroup reaches the exit. The autotuner picks the
// grid size at plan time.
//
// Per matrix: rows live in registers (position-tracked partial pivoting exactly as the
// register leaf of getrf.hip: largest |a|, lowest position on ties, reciprocal scaling,
// fused rank-1 updates); per column a 64-bit key argmax (DPP in the wave, one slot per wave
// through LDS) and two barriers (one wave: wave-level only); the right-hand sides ride
// along with their rows, then a back substitution in 32-column blocks. Outputs: packed LU
// (rows stored at their final positions), 1-based ipiv, X, info (first zero pivot).
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
constexpr int kWG = 512;
constexpr int kBS = 32;      // back-substitution block

struct Item
{
    int bin;     // 0: n<=64 (8 per item), 1: n<=128 (4 per item), 2: n<=256 (1 per item)
    int first;   // index into the plan's matrix order
    int count;   // matrices in this item
};

__device__ inline unsigned long long KeyOf(float v, int pos)
{
    const float a = fabsf(v);
    const unsigned hi = (a == a) ? __float_as_uint(a) + 1u : 0u;  // NaN never wins
    return (1ull << 63) | (static_cast<unsigned long long>(hi) << 32) | static_cast<unsigned>(~pos);
}

template <int CTRL>
__device__ inline unsigned long long Dpp64(unsigned long long v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(v & 0xffffffffull), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(v >> 32), CTRL, 0xF, 0xF, false);
    return (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) | static_cast<unsigned>(lo);
}
__device__ inline unsigned long long Lane64(unsigned long long v, int lane)
{
    const unsigned lo = __builtin_amdgcn_readlane(static_cast<int>(v & 0xffffffffull), lane);
    const unsigned hi = __builtin_amdgcn_readlane(static_cast<int>(v >> 32), lane);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}
__device__ inline unsigned long long WaveMax64(unsigned long long v)
{
    auto mx = [](unsigned long long a, unsigned long long b) { return a > b ? a : b; };
    v = mx(v, Dpp64<0xB1>(v));   // quad_perm [1,0,3,2]
    v = mx(v, Dpp64<0x4E>(v));   // quad_perm [2,3,0,1]
    v = mx(v, Dpp64<0x124>(v));  // row_ror:4
    v = mx(v, Dpp64<0x128>(v));  // row_ror:8
    return mx(mx(Lane64(v, 0), Lane64(v, 16)), mx(Lane64(v, 32), Lane64(v, 48)));
}

// LDS of one group (one matrix).
template <int NMAX, int TPR>
struct GroupLds
{
    static constexpr int NWG = NMAX * TPR / 64;  // waves in the group
    static constexpr int C = NMAX / TPR;         // columns per thread
    unsigned long long key[NWG];
    float val[NWG];
    float prow[TPR][C];  // pivot row, de-interleaved by column owner
    float rhs[kNrhsMax];
    float l[TPR > 1 ? NMAX : 1];
    float U[kBS][kBS + 1];
    float y[kBS][kNrhsMax];
    float part[TPR > 1 ? NMAX : 1][kNrhsMax];
};

// LU with partial pivoting of one n x n matrix (n <= NMAX) plus the solve of nrhs right-hand
// sides, by G = NMAX * TPR threads. Thread t owns row r = t % NMAX and the columns
// j = TPR * jj + h, h = t / NMAX (C = NMAX / TPR of them) in registers. Pivoting is
// implicit: rows keep their registers and track the position they occupy; the pivot of
// column k is the candidate with the largest |a| and, on ties, the lowest position (the
// LAPACK first-row rule), scaling is by the reciprocal when |d| >= FLT_MIN (division
// otherwise), updates are fused multiply-adds -- as in the register leaf of the dense
// factorization. The right-hand sides ride along with their rows (forward substitution
// during elimination); back substitution runs in blocks of 32 columns.
template <int NMAX, int TPR, bool WAVE>
__device__ __attribute__((always_inline)) void solve_group(int n, int nLoop, int nrhs, float* __restrict__ A,
                                                         float* __restrict__ B, int* __restrict__ ipiv,
                                                         int* __restrict__ info, int t, bool live,
                                                         GroupLds<NMAX, TPR>& L)
{
    // nLoop: the largest order among the groups that synchronize together (workgroup
    // barriers), so every group runs the same number of barrier steps.
    constexpr int C = NMAX / TPR;
    constexpr int NWG = NMAX * TPR / 64;
    constexpr int CH = 8;  // columns per update chunk (per thread)
    auto sync = [&]()
    {
        if (WAVE)
        {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        else
            __syncthreads();
    };
    const int r = t % NMAX;
    const int h = __builtin_amdgcn_readfirstlane(t / NMAX);  // uniform per wave (NMAX % 64 == 0)
    const int wig = __builtin_amdgcn_readfirstlane(t >> 6);
    const int lane = t & 63;
    const bool rowLive = live && r < n;
    float a[C];
    {
        // Loads in chunks: each chunk's addresses die before the next chunk issues.
        const float* src = A + r + static_cast<size_t>(h) * n;
        const size_t stride = static_cast<size_t>(TPR) * n;
#pragma unroll
        for (int c0 = 0; c0 < C; c0 += 16)
        {
#pragma unroll
            for (int jj = c0; jj < c0 + 16; ++jj)
            {
                const int j = TPR * jj + h;
                a[jj] = (rowLive && j < n) ? src[jj * stride] : 0.0f;
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    float b[kNrhsMax];  // every thread of a row carries the right-hand sides
#pragma unroll
    for (int q = 0; q < kNrhsMax; ++q) b[q] = (rowLive && q < nrhs) ? B[static_cast<size_t>(q) * n + r] : 0.0f;
    int pos = rowLive ? r : INT_MAX;
    float cur = a[0];  // value in column k when this thread owns column k
    int zero = 0;
    const int steps = live ? n : 0;

    for (int k = 0; k < NMAX; ++k)
    {
        if (WAVE && k >= steps) break;   // a wave-private group stops on its own
        if (!WAVE && k >= nLoop) break;  // uniform across the workgroup
        const bool active = k < steps;
        const bool ownK = (k % TPR) == h;
        // (1) candidates of column k: the wave's best key, then one slot per wave
        const unsigned long long myKey =
            (active && ownK && pos >= k && pos != INT_MAX) ? KeyOf(cur, pos) : 0ull;
        const unsigned long long wmax = WaveMax64(myKey);
        if (myKey != 0ull && myKey == wmax)
        {
            L.key[wig] = wmax;
            L.val[wig] = cur;
        }
        if (lane == 0 && wmax == 0ull) L.key[wig] = 0ull;
        sync();
        unsigned long long top = L.key[0];
        int tw = 0;
#pragma unroll
        for (int w = 1; w < NWG; ++w)
        {
            const unsigned long long kw = L.key[w];
            if (kw > top)
            {
                top = kw;
                tw = w;
            }
        }
        const int bpos = top == 0ull ? k : static_cast<int>(~static_cast<unsigned>(top & 0xffffffffull));
        const float d = top == 0ull ? 0.0f : L.val[tw];
        // (2) publish the pivot row and this row's multiplier
        float l = 0.0f;
        if (active)
        {
            if (pos == bpos)
            {
#pragma unroll
                for (int jj = 0; jj < C; ++jj) L.prow[h][jj] = a[jj];
                if (h == 0)
#pragma unroll
                    for (int q = 0; q < kNrhsMax; ++q) L.rhs[q] = b[q];
            }
            if (pos == bpos)
                pos = k;
            else if (pos == k)
                pos = bpos;
            if (t == 0) ipiv[k] = bpos + 1;
            if (d == 0.0f && zero == 0) zero = k + 1;
            if (ownK)
            {
                const bool below = pos > k && pos != INT_MAX && d != 0.0f;
                const float q = fabsf(d) >= FLT_MIN ? cur * (1.0f / d) : cur / d;
                l = below ? q : 0.0f;
                if (TPR > 1) L.l[r] = l;
            }
        }
        sync();
        // (3) rank-1 update of the columns right of k (whole chunks left of k are skipped)
        if (active)
        {
            if (TPR > 1 && !ownK) l = L.l[r];
            float next = 0.0f;
#pragma unroll
            for (int c0 = 0; c0 < C; c0 += CH)
            {
                // Columns of this thread in the chunk: TPR*c0 + h .. TPR*(c0+CH-1) + h
                // (h is uniform per wave, k per group: every branch below is uniform).
                const int jlo = TPR * c0 + h, jhi = TPR * (c0 + CH - 1) + h;
                if (jhi < k) continue;  // finished columns
                if (jlo > k + 1)
                {
                    // all columns right of k + 1: plain rank-1 update
#pragma unroll
                    for (int jj = c0; jj < c0 + CH; ++jj) a[jj] = fmaf(-l, L.prow[h][jj], a[jj]);
                }
                else
                {
#pragma unroll
                    for (int jj = c0; jj < c0 + CH; ++jj)
                    {
                        const int j = TPR * jj + h;
                        const float u = L.prow[h][jj];
                        if (j > k) a[jj] = fmaf(-l, u, a[jj]);
                        if (j == k && ownK && l != 0.0f) a[jj] = l;  // store the multiplier
                        if (j == k + 1) next = a[jj];
                    }
                }
                __builtin_amdgcn_sched_barrier(0);  // keep each chunk's pivot-row loads local
            }
#pragma unroll
            for (int q = 0; q < kNrhsMax; ++q) b[q] = fmaf(-l, L.rhs[q], b[q]);
            cur = next;
        }
    }

    // Back substitution U x = y over final positions, blocks of kBS columns from the end of
    // this matrix; the number of block steps is uniform across the synchronizing groups.
    const int nb = live ? n : 0;
    const int blocks = ((WAVE ? nb : nLoop) + kBS - 1) / kBS;
    for (int bi = 0; bi < blocks; ++bi)
    {
        const int c1 = nb - bi * kBS;  // <= 0: this matrix has no block left (barriers only)
        const int c0 = c1 - kBS > 0 ? c1 - kBS : 0;
        const bool inBlk = rowLive && pos >= c0 && pos < c1;
        // (a) diagonal block rows and their right-hand sides to LDS
        if (inBlk)
        {
#pragma unroll
            for (int cc = 0; cc < C; cc += CH)
            {
                if (TPR * (cc + CH) <= c0 || TPR * cc >= c1) continue;  // uniform
#pragma unroll
                for (int jj = cc; jj < cc + CH; ++jj)
                {
                    const int j = TPR * jj + h;
                    if (j >= c0 && j < c1) L.U[pos - c0][j - c0] = a[jj];
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            if (h == 0)
#pragma unroll
                for (int q = 0; q < kNrhsMax; ++q) L.y[pos - c0][q] = b[q];
        }
        sync();
        // (b) one wave solves the block: lane i holds row i
        if (wig == 0 && c1 > 0)
        {
            const int bs = c1 - c0;
            float yv[kNrhsMax];
#pragma unroll
            for (int q = 0; q < kNrhsMax; ++q) yv[q] = lane < bs ? L.y[lane][q] : 0.0f;
            for (int s2 = bs - 1; s2 >= 0; --s2)
            {
                const float dss = L.U[s2][s2];
                const float u = lane < s2 ? L.U[lane][s2] : 0.0f;
#pragma unroll
                for (int q = 0; q < kNrhsMax; ++q)
                {
                    const float xs = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(yv[q]), s2)) / dss;
                    if (lane == s2) yv[q] = xs;
                    yv[q] = fmaf(-u, xs, yv[q]);
                    if (lane == s2) yv[q] = xs;
                }
            }
            if (lane < bs)
#pragma unroll
                for (int q = 0; q < kNrhsMax; ++q) L.y[lane][q] = yv[q];
        }
        sync();
        // (c) rows of the block take their solution; rows above fold the block in
        if (inBlk)
        {
#pragma unroll
            for (int q = 0; q < kNrhsMax; ++q) b[q] = L.y[pos - c0][q];
        }
        float part[kNrhsMax] = {};
        const bool above = rowLive && pos < c0;
        if (above)
        {
#pragma unroll
            for (int cc = 0; cc < C; cc += CH)
            {
                if (TPR * (cc + CH) <= c0 || TPR * cc >= c1) continue;  // uniform
#pragma unroll
                for (int jj = cc; jj < cc + CH; ++jj)
                {
                    const int j = TPR * jj + h;
                    if (j >= c0 && j < c1)
#pragma unroll
                        for (int q = 0; q < kNrhsMax; ++q) part[q] = fmaf(a[jj], L.y[j - c0][q], part[q]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (TPR > 1)
        {
            if (above && h == 1)
#pragma unroll
                for (int q = 0; q < kNrhsMax; ++q) L.part[r][q] = part[q];
            sync();
            if (above && h == 0)
#pragma unroll
                for (int q = 0; q < kNrhsMax; ++q) part[q] += L.part[r][q];
        }
        if (above && h == 0)
#pragma unroll
            for (int q = 0; q < kNrhsMax; ++q) b[q] -= part[q];
        sync();
    }

    if (rowLive)
    {
        float* dst = A + pos + static_cast<size_t>(h) * n;
        const size_t stride = static_cast<size_t>(TPR) * n;
#pragma unroll
        for (int c0 = 0; c0 < C; c0 += 16)
        {
#pragma unroll
            for (int jj = c0; jj < c0 + 16; ++jj)
            {
                const int j = TPR * jj + h;
                if (j < n) dst[jj * stride] = a[jj];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (h == 0)
#pragma unroll
            for (int q = 0; q < kNrhsMax; ++q)
                if (q < nrhs) B[static_cast<size_t>(q) * n + pos] = b[q];
    }
    if (live && t == 0) *info = zero;
}

__global__ void reset_counter_kernel(int* counter) { *counter = 0; }

union VbLds
{
    GroupLds<64, 1> s[8];
    GroupLds<128, 1> m[4];
    GroupLds<256, 2> l;
};

__global__ __launch_bounds__(kWG) void vbatched_kernel(const Item* __restrict__ items, int numItems,
                                                       const int* __restrict__ order, const int* __restrict__ nArr,
                                                       const long long* __restrict__ pivOff, int nrhs,
                                                       float* const* __restrict__ Aarr, float* const* __restrict__ Barr,
                                                       int* __restrict__ ipiv, int* __restrict__ info,
                                                       int* __restrict__ counter)
{
    __shared__ VbLds lds;
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
        if (item.bin == 0)  // 8 matrices, one wave64 each, wave-level synchronization only
        {
            const int g = tid >> 6, t = tid & 63;
            const bool live = g < item.count;
            const int mi = live ? order[item.first + g] : 0;
            const int n = live ? nArr[mi] : 0;
            solve_group<64, 1, true>(n, n, nrhs, live ? Aarr[mi] : nullptr, live ? Barr[mi] : nullptr,
                                     ipiv + (live ? pivOff[mi] : 0), info + mi, t, live, lds.s[g]);
        }
        else if (item.bin == 1)  // 4 matrices, 128 threads each, in lockstep (largest first)
        {
            const int g = tid >> 7, t = tid & 127;
            const bool live = g < item.count;
            const int mi = live ? order[item.first + g] : 0;
            const int nLead = nArr[order[item.first]];
            solve_group<128, 1, false>(live ? nArr[mi] : 0, nLead, nrhs, live ? Aarr[mi] : nullptr,
                                       live ? Barr[mi] : nullptr, ipiv + (live ? pivOff[mi] : 0), info + mi, t, live,
                                       lds.m[g]);
        }
        else  // 1 matrix, two threads per row
        {
            const int mi = order[item.first];
            const int n = nArr[mi];
            solve_group<256, 2, false>(n, n, nrhs, Aarr[mi], Barr[mi], ipiv + pivOff[mi], info + mi, tid, true,
                                       lds.l);
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
        const int per = bin == 0 ? 8 : (bin == 1 ? 4 : 1);
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
