// Single-launch triangular solve of one right-hand side with packed LU factors.
//
// The blocked solve (a diagonal TRSM and a GEMV per block row, two launches per block)
// is bound by launch latency for one vector: n = 16384 with 128-row blocks is 256
// dependent launches per solve. Here one launch covers the whole triangle. Workgroup b
// owns row block b: it streams its off-diagonal tiles of the factor while earlier blocks
// are still being solved, folds in each solved block x_j as soon as it is published, then
// solves its own diagonal block and publishes x_b. The only serial part left is the chain
// of diagonal solves plus one device-scope round trip per block.
//
// Publication: every solved entry is written as one 64-bit word (epoch << 32 | float
// bits) with a relaxed agent-scope atomic store, and readers poll the words themselves, so
// no separate flag or fence is needed and a stale word can never pass as current. Block
// order comes from a ticket counter (a workgroup that draws ticket t only waits for
// tickets < t, which were drawn by running workgroups), so progress does not depend on
// the hardware's dispatch order. Every wait is bounded; a wait that runs out raises the
// error word and the workgroup finishes without blocking anyone.
//
// The algorithm is forward / backward substitution (LAPACK strsv, Uplo L/U, Trans N), the
// same two triangle stages the reference's solve runs after its row interchanges
// (src/sgetrs/sgetrs_kernel.cpp:55-57: laswp, lower trsm, upper trsm).
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "internal.h"
#include "trsv.h"

namespace ops
{
namespace
{

__device__ inline unsigned long long PollWord(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publication words of one solved entry: a float takes one word (epoch << 32 | bits), a
// double two (low and high halves, each tagged with the epoch).
template <typename V>
struct Words
{
    static constexpr int count = sizeof(V) / 4;
};

__device__ inline float Broadcast(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ inline double Broadcast(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// NB rows per block, 256 threads: wave w streams columns [w*NB/4, (w+1)*NB/4) of every
// off-diagonal tile, lane l owns rows l, l+64, ... of the block. V is the vector type
// (float, or double for a solve carried out in FP64 on FP32 factors).
template <int NB, bool UPPER, typename V>
__global__ __launch_bounds__(256) void trsv_chain_kernel(int n, const float* __restrict__ T, int ldt,
                                                         V* __restrict__ x, unsigned long long* __restrict__ pub,
                                                         unsigned* __restrict__ ticket, unsigned base, unsigned epoch,
                                                         int* __restrict__ err, unsigned spinLimit)
{
    constexpr int RPL = NB / 64;  // rows per lane
    constexpr int CPW = NB / 4;   // tile columns per wave
    constexpr int NW = Words<V>::count;
    __shared__ float s_diag[NB][NB + 1];
    __shared__ V s_part[4][NB];
    __shared__ V s_rinv[NB];
    __shared__ unsigned s_id;
    __shared__ int s_fail;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nblk = (n + NB - 1) / NB;
    if (tid == 0) s_id = atomicAdd(ticket, 1u) - base;
    __syncthreads();
    const int id = static_cast<int>(s_id);
    const int lb = UPPER ? nblk - 1 - id : id;
    const int r0 = lb * NB;
    const int rows = min(NB, n - r0);

    // Diagonal tile (row-major in LDS); rows past n become identity rows.
    for (int e = tid; e < NB * NB; e += 256)
    {
        const int c = e / NB, r = e % NB;
        s_diag[r][c] = (r < rows && c < rows) ? T[static_cast<size_t>(r0 + c) * ldt + r0 + r] : (r == c ? 1.0f : 0.0f);
    }
    V bv[RPL];
#pragma unroll
    for (int q = 0; q < RPL; ++q)
    {
        const int r = lane + 64 * q;
        bv[q] = (wave == 0 && r < rows) ? x[r0 + r] : V(0);
    }

    const int ndep = UPPER ? nblk - 1 - lb : lb;
    V acc[RPL];
    float cur[RPL][CPW], nxt[RPL][CPW];
#pragma unroll
    for (int q = 0; q < RPL; ++q) acc[q] = V(0);
    auto loadTile = [&](int k, float (&buf)[RPL][CPW])
    {
        const int j = UPPER ? nblk - 1 - k : k;
        const int c0 = j * NB + wave * CPW;
#pragma unroll
        for (int c = 0; c < CPW; ++c)
#pragma unroll
            for (int q = 0; q < RPL; ++q)
            {
                const int r = lane + 64 * q;
                buf[q][c] = (r < rows && c0 + c < n) ? T[static_cast<size_t>(c0 + c) * ldt + r0 + r] : 0.0f;
            }
    };
    if (ndep > 0) loadTile(0, cur);
    bool failed = false;
    for (int k = 0; k < ndep && !failed; ++k)  // after one expired wait, stop waiting
    {
        if (k + 1 < ndep) loadTile(k + 1, nxt);
        const int j = UPPER ? nblk - 1 - k : k;
        const int idx = j * NB + wave * CPW + lane;
        const bool mine = lane < CPW && idx < n;
        unsigned long long w[NW];
#pragma unroll
        for (int h = 0; h < NW; ++h) w[h] = static_cast<unsigned long long>(epoch) << 32;
        unsigned spins = 0;
        while (true)
        {
            if (mine)
#pragma unroll
                for (int h = 0; h < NW; ++h) w[h] = PollWord(pub + static_cast<size_t>(idx) * NW + h);
            bool ok = true;
#pragma unroll
            for (int h = 0; h < NW; ++h) ok = ok && static_cast<unsigned>(w[h] >> 32) == epoch;
            if (__all(ok)) break;
            if (++spins > spinLimit)
            {
                failed = true;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        V yv;
        if constexpr (NW == 1)
            yv = __uint_as_float(static_cast<unsigned>(w[0]));
        else
            yv = __longlong_as_double(static_cast<long long>((w[1] << 32) | (w[0] & 0xffffffffull)));
#pragma unroll
        for (int c = 0; c < CPW; ++c)
        {
            const V ys = Broadcast(yv, c);
#pragma unroll
            for (int q = 0; q < RPL; ++q) acc[q] = fma(static_cast<V>(cur[q][c]), ys, acc[q]);
        }
#pragma unroll
        for (int c = 0; c < CPW; ++c)
#pragma unroll
            for (int q = 0; q < RPL; ++q) cur[q][c] = nxt[q][c];
    }
    if (tid == 0) s_fail = 0;
    __syncthreads();
    if (failed && lane == 0)
    {
        s_fail = 1;
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int q = 0; q < RPL; ++q) s_part[wave][lane + 64 * q] = acc[q];
    if (UPPER)
        for (int r = tid; r < NB; r += 256) s_rinv[r] = V(1) / static_cast<V>(s_diag[r][r]);
    __syncthreads();
    if (wave != 0) return;

    V xv[RPL];
#pragma unroll
    for (int q = 0; q < RPL; ++q)
    {
        const int r = lane + 64 * q;
        xv[q] = bv[q] - (((s_part[0][r] + s_part[1][r]) + s_part[2][r]) + s_part[3][r]);
    }
    if (!UPPER)
    {
        // unit lower: x_i -= L_is x_s for i > s
#pragma unroll
        for (int s = 0; s < NB; ++s)
        {
            const V ys = Broadcast(xv[s / 64], s % 64);
#pragma unroll
            for (int q = 0; q < RPL; ++q)
            {
                const int r = lane + 64 * q;
                const V l = r > s ? static_cast<V>(s_diag[r][s]) : V(0);
                xv[q] = fma(-l, ys, xv[q]);
            }
        }
    }
    else
    {
        // non-unit upper, bottom to top: x_s = x_s / U_ss (as a product with the rounded
        // reciprocal), then x_i -= U_is x_s for i < s
#pragma unroll
        for (int s = NB - 1; s >= 0; --s)
        {
            const V ys = Broadcast(xv[s / 64], s % 64) * s_rinv[s];
#pragma unroll
            for (int q = 0; q < RPL; ++q)
            {
                const int r = lane + 64 * q;
                const V u = r < s ? static_cast<V>(s_diag[r][s]) : V(0);
                xv[q] = r == s ? ys : fma(-u, ys, xv[q]);
            }
        }
    }
    // A block whose wait ran out still publishes (NaN), so its dependents finish too and
    // the failure is visible in the result.
    if (s_fail)
#pragma unroll
        for (int q = 0; q < RPL; ++q) xv[q] = static_cast<V>(__int_as_float(0x7fc00000));
#pragma unroll
    for (int q = 0; q < RPL; ++q)
    {
        const int r = lane + 64 * q;
        if (r < rows)
        {
            x[r0 + r] = xv[q];
            const unsigned long long tag = static_cast<unsigned long long>(epoch) << 32;
            if constexpr (NW == 1)
            {
                const unsigned long long word = tag | __float_as_uint(xv[q]);
                __hip_atomic_store(pub + r0 + r, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            else
            {
                const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(xv[q]));
                __hip_atomic_store(pub + 2 * static_cast<size_t>(r0 + r), tag | (b & 0xffffffffull), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(pub + 2 * static_cast<size_t>(r0 + r) + 1, tag | (b >> 32), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

__global__ void trsv_reset_kernel(int words, unsigned long long* pub, unsigned* ticket, int* err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) pub[i] = 0ull;
    if (i == 0)
    {
        *ticket = 0u;
        *err = 0;
    }
}

// Rows per block: 64 (one row per lane) or 128 (two). Larger blocks halve the number of
// serial steps but lengthen each diagonal solve. OPS_SOLVER_TRSV_NB=64|128 overrides the
// size-based choice (tuning aid).
int ChainBlock(int n)
{
    static const int forced = []()
    {
        const char* v = std::getenv("OPS_SOLVER_TRSV_NB");
        return v ? std::atoi(v) : 0;
    }();
    if (forced == 64 || forced == 128) return forced;
    return n >= 4096 ? 128 : 64;
}
}  // namespace

size_t TrsvChainStateBytes(int n) { return (static_cast<size_t>(n) * 16 + 256 + 255) / 256 * 256; }

TrsvChainState TrsvChainLayout(void* base, int n)
{
    TrsvChainState s{};
    char* p = static_cast<char*>(base);
    s.pub = reinterpret_cast<unsigned long long*>(p);
    s.ticket = reinterpret_cast<unsigned*>(p + static_cast<size_t>(n) * 16);
    s.err = reinterpret_cast<int*>(p + static_cast<size_t>(n) * 16 + 64);
    s.n = n;
    s.launches = 0;
    return s;
}

hipError_t TrsvChainReset(TrsvChainState& s, hipStream_t stream)
{
    s.launches = 0;
    hipLaunchKernelGGL(trsv_reset_kernel, dim3((2 * s.n + 255) / 256), dim3(256), 0, stream, 2 * s.n, s.pub, s.ticket,
                       s.err);
    return hipGetLastError();
}

template <typename V>
hipError_t LaunchChain(TrsvChainState& s, bool upper, const float* T, int ldt, V* x, hipStream_t stream)
{
    const int n = s.n;
    if (n <= 0) return hipSuccess;
    const int nb = ChainBlock(n);
    const int nblk = (n + nb - 1) / nb;
    // Tickets are drawn from one 32-bit counter: start over before it could wrap.
    if ((s.launches + 1) * static_cast<long long>(nblk) >= (1ll << 31))
    {
        hipError_t e = TrsvChainReset(s, stream);
        if (e != hipSuccess) return e;
    }
    const unsigned epoch = static_cast<unsigned>(s.launches) + 1u;
    const unsigned base = static_cast<unsigned>(s.launches) * static_cast<unsigned>(nblk);
    const unsigned spinLimit = SpinLimit();
#define CHAIN(NB, UP)                                                                                             \
    hipLaunchKernelGGL((trsv_chain_kernel<NB, UP, V>), dim3(nblk), dim3(256), 0, stream, n, T, ldt, x, s.pub, s.ticket, \
                       base, epoch, s.err, spinLimit)
    if (nb == 128)
    {
        if (upper)
            CHAIN(128, true);
        else
            CHAIN(128, false);
    }
    else
    {
        if (upper)
            CHAIN(64, true);
        else
            CHAIN(64, false);
    }
#undef CHAIN
    ++s.launches;
    return hipGetLastError();
}

hipError_t TrsvChain(TrsvChainState& s, bool upper, const float* T, int ldt, float* x, hipStream_t stream)
{
    return LaunchChain<float>(s, upper, T, ldt, x, stream);
}

hipError_t TrsvChain(TrsvChainState& s, bool upper, const float* T, int ldt, double* x, hipStream_t stream)
{
    return LaunchChain<double>(s, upper, T, ldt, x, stream);
}

}  // namespace ops
