// Complex64 LU family on gfx950: cgetrf, device cgetrs (trans N/T/C), batched direct
// inverse (cmatinv) and batched inverse from factors (cgetri_batched).
//
// Reference semantics (src/cgetrf/cgetrf_kernel.cpp:376-411, ComplexSolverGetrf):
//   * pivot = row with the largest L1 score |re|+|im| under the unsigned key of
//     src/utils/complex_solver_kernel.hpp:47-63 (NaN inactive, both zeros equal, an
//     overflow-safe half sum), lowest row on ties;
//   * multipliers are accurate quotients (the reference's Divide is specified to 4 ulp,
//     test/complex_solver/complex_solver_test.cpp:520-556): here a * (1/d) evaluated in
//     FP64 and rounded once;
//   * rank-1 updates use the fused order of SubMul (complex_solver_kernel.hpp:65-68);
//   * the trailing update is built from real MFMA products, C -= A B as
//     Re += Ar Br + (-Ai) Bi, Im += Ar Bi + Ai Br (the 4-product form of
//     src/utils/complex_solver_cube.cpp:206-214), FP32 v_mfma_f32_32x32x2_f32, FP32 accumulation.
// The blocked structure mirrors the real path: register-resident leaves with position
// tracking, in-panel interchange + TRSM + update, then outer interchanges, U12 solve and
// the MFMA trailing update.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstdint>

#include "internal.h"

namespace ops
{
namespace
{
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ inline uint32_t PivotScore(float2 v)
{
    const float r = fabsf(v.x), i = fabsf(v.y);
    float s = r + i;
    if (!(s >= 0.0f)) return 0u;
    if (s == 0.0f) return 1u;
    if (s > FLT_MAX)
    {
        s = r * 0.5f + i * 0.5f;
        return __float_as_uint(s) + 0x00800001u;
    }
    return __float_as_uint(s) + 1u;
}

__device__ inline float2 SubMul(float2 x, float2 a, float2 b)
{
    return make_float2(fmaf(a.y, b.y, fmaf(-a.x, b.x, x.x)), fmaf(-a.y, b.x, fmaf(-a.x, b.y, x.y)));
}

// 1/d in FP64 (exact enough that a * (1/d) rounds within one ulp of a / d).
__device__ inline double2 Recip64(float2 d)
{
    const double dr = d.x, di = d.y;
    const double den = dr * dr + di * di;
    return make_double2(dr / den, -di / den);
}

__device__ inline float2 Mul64(float2 a, double2 r)
{
    const double ar = a.x, ai = a.y;
    return make_float2(static_cast<float>(ar * r.x - ai * r.y), static_cast<float>(ar * r.y + ai * r.x));
}

__device__ inline float2 Conj(float2 v, bool c) { return c ? make_float2(v.x, -v.y) : v; }

// ---------------------------------------------------------------------------------------
// Leaf (see getrf.hip for the real counterpart): W complex columns, R rows per thread.
// ---------------------------------------------------------------------------------------
template <int T, int R, int W>
__global__ __launch_bounds__(T) void cleaf_kernel(float2* __restrict__ A, int lda, int m, int j0, int c0, int wact,
                                                  int pivoting, int newPanel, int* __restrict__ ipiv,
                                                  int* __restrict__ info, int* __restrict__ rowsrc,
                                                  int* __restrict__ leafmap)
{
    constexpr int NW = T / 64;
    __shared__ float2 s_prow[W];
    __shared__ uint32_t s_best[NW];
    __shared__ int s_bpos[NW];
    __shared__ int s_count;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int mp = m - c0;
    float2* Ap = A + static_cast<size_t>(c0) * lda + c0;
    float2 a[R][W];
    int pos[R];
#pragma unroll
    for (int i = 0; i < R; ++i)
    {
        const int lr = tid + i * T;
        const bool valid = lr < mp;
#pragma unroll
        for (int j = 0; j < W; ++j)
            a[i][j] = (valid && j < wact) ? Ap[static_cast<size_t>(j) * lda + lr] : make_float2(0.0f, 0.0f);
        pos[i] = valid ? lr : INT_MAX;
    }
    if (tid == 0) s_count = 0;
    int zeroCol = 0;
#pragma unroll
    for (int j = 0; j < W; ++j)
    {
        if (j < wact)
        {
            int bpos = j;
            if (pivoting)
            {
                uint32_t best = 0;
                int bp = INT_MAX;
#pragma unroll
                for (int i = 0; i < R; ++i)
                    if (pos[i] >= j && pos[i] != INT_MAX)
                    {
                        const uint32_t s = PivotScore(a[i][j]);
                        if (s > best || (s == best && pos[i] < bp))
                        {
                            best = s;
                            bp = pos[i];
                        }
                    }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1)
                {
                    const uint32_t ob = __shfl_xor(best, off);
                    const int op = __shfl_xor(bp, off);
                    if (ob > best || (ob == best && op < bp))
                    {
                        best = ob;
                        bp = op;
                    }
                }
                if (lane == 0)
                {
                    s_best[wid] = best;
                    s_bpos[wid] = bp;
                }
                __syncthreads();
                best = s_best[0];
                bp = s_bpos[0];
#pragma unroll
                for (int w = 1; w < NW; ++w)
                {
                    const uint32_t ob = s_best[w];
                    const int op = s_bpos[w];
                    if (ob > best || (ob == best && op < bp))
                    {
                        best = ob;
                        bp = op;
                    }
                }
                bpos = bp == INT_MAX ? j : bp;
            }
#pragma unroll
            for (int i = 0; i < R; ++i)
                if (pos[i] == bpos)
                {
#pragma unroll
                    for (int k = 0; k < W; ++k) s_prow[k] = a[i][k];
                }
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                if (pos[i] == bpos)
                    pos[i] = j;
                else if (pos[i] == j)
                    pos[i] = bpos;
            }
            if (tid == 0 && ipiv != nullptr) ipiv[c0 + j] = c0 + bpos + 1;
            __syncthreads();
            const float2 d = s_prow[j];
            const bool zero = d.x == 0.0f && d.y == 0.0f;
            if (zero && zeroCol == 0) zeroCol = c0 + j + 1;
            const double2 rinv = Recip64(zero ? make_float2(1.0f, 0.0f) : d);
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                const bool below = pos[i] > j && pos[i] != INT_MAX && !zero;
                const float2 q = Mul64(a[i][j], rinv);
                const float2 l = below ? q : make_float2(0.0f, 0.0f);
                a[i][j] = below ? q : a[i][j];
#pragma unroll
                for (int k = 1; k < W; ++k)  // constant bounds: unrolled with the j loop
                    if (k > j) a[i][k] = SubMul(a[i][k], l, s_prow[k]);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i)
    {
        const int lr = tid + i * T;
        if (lr < mp)
        {
            const int dst = pos[i];
#pragma unroll
            for (int j = 0; j < W; ++j)
                if (j < wact) Ap[static_cast<size_t>(j) * lda + dst] = a[i][j];
            if (dst != lr)
            {
                const int slot = atomicAdd(&s_count, 1);
                leafmap[1 + 2 * slot] = c0 + dst;
                leafmap[2 + 2 * slot] = c0 + lr;
            }
        }
    }
    int srcv[R];
#pragma unroll
    for (int i = 0; i < R; ++i)
    {
        const int lr = tid + i * T;
        srcv[i] = (lr < mp) ? (newPanel ? c0 + lr : rowsrc[c0 + lr]) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R; ++i)
    {
        const int lr = tid + i * T;
        if (lr < mp) rowsrc[c0 + pos[i]] = srcv[i];
    }
    if (tid == 0)
    {
        leafmap[0] = s_count;
        if (zeroCol != 0 && *info == 0) *info = zeroCol;
    }
}

__global__ __launch_bounds__(128) void cpanel_swap_trsm_kernel(float2* __restrict__ A, int lda, int j0, int c0,
                                                               int wact, const int* __restrict__ leafmap)
{
    __shared__ float2 s_x[64];
    const int left = c0 - j0;
    const int idx = blockIdx.x;
    const int c = idx < left ? j0 + idx : c0 + wact + (idx - left);
    float2* col = A + static_cast<size_t>(c) * lda;
    const int cnt = leafmap[0];
    const int t = threadIdx.x;
    float2 v = make_float2(0.0f, 0.0f);
    int dst = -1;
    if (t < cnt)
    {
        dst = leafmap[1 + 2 * t];
        v = col[leafmap[2 + 2 * t]];
    }
    __syncthreads();
    if (dst >= 0) col[dst] = v;
    __syncthreads();
    if (c < c0) return;
    if (t < wact) s_x[t] = col[c0 + t];
    __syncthreads();
    const float2* L = A + static_cast<size_t>(c0) * lda + c0;
    for (int s = 0; s < wact - 1; ++s)
    {
        if (t > s && t < wact) s_x[t] = SubMul(s_x[t], L[static_cast<size_t>(s) * lda + t], s_x[s]);
        __syncthreads();
    }
    if (t < wact) col[c0 + t] = s_x[t];
}

// ---------------------------------------------------------------------------------------
// Tall complex leaf over several one-wave workgroups (the design of leaf_multi_kernel in
// getrf.hip): rows [c0, m) split over P = gridDim.x workgroups of 64*R rows; per column each
// wave reduces its candidates to one 64-bit key (pivot score, then the lowest position),
// publishes its candidate's row (two tagged 64-bit words per complex value: real, imag) and
// the key, polls the P keys, fetches the winning row and eliminates with the FP64
// reciprocal and SubMul order of cleaf_kernel. The interchanges are written to leafmap for
// cpanel_swap_trsm_kernel, exactly as cleaf_kernel does. Spins are bounded; on a timeout the
// workgroup raises *err and fills its rows with NaN.
// ---------------------------------------------------------------------------------------
// Slots, per column parity: the workgroups' keys (one word each: 16-bit tag, 32-bit pivot
// score, 16-bit position rank; contiguous, so one poll of all keys touches few cache lines),
// then their candidate rows (kCRowWords words each: 32-bit tag, 32-bit value).
constexpr int kCRowWords = 64;  // 2 words per complex column (W <= 32)
constexpr int kCMaxGroups = 128;
constexpr int kCKeyWords = 2 * kCMaxGroups;  // both parities
constexpr size_t kCSlotBytes =
    (static_cast<size_t>(kCKeyWords) + 2ull * kCMaxGroups * kCRowWords) * sizeof(unsigned long long);

template <int CTRL>
__device__ inline unsigned CDppU32(unsigned v)
{
    return static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, 0xF, 0xF, false));
}
__device__ inline unsigned CWaveMaxU32(unsigned v)
{
    auto mx = [](unsigned a, unsigned b) { return a > b ? a : b; };
    v = mx(v, CDppU32<0xB1>(v));   // quad_perm [1,0,3,2]
    v = mx(v, CDppU32<0x4E>(v));   // quad_perm [2,3,0,1]
    v = mx(v, CDppU32<0x124>(v));  // row_ror:4
    v = mx(v, CDppU32<0x128>(v));  // row_ror:8
    const unsigned r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const unsigned r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return mx(mx(r0, r1), mx(r2, r3));
}
// Wave-wide maximum of a 64-bit key, reduced word by word: the high words first (one 32-bit
// DPP max per step), then, only when several lanes hold that high word, their low words
// (the keys hold a pivot score in the high word and a position in the low one, so the
// second pass runs only on exact score ties).
__device__ inline unsigned long long CWaveMax64(unsigned long long v)
{
    const unsigned hi = static_cast<unsigned>(v >> 32), lo = static_cast<unsigned>(v);
    const unsigned mh = CWaveMaxU32(hi);
    const unsigned long long tie = __ballot(hi == mh);
    unsigned ml;
    if (__popcll(tie) == 1)  // uniform
        ml = __builtin_amdgcn_readlane(lo, __builtin_ctzll(tie));
    else
        ml = CWaveMaxU32(hi == mh ? lo : 0u);
    return (static_cast<unsigned long long>(mh) << 32) | ml;
}
__device__ inline unsigned long long CLoadSlot(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool LOCAL = false>
__device__ inline void CStoreSlot(unsigned long long* p, unsigned tag, unsigned payload)
{
    __hip_atomic_store(p, (static_cast<unsigned long long>(tag) << 32) | payload, __ATOMIC_RELAXED,
                       LOCAL ? __HIP_MEMORY_SCOPE_WORKGROUP : __HIP_MEMORY_SCOPE_AGENT);
}

// FULL: wact == W (no column of the unrolled loop is skipped at run time). SPREAD: 8
// workgroups per participant, only ids = 0 mod 8 take part, all on one XCD, and the exchange
// words are stored at workgroup scope into its L2 (as leaf_multi_kernel's SPREAD).
template <int R, int W, bool FULL, bool SPREAD = false>
__global__ __launch_bounds__(64) void cleaf_multi_kernel(float2* __restrict__ A, int lda, int m, int c0, int wact,
                                                         int pivoting, int newPanel, int* __restrict__ ipiv,
                                                         int* __restrict__ info, int* __restrict__ rowsrc,
                                                         int* __restrict__ leafmap, unsigned long long* slots,
                                                         unsigned tagBase, int* err, unsigned spinLimit)
{
    constexpr int T = 64;
    static_assert(2 * W <= kCRowWords && 2 * W <= T, "slot layout");
    __shared__ float s_row[2 * W];   // this wave's candidate row (re, im interleaved)
    __shared__ float s_prow[2 * W];  // the pivot row
    __shared__ int s_bpos[W];
    __shared__ int s_count;
    const int lane = threadIdx.x;
    if (SPREAD && (blockIdx.x & 7) != 0) return;  // not a participant
    const int P = SPREAD ? static_cast<int>(gridDim.x >> 3) : static_cast<int>(gridDim.x);
    const int g = SPREAD ? static_cast<int>(blockIdx.x >> 3) : static_cast<int>(blockIdx.x);
    const int mp = m - c0;
    float2* Ap = A + static_cast<size_t>(c0) * lda + c0;
    const int rbase = g * T * R;
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;

    float2 a[R][W];
    int pos[R], srcv[R];
#pragma unroll
    for (int i = 0; i < R; ++i)
    {
        const int lr = rbase + lane + i * T;
        const bool valid = lr < mp;
#pragma unroll
        for (int j = 0; j < W; ++j)
            a[i][j] = (valid && j < wact) ? Ap[static_cast<size_t>(j) * lda + lr] : make_float2(0.0f, 0.0f);
        pos[i] = valid ? lr : INT_MAX;
        // read before any workgroup stores (stores follow the last column's exchange)
        srcv[i] = valid ? (newPanel ? c0 + lr : rowsrc[c0 + lr]) : 0;
    }
    if (lane == 0) s_count = 0;
    int zeroCol = 0;
    bool aborted = false;

#pragma unroll
    for (int j = 0; j < W; ++j)
    {
        if (FULL || j < wact)  // uniform (one wave)
        {
            const unsigned tag = tagBase + static_cast<unsigned>(j);
            unsigned long long* keys = slots + (j & 1) * kCMaxGroups;
            unsigned long long* rows = slots + kCKeyWords + static_cast<size_t>(j & 1) * kCMaxGroups * kCRowWords;
            // After a timeout the remaining columns run without waiting (no exchange, no
            // interchange): the column loop stays one straight unrolled sequence, and the
            // factors end as NaN.
            const unsigned lim = aborted ? 0u : spinLimit;
            unsigned long long best = 0ull;
            int bi = -1;
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                // (score << 32) | (0xffff - position): positions are below 16384, so a
                // candidate's key is never 0, even with score 0 (NaN)
                const bool cand = pos[i] != INT_MAX && (pivoting ? pos[i] >= j : pos[i] == j);
                const unsigned long long k =
                    cand ? (static_cast<unsigned long long>(PivotScore(a[i][j])) << 32) |
                               static_cast<unsigned>(0xffff - pos[i])
                         : 0ull;
                bi = k > best ? i : bi;
                best = k > best ? k : best;
            }
            const unsigned long long top = CWaveMax64(best);
            if (best == top && bi >= 0)
#pragma unroll
                for (int i = 0; i < R; ++i)
                    if (i == bi)
#pragma unroll
                        for (int k = 0; k < W; ++k)
                        {
                            s_row[2 * k] = a[i][k].x;
                            s_row[2 * k + 1] = a[i][k].y;
                        }
            __syncthreads();
            if (lane < 2 * wact)
                CStoreSlot<SPREAD>(rows + g * kCRowWords + lane, tag,
                                   __float_as_uint(s_row[lane < 2 * W ? lane : 0]));
            if (lane == 0)
                __hip_atomic_store(keys + g,
                                   (static_cast<unsigned long long>(tag & 0xffffu) << 48) | ((top >> 32) << 16) |
                                       (top & 0xffffull),
                                   __ATOMIC_RELAXED, SPREAD ? __HIP_MEMORY_SCOPE_WORKGROUP : __HIP_MEMORY_SCOPE_AGENT);
            // the P keys: lane l reads workgroups l and l + 64
            unsigned long long key0 = 0ull, key1 = 0ull;
            bool ok = false;
            const unsigned long long t16 = tag & 0xffffu;
            for (unsigned spins = 0; spins < lim; ++spins)
            {
                const unsigned long long w0 = lane < P ? CLoadSlot(keys + lane) : t16 << 48;
                const unsigned long long w1 = lane + 64 < P ? CLoadSlot(keys + lane + 64) : t16 << 48;
                key0 = (((w0 >> 16) & 0xffffffffull) << 32) | (w0 & 0xffffull);
                key1 = (((w1 >> 16) & 0xffffffffull) << 32) | (w1 & 0xffffull);
                if (__all((w0 >> 48) == t16 && (w1 >> 48) == t16))
                {
                    ok = true;
                    break;
                }
            }
            const unsigned long long gtop = CWaveMax64(key0 > key1 ? key0 : key1);
            const unsigned long long hit0 = __ballot(ok && lane < P && key0 == gtop);
            const unsigned long long hit1 = __ballot(ok && lane + 64 < P && key1 == gtop);
            const int winner = hit0 != 0ull   ? __ffsll(static_cast<long long>(hit0)) - 1
                               : hit1 != 0ull ? 64 + __ffsll(static_cast<long long>(hit1)) - 1
                                              : 0;
            unsigned long long rw = 0ull;
            for (unsigned spins = 0; ok; ++spins)
            {
                rw = lane < 2 * wact ? CLoadSlot(rows + winner * kCRowWords + lane)
                                     : static_cast<unsigned long long>(tag) << 32;
                if (__all(static_cast<unsigned>(rw >> 32) == tag)) break;
                if (spins >= spinLimit) ok = false;
            }
            if (!ok)
            {
                if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                aborted = true;
            }
            if (lane < 2 * W) s_prow[lane] = __uint_as_float(static_cast<unsigned>(rw));
            __syncthreads();
            const int bpos = aborted || gtop == 0ull ? j : 0xffff - static_cast<int>(gtop & 0xffffull);
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                if (pos[i] == bpos)
                    pos[i] = j;
                else if (pos[i] == j)
                    pos[i] = bpos;
            }
            if (lane == 0)
            {
                s_bpos[j] = bpos;
                if (g == 0 && ipiv != nullptr) ipiv[c0 + j] = c0 + bpos + 1;
            }
            const float2 d = make_float2(s_prow[2 * j], s_prow[2 * j + 1]);
            const bool zero = d.x == 0.0f && d.y == 0.0f;
            if (zero && zeroCol == 0 && !aborted) zeroCol = c0 + j + 1;
            const double2 rinv = Recip64(zero ? make_float2(1.0f, 0.0f) : d);
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                const bool below = pos[i] > j && pos[i] != INT_MAX && !zero;
                const float2 q = Mul64(a[i][j], rinv);
                const float2 l = below ? q : make_float2(0.0f, 0.0f);
                a[i][j] = below ? q : a[i][j];
#pragma unroll
                for (int k = 1; k < W; ++k)  // constant bounds: unrolled with the j loop
                    if (k > j) a[i][k] = SubMul(a[i][k], l, make_float2(s_prow[2 * k], s_prow[2 * k + 1]));
            }
        }
    }
    if (aborted)
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int j = 0; j < W; ++j) a[i][j] = make_float2(__builtin_nanf(""), __builtin_nanf(""));
#pragma unroll
    for (int i = 0; i < R; ++i)
    {
        const int lr = rbase + lane + i * T;
        if (lr < mp)
        {
            const int dst = pos[i];
#pragma unroll
            for (int j = 0; j < W; ++j)
                if (j < wact) Ap[static_cast<size_t>(j) * lda + dst] = a[i][j];
            rowsrc[c0 + dst] = srcv[i];
        }
    }
    if (g != 0) return;
    if (lane == 0 && zeroCol != 0 && *info == 0) *info = zeroCol;
    if (aborted)
    {
        if (lane == 0) leafmap[0] = 0;
        return;
    }
    __syncthreads();
    // workgroup 0 lists the moved rows for cpanel_swap_trsm_kernel (undoing the interchange
    // sequence from each touched position, as leaf_multi_kernel does)
    {
        int p = -1;
        if (lane < W)
            p = lane < wact ? lane : -1;
        else if (lane - W < wact)
        {
            const int jj = lane - W;
            p = s_bpos[jj] >= wact ? s_bpos[jj] : -1;
            for (int t = 0; t < jj && p >= 0; ++t)
                if (s_bpos[t] == p) p = -1;
        }
        if (p >= 0)
        {
            int x = p;
            for (int jj = wact - 1; jj >= 0; --jj)
            {
                const int b = s_bpos[jj];
                x = x == jj ? b : (x == b ? jj : x);
            }
            if (x != p)
            {
                const int slot = atomicAdd(&s_count, 1);
                leafmap[1 + 2 * slot] = c0 + p;
                leafmap[2 + 2 * slot] = c0 + x;
            }
        }
    }
    __syncthreads();
    if (lane == 0) leafmap[0] = s_count;
}

__global__ void cclear_slots_kernel(unsigned long long* slots, int words, int* err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) slots[i] = 0ull;
    if (i == 0) *err = 0;
}

__global__ __launch_bounds__(1024) void ccompact_moves_kernel(int m, int j0, const int* __restrict__ rowsrc,
                                                              int* __restrict__ moved)
{
    __shared__ int s_count;
    if (threadIdx.x == 0) s_count = 0;
    __syncthreads();
    for (int r = j0 + threadIdx.x; r < m; r += blockDim.x)
    {
        const int s = rowsrc[r];
        if (s != r)
        {
            const int slot = atomicAdd(&s_count, 1);
            moved[1 + 2 * slot] = r;
            moved[2 + 2 * slot] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) moved[0] = s_count;
}

constexpr int CNB = 128;  // complex outer block

// Applies a panel's moved-row list to the columns [lo1, hi1) and [lo2, hi2): one wave64 per
// column gathers every source row, then scatters.
__global__ __launch_bounds__(256) void couter_swap_kernel(float2* __restrict__ A, int lda, int lo1, int hi1, int lo2,
                                                          int hi2, const int* __restrict__ moved)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int idx = blockIdx.x * 4 + wave;
    const int n1 = hi1 - lo1;
    if (idx >= n1 + (hi2 - lo2)) return;
    const int c = idx < n1 ? lo1 + idx : lo2 + (idx - n1);
    float2* col = A + static_cast<size_t>(c) * lda;
    const int cnt = moved[0];
    constexpr int MAXQ = (2 * CNB + 63) / 64;
    float2 v[MAXQ];
    int d[MAXQ];
#pragma unroll
    for (int q = 0; q < MAXQ; ++q)
    {
        const int e = lane + 64 * q;
        d[q] = -1;
        if (e < cnt)
        {
            d[q] = moved[1 + 2 * e];
            v[q] = col[moved[2 + 2 * e]];
        }
    }
#pragma unroll
    for (int q = 0; q < MAXQ; ++q)
        if (d[q] >= 0) col[d[q]] = v[q];
}

// C (MxN) -= op(A) (MxK) * B (KxN), complex64, 4 real MFMA products per block.
// op(A): transA 0 = A, 1 = A^T, 2 = A^H. Tile 64x64 per workgroup, one 32x32 block per wave.
template <int TRANSA>
__global__ __launch_bounds__(256) void cgemm_sub_kernel(int M, int N, int K, const float2* __restrict__ A, int lda,
                                                        const float2* __restrict__ B, int ldb, float2* __restrict__ C,
                                                        int ldc, int tilesM)
{
    __shared__ float Ar[16][64 + 4], Ai[16][64 + 4], Br[16][64 + 4], Bi[16][64 + 4];
    const int tm = blockIdx.x % tilesM, tn = blockIdx.x / tilesM;
    const int row0 = tm * 64, col0 = tn * 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    f32x16 accR, accI;
#pragma unroll
    for (int v = 0; v < 16; ++v)
    {
        accR[v] = 0.0f;
        accI[v] = 0.0f;
    }
    for (int k0 = 0; k0 < K; k0 += 16)
    {
        // A tile: 64 rows x 16 k
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            const int e = tid + q * 256;  // 1024 elements
            int r, k;
            if (TRANSA == 0)
            {
                r = e & 63;
                k = e >> 6;
            }
            else
            {
                k = e & 15;
                r = e >> 4;
            }
            float2 v = make_float2(0.0f, 0.0f);
            if (row0 + r < M && k0 + k < K)
            {
                v = TRANSA == 0 ? A[static_cast<size_t>(k0 + k) * lda + row0 + r]
                                : A[static_cast<size_t>(row0 + r) * lda + k0 + k];
                if (TRANSA == 2) v.y = -v.y;
            }
            Ar[k][r] = v.x;
            Ai[k][r] = v.y;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            const int e = tid + q * 256;
            const int k = e & 15, j = e >> 4;
            float2 v = make_float2(0.0f, 0.0f);
            if (col0 + j < N && k0 + k < K) v = B[static_cast<size_t>(col0 + j) * ldb + k0 + k];
            Br[k][j] = v.x;
            Bi[k][j] = v.y;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk += 2)
        {
            const int k = kk + (lane >> 5);
            const float ar = Ar[k][wm * 32 + (lane & 31)], ai = Ai[k][wm * 32 + (lane & 31)];
            const float br = Br[k][wn * 32 + (lane & 31)], bi = Bi[k][wn * 32 + (lane & 31)];
            accR = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, br, accR, 0, 0, 0);
            accR = __builtin_amdgcn_mfma_f32_32x32x2f32(-ai, bi, accR, 0, 0, 0);
            accI = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, bi, accI, 0, 0, 0);
            accI = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, br, accI, 0, 0, 0);
        }
        __syncthreads();
    }
    const int j = col0 + wn * 32 + (lane & 31);
    if (j >= N) return;
    float2* cj = C + static_cast<size_t>(j) * ldc;
#pragma unroll
    for (int v = 0; v < 16; ++v)
    {
        const int i = row0 + wm * 32 + (v & 3) + 8 * (v >> 2) + 4 * (lane >> 5);
        if (i < M)
        {
            float2 c = cj[i];
            c.x -= accR[v];
            c.y -= accI[v];
            cj[i] = c;
        }
    }
}

template <int CTRL>
__device__ inline float2 QuadBroadcast2(float2 v)
{
    return make_float2(__int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v.x), CTRL, 0xF, 0xF, false)),
                       __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v.y), CTRL, 0xF, 0xF, false)));
}

// B <- op(T)^{-1} B on one diagonal block of kb <= 64 rows, complex; op: 0 = T, 1 = T^T,
// 2 = T^H. Four lanes (a quad) share one column of B, lane q holding rows 16q .. 16q+15 in
// registers; the unknown of step s is broadcast in the quad with DPP, so a lane's serial
// chain is 16 updates per step. op(T) and the FP64 reciprocals of its diagonal sit in LDS.
// Per element: the same SubMul step per term, in pivot order, and the same FP64-reciprocal
// scaling per pivot as the unblocked complex solve; rows outside the triangle are kept by
// a select.
template <bool FWD, bool UNIT, int TRANS>
__global__ __launch_bounds__(256) void ctrsm_diag64_kernel(int kb, int N, const float2* __restrict__ T, int ldt,
                                                           float2* __restrict__ B, int ldb)
{
    __shared__ float2 Ts[64][64];  // Ts[s][i] = op(T)(i, s)
    __shared__ double2 Dinv[64];
    auto opT = [&](int i, int s2)
    { return Conj(TRANS ? T[static_cast<size_t>(i) * ldt + s2] : T[static_cast<size_t>(s2) * ldt + i], TRANS == 2); };
    for (int e = threadIdx.x; e < 64 * 64; e += 256)
    {
        const int s2 = e >> 6, i = e & 63;
        Ts[s2][i] = (i < kb && s2 < kb) ? opT(i, s2) : make_float2(0.0f, 0.0f);
    }
    if (!UNIT)
        for (int t = threadIdx.x; t < kb; t += 256) Dinv[t] = Recip64(opT(t, t));
    __syncthreads();
    const int col = blockIdx.x * 64 + (threadIdx.x >> 2), q = threadIdx.x & 3;
    const bool live = col < N;
    float2* b = B + static_cast<size_t>(live ? col : 0) * ldb;
    float2 x[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) x[t] = (live && 16 * q + t < kb) ? b[16 * q + t] : make_float2(0.0f, 0.0f);
#pragma unroll
    for (int u = 0; u < 64; ++u)
    {
        const int s2 = FWD ? u : 63 - u;
        if (s2 >= kb) continue;  // uniform
        const int o = s2 >> 4, so = s2 & 15;
        float2 xs;
        if (o == 0)
            xs = QuadBroadcast2<0x00>(x[so]);
        else if (o == 1)
            xs = QuadBroadcast2<0x55>(x[so]);
        else if (o == 2)
            xs = QuadBroadcast2<0xAA>(x[so]);
        else
            xs = QuadBroadcast2<0xFF>(x[so]);
        if (!UNIT) xs = Mul64(xs, Dinv[s2]);
        if (q == o) x[so] = xs;
#pragma unroll
        for (int t = 0; t < 16; ++t)
        {
            const int row = 16 * q + t;
            const bool upd = FWD ? row > s2 : row < s2;
            const float2 nx = SubMul(x[t], Ts[s2][row], xs);
            x[t] = upd ? nx : x[t];
        }
    }
#pragma unroll
    for (int t = 0; t < 16; ++t)
        if (live && 16 * q + t < kb) b[16 * q + t] = x[t];
}

hipError_t CGemmSub(int M, int N, int K, const float2* A, int lda, int transA, const float2* B, int ldb, float2* C,
                    int ldc, hipStream_t stream)
{
    if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
    const int tilesM = (M + 63) / 64, tilesN = (N + 63) / 64;
    const dim3 grid(tilesM * tilesN);
    if (transA == 0)
        hipLaunchKernelGGL(cgemm_sub_kernel<0>, grid, dim3(256), 0, stream, M, N, K, A, lda, B, ldb, C, ldc, tilesM);
    else if (transA == 1)
        hipLaunchKernelGGL(cgemm_sub_kernel<1>, grid, dim3(256), 0, stream, M, N, K, A, lda, B, ldb, C, ldc, tilesM);
    else
        hipLaunchKernelGGL(cgemm_sub_kernel<2>, grid, dim3(256), 0, stream, M, N, K, A, lda, B, ldb, C, ldc, tilesM);
    return hipGetLastError();
}

template <bool FWD, bool UNIT, int TRANS>
hipError_t LaunchCTrsmDiag(int kb, int N, const float2* T, int ldt, float2* B, int ldb, hipStream_t stream)
{
    hipLaunchKernelGGL((ctrsm_diag64_kernel<FWD, UNIT, TRANS>), dim3((N + 63) / 64), dim3(256), 0, stream, kb, N, T,
                       ldt, B, ldb);
    return hipGetLastError();
}

hipError_t CTrsmDiag(int kb, int N, const float2* T, int ldt, bool fwd, bool unit, int trans, float2* B, int ldb,
                     hipStream_t stream)
{
#define CD_CASE(F, U, TR) \
    if (fwd == F && unit == U && trans == TR) return LaunchCTrsmDiag<F, U, TR>(kb, N, T, ldt, B, ldb, stream)
    CD_CASE(true, true, 0);
    CD_CASE(true, false, 0);
    CD_CASE(false, true, 0);
    CD_CASE(false, false, 0);
    CD_CASE(true, true, 1);
    CD_CASE(true, false, 1);
    CD_CASE(false, true, 1);
    CD_CASE(false, false, 1);
    CD_CASE(true, true, 2);
    CD_CASE(true, false, 2);
    CD_CASE(false, true, 2);
    CD_CASE(false, false, 2);
#undef CD_CASE
    return hipErrorInvalidValue;
}

// B <- op(T)^{-1} B, complex, any K: 64-row diagonal solves interleaved with complex GEMM
// updates of the rows still to be solved.
hipError_t CTrsmBlock(int K, int N, const float2* T, int ldt, bool lower, bool unit, int trans, float2* B, int ldb,
                      hipStream_t stream)
{
    if (K <= 0 || N <= 0) return hipSuccess;
    const bool fwd = lower != (trans != 0);
    const int nblk = (K + 63) / 64;
    for (int q = 0; q < nblk; ++q)
    {
        const int bi = fwd ? q : nblk - 1 - q;
        const int k0 = bi * 64, kb = std::min(64, K - k0), k1 = k0 + kb;
        hipError_t e = CTrsmDiag(kb, N, T + static_cast<size_t>(k0) * ldt + k0, ldt, fwd, unit, trans, B + k0, ldb,
                                 stream);
        if (e != hipSuccess) return e;
        if (fwd && k1 < K)
        {
            const float2* Ablk = trans ? T + static_cast<size_t>(k1) * ldt + k0 : T + static_cast<size_t>(k0) * ldt + k1;
            e = CGemmSub(K - k1, N, kb, Ablk, ldt, trans, B + k0, ldb, B + k1, ldb, stream);
        }
        else if (!fwd && k0 > 0)
        {
            const float2* Ablk = trans ? T + k0 : T + static_cast<size_t>(k0) * ldt;
            e = CGemmSub(k0, N, kb, Ablk, ldt, trans, B + k0, ldb, B, ldb, stream);
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Workspace: ints [0, m) row sources, two moved-row lists of 4 CNB + 8 (panel k uses list
// k % 2, so the look-ahead panel can build its list while the previous one is still being
// applied), the leaf's map (136); then, 8-byte aligned, the exchange slots of the
// multi-workgroup leaf and its timeout flag.
constexpr int kCLeafMultiRows = 2048;  // panels taller than this spread their leaves
constexpr int kCMovedInts = 4 * CNB + 8;
struct CWork
{
    int* rowsrc;
    int* moved[2];
    int* leafmap;
    unsigned long long* slots;
    int* err;
    unsigned* epoch;  // host counter: tag base of the next multi-workgroup leaf
};
inline size_t CWorkInts(int m) { return static_cast<size_t>(m) + 2 * kCMovedInts + 1 + 2 * 64 + 7; }
CWork CLayout(void* base, int m)
{
    int* p = static_cast<int*>(base);
    CWork w;
    w.rowsrc = p;
    w.moved[0] = p + m;
    w.moved[1] = w.moved[0] + kCMovedInts;
    w.leafmap = w.moved[1] + kCMovedInts;
    uintptr_t q = reinterpret_cast<uintptr_t>(p + CWorkInts(m));
    q = (q + 7) & ~static_cast<uintptr_t>(7);
    w.slots = reinterpret_cast<unsigned long long*>(q);
    w.err = reinterpret_cast<int*>(q + kCSlotBytes);
    w.epoch = nullptr;
    return w;
}

template <int T, int R, int W>
hipError_t LaunchCLeaf(float2* A, int lda, int m, int j0, int c0, int wact, int pivoting, int newPanel, int* ipiv,
                       int* info, const CWork& w, hipStream_t stream)
{
    hipLaunchKernelGGL((cleaf_kernel<T, R, W>), dim3(1), dim3(T), 0, stream, A, lda, m, j0, c0, wact, pivoting,
                       newPanel, ipiv, info, w.rowsrc, w.leafmap);
    return hipGetLastError();
}

int CLeafWidth(int mp)
{
    if (mp > kCLeafMultiRows) return 32;
    if (mp <= 512) return 32;
    if (mp <= 2048) return 16;
    if (mp <= 4096) return 8;
    if (mp <= 8192) return 4;
    return 2;
}

hipError_t CLeaf(float2* A, int lda, int m, int j0, int c0, int wact, int pivoting, int newPanel, int* ipiv, int* info,
                 const CWork& w, hipStream_t stream)
{
    const int mp = m - c0;
    if (mp > kCLeafMultiRows)
    {
        constexpr int R = 2, W = 32;
        const int groups = (mp + 64 * R - 1) / (64 * R);
        if (groups > kCMaxGroups || wact > W || w.epoch == nullptr) return hipErrorInvalidValue;
        // Tags: W per leaf; key words carry 16 bits of them, which must not wrap within one
        // factorization (slots cleared at its start): at most n / W leaves.
        unsigned tagBase = (*w.epoch)++ * static_cast<unsigned>(W);
        if (tagBase + W > 0xffffu) return hipErrorInvalidValue;
        unsigned spinLimit = SpinLimit();
        // Cooperative launch: the workgroups wait for each other every column, so all of
        // them must be resident at once.
        int* rowsrc = w.rowsrc;
        int* leafmap = w.leafmap;
        unsigned long long* slots = w.slots;
        int* err = w.err;
        void* args[] = {&A, &lda, &m, &c0, &wact, &pivoting, &newPanel, &ipiv, &info, &rowsrc, &leafmap, &slots,
                        &tagBase, &err, &spinLimit};
        if (wact == W && LeafSpread())
            return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&cleaf_multi_kernel<R, W, true, true>),
                                              dim3(8 * groups), dim3(64), args, 0, stream);
        const void* fn = wact == W ? reinterpret_cast<const void*>(&cleaf_multi_kernel<R, W, true>)
                                   : reinterpret_cast<const void*>(&cleaf_multi_kernel<R, W, false>);
        return hipLaunchCooperativeKernel(fn, dim3(groups), dim3(64), args, 0, stream);
    }
#define CLEAF(T, R, W) LaunchCLeaf<T, R, W>(A, lda, m, j0, c0, wact, pivoting, newPanel, ipiv, info, w, stream)
    if (mp <= 256) return CLEAF(256, 1, 32);
    if (mp <= 512) return CLEAF(512, 1, 32);
    if (mp <= 1024) return CLEAF(512, 2, 16);
    if (mp <= 2048) return CLEAF(512, 4, 16);
    if (mp <= 4096) return CLEAF(512, 8, 8);
    if (mp <= 8192) return CLEAF(512, 16, 4);
    return CLEAF(512, 32, 2);
#undef CLEAF
}

#define CRET(e)                          \
    do                                   \
    {                                    \
        hipError_t _e = (e);             \
        if (_e != hipSuccess) return _e; \
    } while (0)

}  // namespace

size_t ComplexGetrfWorkspaceBytes(int m, int n)
{
    (void)n;
    return CWorkInts(std::max(m, 1)) * sizeof(int) + 8 + kCSlotBytes + 16;
}

hipError_t ComplexGetrf(int m, int n, float2* a, int lda, int* ipiv, int* info, void* work, hipStream_t stream,
                        const LookAhead* la)
{
    if (m > 16384) return hipErrorInvalidValue;
    CWork w = CLayout(work, m);
    unsigned epoch = 1;
    w.epoch = &epoch;
    FillInfo(info, 1, 0, stream);
    if (m > kCLeafMultiRows)
    {
        constexpr int words = static_cast<int>(kCSlotBytes / sizeof(unsigned long long));
        hipLaunchKernelGGL(cclear_slots_kernel, dim3((words + 255) / 256), dim3(256), 0, stream, w.slots, words, w.err);
    }
    const int kmax = std::min(m, n);
    if (kmax == 0) return hipGetLastError();
    const int pivoting = ipiv != nullptr;
    const int npanel = (kmax + CNB - 1) / CNB;

    // Look-ahead, as in the real factorization: the caller's stream s0 carries the critical
    // path (the next block column's update and its panel), the side stream s1 the rest of
    // each trailing update.
    const bool two = la != nullptr && la->side != nullptr && npanel > 1;
    hipStream_t s0 = stream, s1 = two ? la->side : stream;
    if (two)
    {
        CRET(hipEventRecord(la->fork, stream));
        CRET(hipStreamWaitEvent(s1, la->fork, 0));
    }

    // Panel k: its leaves (each followed by the update of the panel's remaining columns),
    // then the list of its interchanges.
    auto panel = [&](int k) -> hipError_t
    {
        const int j0 = k * CNB, nb = std::min(CNB, kmax - j0), panelEnd = j0 + nb;
        for (int c0 = j0; c0 < panelEnd;)
        {
            const int W = CLeafWidth(m - c0);
            const int wact = std::min(W, panelEnd - c0);
            CRET(CLeaf(a, lda, m, j0, c0, wact, pivoting, c0 == j0, ipiv, info, w, s0));
            const int others = nb - wact;
            if (others > 0)
            {
                if (pivoting)
                    hipLaunchKernelGGL(cpanel_swap_trsm_kernel, dim3(others), dim3(128), 0, s0, a, lda, j0, c0, wact,
                                       w.leafmap);
                else if (panelEnd - c0 - wact > 0)
                    CRET(CTrsmBlock(wact, panelEnd - c0 - wact, a + static_cast<size_t>(c0) * lda + c0, lda, true,
                                    true, 0, a + static_cast<size_t>(c0 + wact) * lda + c0, lda, s0));
                CRET(hipGetLastError());
                const int right = panelEnd - c0 - wact;
                if (right > 0 && m - c0 - wact > 0)
                    CRET(CGemmSub(m - c0 - wact, right, wact, a + static_cast<size_t>(c0) * lda + c0 + wact, lda, 0,
                                  a + static_cast<size_t>(c0 + wact) * lda + c0, lda,
                                  a + static_cast<size_t>(c0 + wact) * lda + c0 + wact, lda, s0));
            }
            c0 += wact;
        }
        if (pivoting)
            hipLaunchKernelGGL(ccompact_moves_kernel, dim3(1), dim3(1024), 0, s0, m, j0, w.rowsrc, w.moved[k & 1]);
        return hipGetLastError();
    };
    // Applies panel k to the columns [lo, hi) (interchanges, U12 solve, trailing product)
    // and, when leftSwaps, its interchanges to the columns left of the panel.
    auto update = [&](int k, int lo, int hi, bool leftSwaps, hipStream_t st) -> hipError_t
    {
        const int j0 = k * CNB, nb = std::min(CNB, kmax - j0), pe = j0 + nb;
        const int cols = hi - lo, left = leftSwaps ? j0 : 0;
        if (pivoting && cols + left > 0)
            hipLaunchKernelGGL(couter_swap_kernel, dim3((cols + left + 3) / 4), dim3(256), 0, st, a, lda, 0, left, lo,
                               hi, w.moved[k & 1]);
        CRET(hipGetLastError());
        if (cols <= 0) return hipSuccess;
        CRET(CTrsmBlock(nb, cols, a + static_cast<size_t>(j0) * lda + j0, lda, true, true, 0,
                        a + static_cast<size_t>(lo) * lda + j0, lda, st));
        if (m - pe <= 0) return hipSuccess;
        ProfileScope prof(ProfileKernel::ComplexGemm, 8.0 * (m - pe) * cols * nb, st);
        return CGemmSub(m - pe, cols, nb, a + static_cast<size_t>(j0) * lda + pe, lda, 0,
                        a + static_cast<size_t>(lo) * lda + j0, lda, a + static_cast<size_t>(lo) * lda + pe, lda, st);
    };

    CRET(panel(0));
    if (two) CRET(hipEventRecord(la->panel[0], s0));
    for (int k = 0; k < npanel; ++k)
    {
        const int pe = std::min(kmax, (k + 1) * CNB);
        const int nbNext = k + 1 < npanel ? std::min(CNB, kmax - pe) : 0;
        // Bulk of the trailing update (s1): everything right of the next block column, and
        // the interchanges of panel k left of it.
        if (two) CRET(hipStreamWaitEvent(s1, la->panel[k & 1], 0));
        CRET(update(k, pe + nbNext, n, true, s1));
        if (two) CRET(hipEventRecord(la->bulk[k & 1], s1));
        // Critical path (s0): the next block column, then its panel.
        if (nbNext > 0)
        {
            if (two && k > 0) CRET(hipStreamWaitEvent(s0, la->bulk[(k - 1) & 1], 0));
            CRET(update(k, pe, pe + nbNext, false, s0));
            CRET(panel(k + 1));
            if (two) CRET(hipEventRecord(la->panel[(k + 1) & 1], s0));
        }
    }
    if (two) CRET(hipStreamWaitEvent(s0, la->bulk[(npanel - 1) & 1], 0));
    if (m > kCLeafMultiRows) FoldWaitError(w.err, info, s0);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// cgetrs (Device, column-major), trans 0 / 1 / 2
// ---------------------------------------------------------------------------------------
namespace
{
template <bool GATHER>
__global__ void cpermute_kernel(int n, int nrhs, const int* __restrict__ perm, const float2* __restrict__ src, int lds,
                                float2* __restrict__ dst, int ldd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n) return;
    const int p = perm ? perm[i] : i;
    if (GATHER)
        dst[static_cast<size_t>(c) * ldd + i] = src[static_cast<size_t>(c) * lds + p];
    else
        dst[static_cast<size_t>(c) * ldd + p] = src[static_cast<size_t>(c) * lds + i];
}
}  // namespace

size_t ComplexGetrsWorkspaceBytes(int n, int nrhs)
{
    return static_cast<size_t>(n) * std::max(nrhs, 1) * sizeof(float2) + (static_cast<size_t>(n) + 64) * sizeof(int);
}

hipError_t ComplexGetrs(int trans, int n, int nrhs, const float2* lu, int lda, const int* ipiv, float2* b, int ldb,
                        void* work, hipStream_t stream)
{
    if (n <= 0 || nrhs <= 0) return hipSuccess;
    if (n > 16384) return hipErrorInvalidValue;
    float2* x = static_cast<float2*>(work);
    int* perm = reinterpret_cast<int*>(x + static_cast<size_t>(n) * nrhs);
    if (ipiv) CRET(IpivToPerm(n, ipiv, perm, stream, true));
    const int* p = ipiv ? perm : nullptr;
    const dim3 grid((n + 255) / 256, nrhs);
    const int NBS = 128;
    if (trans == 0)
    {
        hipLaunchKernelGGL(cpermute_kernel<true>, grid, dim3(256), 0, stream, n, nrhs, p, b, ldb, x, n);
        for (int k = 0; k < n; k += NBS)
        {
            const int kb = std::min(NBS, n - k);
            CRET(CTrsmBlock(kb, nrhs, lu + static_cast<size_t>(k) * lda + k, lda, true, true, 0, x + k, n, stream));
            if (n - k - kb > 0)
                CRET(CGemmSub(n - k - kb, nrhs, kb, lu + static_cast<size_t>(k) * lda + k + kb, lda, 0, x + k, n,
                              x + k + kb, n, stream));
        }
        for (int k = ((n - 1) / NBS) * NBS; k >= 0; k -= NBS)
        {
            const int kb = std::min(NBS, n - k);
            CRET(CTrsmBlock(kb, nrhs, lu + static_cast<size_t>(k) * lda + k, lda, false, false, 0, x + k, n, stream));
            if (k > 0)
                CRET(CGemmSub(k, nrhs, kb, lu + static_cast<size_t>(k) * lda, lda, 0, x + k, n, x, n, stream));
        }
        hipLaunchKernelGGL(cpermute_kernel<true>, grid, dim3(256), 0, stream, n, nrhs, nullptr, x, n, b, ldb);
    }
    else
    {
        hipLaunchKernelGGL(cpermute_kernel<true>, grid, dim3(256), 0, stream, n, nrhs, nullptr, b, ldb, x, n);
        for (int k = 0; k < n; k += NBS)  // op(U) w = b, op(U) lower
        {
            const int kb = std::min(NBS, n - k);
            CRET(CTrsmBlock(kb, nrhs, lu + static_cast<size_t>(k) * lda + k, lda, false, false, trans, x + k, n,
                            stream));
            if (n - k - kb > 0)
                CRET(CGemmSub(n - k - kb, nrhs, kb, lu + static_cast<size_t>(k + kb) * lda + k, lda, trans, x + k, n,
                              x + k + kb, n, stream));
        }
        for (int k = ((n - 1) / NBS) * NBS; k >= 0; k -= NBS)  // op(L) z = w, op(L) upper unit
        {
            const int kb = std::min(NBS, n - k);
            CRET(CTrsmBlock(kb, nrhs, lu + static_cast<size_t>(k) * lda + k, lda, true, true, trans, x + k, n,
                            stream));
            if (k > 0) CRET(CGemmSub(k, nrhs, kb, lu + k, lda, trans, x + k, n, x, n, stream));
        }
        hipLaunchKernelGGL(cpermute_kernel<false>, grid, dim3(256), 0, stream, n, nrhs, p, x, n, b, ldb);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Batched complex inverses, n <= 64: one wave64 group of NL lanes per matrix, lane = row.
// ---------------------------------------------------------------------------------------
namespace
{
template <int NL>
__device__ inline void CGroupArgmax(uint32_t val, int key, uint32_t* bestVal, int* bestKey)
{
    uint32_t b = val;
#pragma unroll
    for (int off = NL / 2; off > 0; off >>= 1) b = max(b, static_cast<uint32_t>(__shfl_xor(b, off)));
    int k = (val == b) ? key : INT_MAX;
#pragma unroll
    for (int off = NL / 2; off > 0; off >>= 1) k = min(k, __shfl_xor(k, off));
    *bestVal = b;
    *bestKey = k;
}

// Column k of a register row (k uniform but not a compile-time constant): a select chain,
// so register arrays stay in registers when the loop over k is not unrolled.
template <int NL>
__device__ inline float2 GetCol(const float2 (&S)[NL], int k)
{
    float2 v = S[0];
#pragma unroll
    for (int j = 1; j < NL; ++j) v = (j == k) ? S[j] : v;
    return v;
}
template <int NL>
__device__ inline void SetCol(float2 (&S)[NL], int k, float2 v)
{
#pragma unroll
    for (int j = 0; j < NL; ++j) S[j] = (j == k) ? v : S[j];
}
// Up to 32 columns the loops over k are unrolled (k becomes a constant and the selects
// fold away); at 64 they stay loops, which keeps those kernels small enough to compile
// quickly and to fit the instruction cache.

// Gauss-Jordan with implicit pivoting (see batched.hip for the real derivation).
template <int NL>
__global__ __launch_bounds__(256) void cmatinv_gj_kernel(int n, const float2* const* __restrict__ Aarr,
                                                         const float2* __restrict__ Abase, int lda, int64_t strideA,
                                                         float2* const* __restrict__ Carr, float2* __restrict__ Cbase,
                                                         int ldc, int64_t strideC, int* __restrict__ info, int batch)
{
    constexpr int GROUPS = 256 / NL;
    __shared__ float2 s_row[GROUPS][NL];
    __shared__ int s_inv[GROUPS][NL];
    const int tid = threadIdx.x;
    const int g = tid / NL, lane = tid % NL;
    const int mat = blockIdx.x * GROUPS + g;
    const bool live = mat < batch;
    const float2* A = nullptr;
    float2* C = nullptr;
    if (live)
    {
        A = Aarr ? Aarr[mat] : Abase + static_cast<int64_t>(mat) * strideA;
        C = Carr ? Carr[mat] : Cbase + static_cast<int64_t>(mat) * strideC;
    }
    const bool rowLive = live && lane < n;
    float2 S[NL];
#pragma unroll
    for (int j = 0; j < NL; ++j)
        S[j] = (rowLive && j < n) ? A[static_cast<int64_t>(j) * lda + lane] : make_float2(0.0f, 0.0f);
    double2 sigma = make_double2(1.0, 0.0);
    int pos = lane;
    bool used = false;
    int zero = 0;
    // Slot j of S holds column (j + steps done) % NL: each step pivots slot 0 and rotates the
    // slots down by one (the processed column moves to the last slot), so no slot is ever
    // addressed by a run-time index, also where the step loop is not unrolled.
#pragma unroll(NL <= 32 ? NL : 1)
    for (int k = 0; k < NL; ++k)
    {
        if (k < n)
        {
            const float2 sk = S[0];
            const uint32_t cand = (rowLive && !used) ? PivotScore(sk) : 0u;
            uint32_t best;
            int key;
            CGroupArgmax<NL>(rowLive && !used ? cand : 0u, (rowLive && !used) ? ((pos << 8) | lane) : INT_MAX,
                             &best, &key);
            const int p = key & 0xFF;
            const int ppos = key >> 8;
            const bool isP = lane == p;
            if (pos == k) pos = ppos;
            if (isP) pos = k;
            if (isP)
            {
#pragma unroll
                for (int j = 0; j < NL; ++j) s_row[g][j] = S[j];
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            const float2 d = s_row[g][0];
            const bool z = d.x == 0.0f && d.y == 0.0f;
            if (z && zero == 0) zero = k + 1;
            const double2 r = Recip64(z ? make_float2(1.0f, 0.0f) : d);
            const float2 c = (isP || z) ? make_float2(0.0f, 0.0f) : Mul64(sk, r);
            const float2 last = z ? sk : (isP ? make_float2(1.0f, 0.0f) : make_float2(-c.x, -c.y));
#pragma unroll
            for (int j = 1; j < NL; ++j) S[j - 1] = SubMul(S[j], c, s_row[g][j]);
            S[NL - 1] = last;
            if (!z && isP) sigma = r;
            if (isP) used = true;
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (rowLive) s_inv[g][pos] = lane;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (rowLive)
    {
#pragma unroll
        for (int kk = 0; kk < NL; ++kk)
        {
            const int j = kk + n < NL ? kk + n : kk + n - NL;  // the column slot kk holds
            if (j < n) C[static_cast<int64_t>(s_inv[g][j]) * ldc + pos] = Mul64(S[kk], sigma);
        }
    }
    if (live && lane == 0) info[mat] = zero;
}

template <int NL>
__global__ __launch_bounds__(256) void cgetri_small_kernel(int n, const float2* const* __restrict__ Aarr, int lda,
                                                           const int* __restrict__ piv,
                                                           float2* const* __restrict__ Carr, int ldc,
                                                           int* __restrict__ info, int batch)
{
    constexpr int GROUPS = 256 / NL;
    __shared__ float2 s_row[GROUPS][NL];
    __shared__ int s_perm[GROUPS][NL];
    const int tid = threadIdx.x;
    const int g = tid / NL, lane = tid % NL;
    const int mat = blockIdx.x * GROUPS + g;
    const bool live = mat < batch;
    const bool rowLive = live && lane < n;
    const float2* A = live ? Aarr[mat] : nullptr;
    float2* C = live ? Carr[mat] : nullptr;
    // The factor entries of this lane's row are read from memory one column per step (the
    // next column is fetched a step ahead), so no register array is indexed by the step.
    auto factor = [&](int q) -> float2
    { return (rowLive && q >= 0 && q < n) ? A[static_cast<int64_t>(q) * lda + lane] : make_float2(0.0f, 0.0f); };
    if (live && lane == 0)
    {
        for (int i = 0; i < n; ++i) s_perm[g][i] = i;
        for (int k = 0; k < n; ++k)
        {
            int p = piv ? piv[static_cast<int64_t>(mat) * n + k] - 1 : k;
            p = p < k || p >= n ? k : p;
            const int t = s_perm[g][k];
            s_perm[g][k] = s_perm[g][p];
            s_perm[g][p] = t;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    float2 X[NL];
    const int pr = rowLive ? s_perm[g][lane] : -1;
#pragma unroll
    for (int j = 0; j < NL; ++j) X[j] = make_float2(j == pr ? 1.0f : 0.0f, 0.0f);
    float2 fnext = factor(0);
#pragma unroll(NL <= 32 ? NL : 1)
    for (int q = 0; q < NL; ++q)
    {
        if (q < n)
        {
            const float2 fq = fnext;
            fnext = factor(q + 1);
            if (lane == q)
            {
#pragma unroll
                for (int j = 0; j < NL; ++j) s_row[g][j] = X[j];
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (lane > q && rowLive)
            {
                const float2 l = fq;
#pragma unroll
                for (int j = 0; j < NL; ++j) X[j] = SubMul(X[j], l, s_row[g][j]);
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    fnext = factor(n - 1);
#pragma unroll(NL <= 32 ? NL : 1)
    for (int q = NL - 1; q >= 0; --q)
    {
        if (q < n)
        {
            const float2 fq = fnext;
            fnext = factor(q - 1);
            if (lane == q)
            {
                const float2 d = fq;
                const bool z = d.x == 0.0f && d.y == 0.0f;
                const double2 r = Recip64(z ? make_float2(1.0f, 0.0f) : d);
#pragma unroll
                for (int j = 0; j < NL; ++j)
                {
                    if (!z) X[j] = Mul64(X[j], r);
                    s_row[g][j] = X[j];
                }
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (lane < q && rowLive)
            {
                const float2 uq = fq;
#pragma unroll
                for (int j = 0; j < NL; ++j) X[j] = SubMul(X[j], uq, s_row[g][j]);
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    const float2 diag = factor(lane);
    const unsigned long long zmask = __ballot(rowLive && diag.x == 0.0f && diag.y == 0.0f);
    const int shift = (g * NL) % 64;
    const unsigned long long gmask = (NL >= 64) ? ~0ull : ((~0ull >> (64 - NL)) << shift);
    const unsigned long long mine = zmask & gmask;
    const int zero = mine ? (__ffsll(static_cast<long long>(mine)) - 1 - shift) + 1 : 0;
    if (rowLive)
    {
#pragma unroll
        for (int j = 0; j < NL; ++j)
            if (j < n) C[static_cast<int64_t>(j) * ldc + lane] = X[j];
    }
    if (live && lane == 0) info[mat] = zero;
}

template <int NL>
hipError_t LaunchCMatinv(int n, const float2* const* Aarr, const float2* Abase, int lda, int64_t strideA,
                         float2* const* Carr, float2* Cbase, int ldc, int64_t strideC, int* info, int batch,
                         hipStream_t stream)
{
    constexpr int GROUPS = 256 / NL;
    hipLaunchKernelGGL(cmatinv_gj_kernel<NL>, dim3((batch + GROUPS - 1) / GROUPS), dim3(256), 0, stream, n, Aarr,
                       Abase, lda, strideA, Carr, Cbase, ldc, strideC, info, batch);
    return hipGetLastError();
}

template <int NL>
hipError_t LaunchCGetri(int n, const float2* const* Aarr, int lda, const int* piv, float2* const* Carr, int ldc,
                        int* info, int batch, hipStream_t stream)
{
    constexpr int GROUPS = 256 / NL;
    hipLaunchKernelGGL(cgetri_small_kernel<NL>, dim3((batch + GROUPS - 1) / GROUPS), dim3(256), 0, stream, n, Aarr,
                       lda, piv, Carr, ldc, info, batch);
    return hipGetLastError();
}
}  // namespace

hipError_t CMatinvBatched(int n, const float2* const* Aarr, const float2* Abase, int lda, int64_t strideA,
                          float2* const* Carr, float2* Cbase, int ldc, int64_t strideC, int* info, int batch,
                          hipStream_t stream)
{
    if (batch <= 0) return hipSuccess;
    if (n <= 8) return LaunchCMatinv<8>(n, Aarr, Abase, lda, strideA, Carr, Cbase, ldc, strideC, info, batch, stream);
    if (n <= 16)
        return LaunchCMatinv<16>(n, Aarr, Abase, lda, strideA, Carr, Cbase, ldc, strideC, info, batch, stream);
    if (n <= 32)
        return LaunchCMatinv<32>(n, Aarr, Abase, lda, strideA, Carr, Cbase, ldc, strideC, info, batch, stream);
    if (n <= 64)
        return LaunchCMatinv<64>(n, Aarr, Abase, lda, strideA, Carr, Cbase, ldc, strideC, info, batch, stream);
    return hipErrorNotSupported;  // 64 < n: the caller takes the factor-and-solve route
}

hipError_t CGetriSmallBatched(int n, const float2* const* Aarr, int lda, const int* piv, float2* const* Carr, int ldc,
                              int* info, int batch, hipStream_t stream)
{
    if (batch <= 0) return hipSuccess;
    if (n <= 8) return LaunchCGetri<8>(n, Aarr, lda, piv, Carr, ldc, info, batch, stream);
    if (n <= 16) return LaunchCGetri<16>(n, Aarr, lda, piv, Carr, ldc, info, batch, stream);
    if (n <= 32) return LaunchCGetri<32>(n, Aarr, lda, piv, Carr, ldc, info, batch, stream);
    if (n <= 64) return LaunchCGetri<64>(n, Aarr, lda, piv, Carr, ldc, info, batch, stream);
    return hipErrorNotSupported;
}

}  // namespace ops
