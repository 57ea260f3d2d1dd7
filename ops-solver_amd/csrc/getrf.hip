// Real LU factorization (sgetrf) and the Device column-major solve, gfx950.
//
// Algorithm: right-looking blocked LU with partial pivoting, outer block width NB.
// Reference: src/sgetrf/real_lu_blocked_kernel.cpp:1469-1499 (BlockedFactorLaunch::Run:
// panel, swap+TRSM, trailing product per block). The MI355X design:
//   * panel: a sequence of register-resident leaf factorizations (leaf_getf2_kernel), one
//     workgroup of 256/1024 threads holding every active row of W panel columns in VGPRs.
//     Per column the pivot is a 64-bit (|a|, position) key argmax: DPP within the wave,
//     then the wave winners through LDS with one barrier; the row swap, the scaling and
//     the rank-1 update never leave the CU. Pivot rule, zero-pivot info and reciprocal scaling follow
//     src/sgetrf/real_lu_blocked_kernel.cpp:342-389, :467-482 and
//     src/utils/kernel/real_lu.hpp:419-437.
//   * the panel is halved recursively down to leaves; each leaf applies its row
//     interchanges to the other panel columns itself (a gather/scatter of the <= 2W moved
//     rows), and every split solves the right half's top rows with the left half's L11
//     and updates the rest of the right half with one MFMA product.
//   * after the panel, the composed interchanges of all NB pivots are applied to every
//     column outside the panel (one coalesced gather/scatter per column), then the NB x NB
//     unit-lower TRSM on the U12 block and the MFMA trailing update A22 -= L21 * U12.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cstdlib>
#include <type_traits>

#include "lu_kernels.h"
#include "trsv.h"
#include "tslu.h"
#include "wave.h"

namespace ops
{
namespace
{

constexpr int kMaxRows = 16384;  // largest m the register-resident leaf supports

// Workspace layout (ints): [0, m) panel row sources; two moved-row lists of 4*NBMAX + 8
// (panel k uses list k % 2, so the look-ahead panel can build its list while the previous
// one is still being applied); then the global permutation (m ints).
constexpr int NBMAX = 512;
constexpr int kMovedInts = 4 * NBMAX + 8;
// After the ints (8-byte aligned): the exchange slots of the multi-workgroup leaf and its
// timeout flag. Slots are tagged 64-bit words, per column parity: the workgroups' keys
// (one word each: 16-bit tag, 32-bit magnitude, 16-bit position rank; contiguous, so one
// poll of all keys touches few cache lines), then their candidate rows (kRowWords words
// each: 32-bit tag, 32-bit value).
constexpr int kRowWords = 64;        // per workgroup: up to 64 row words
constexpr int kMaxLeafGroups = 128;  // workgroups of one multi-workgroup leaf
constexpr int kKeyWords = 2 * kMaxLeafGroups;  // both parities
constexpr size_t kSlotBytes =
    (static_cast<size_t>(kKeyWords) + 2ull * kMaxLeafGroups * kRowWords) * sizeof(unsigned long long);
constexpr size_t kLinvPanelBytes = (NBMAX / 64) * 4096 * sizeof(float);  // block inverses of one panel's L11
struct Work
{
    int* rowsrc;
    int* moved[2];
    int* movedAll;  // every panel's moved-row list (deferred interchanges), kMovedInts each
    int* composed;  // m ints: composed row map of the deferred interchanges
    int* perm;
    unsigned long long* slots;
    int* err;
    unsigned* epoch;       // host counter: tag base of the next multi-workgroup leaf
    float* linvNode;       // block inverses for the panel-internal solves (panel stream only)
    float* linvPanel[2];   // block inverses of panel k's L11 in slot k % 2 (both streams read them)
    void* tslu;            // scratch of the tournament leaves
    bool tournament;       // tall leaves pivot by tournament (TsluLeaf)
};
// Panels of one factorization are at least 128 columns wide.
inline size_t MaxPanels(int m) { return static_cast<size_t>(m) / 128 + 2; }
inline size_t WorkInts(int m) { return static_cast<size_t>(m) * 3 + (2 + MaxPanels(m)) * kMovedInts; }
inline Work Layout(void* base, int m)
{
    int* p = static_cast<int*>(base);
    Work w;
    w.rowsrc = p;
    w.moved[0] = p + m;
    w.moved[1] = w.moved[0] + kMovedInts;
    w.perm = w.moved[1] + kMovedInts;
    w.composed = w.perm + m;
    w.movedAll = w.composed + m;
    uintptr_t q = reinterpret_cast<uintptr_t>(p + WorkInts(m));
    q = (q + 7) & ~static_cast<uintptr_t>(7);
    w.slots = reinterpret_cast<unsigned long long*>(q);
    w.err = reinterpret_cast<int*>(q + kSlotBytes);
    w.epoch = nullptr;
    q = (q + kSlotBytes + 16 + 255) & ~static_cast<uintptr_t>(255);
    w.linvNode = reinterpret_cast<float*>(q);
    w.linvPanel[0] = reinterpret_cast<float*>(q + kLinvPanelBytes);
    w.linvPanel[1] = reinterpret_cast<float*>(q + 2 * kLinvPanelBytes);
    w.tslu = reinterpret_cast<void*>(q + 3 * kLinvPanelBytes);
    w.tournament = false;
    return w;
}

// ---------------------------------------------------------------------------------------
// Leaf factorization: columns [c0, c0 + wact) of rows [c0, m), T threads, R rows per
// thread (row tid + i*T), W >= wact columns in registers. Pivoting is implicit: each row
// tracks the position it occupies under the LAPACK interchange sequence, the pivot of
// column j is the candidate (position >= j) with the largest |a| and, on ties, the lowest
// position -- the first-row rule of a physical swap. Per column: one 64-bit key argmax
// per thread, a DPP wave argmax, and one barrier, after which every thread reads the
// wave winners' keys and the chosen candidate's row from LDS (double-buffered by column
// parity, so one barrier per column suffices). At the end each row is stored at its final
// position, which applies the interchanges to the leaf columns.
// ---------------------------------------------------------------------------------------
// FULL: wact == W, so no column of the unrolled loop is skipped at run time (a skip path
// would make the compiler keep two copies of the row registers alive per column).
template <int T, int R, int W, bool FULL>
__global__ __launch_bounds__(T) void leaf_getf2_kernel(float* __restrict__ A, int lda, int m, int j0, int c0, int wact,
                                                       int panelEnd, int pivoting, int newPanel,
                                                       int* __restrict__ ipiv, int* __restrict__ info,
                                                       int* __restrict__ rowsrc)
{
    constexpr int NW = T / 64;
    constexpr int MOVES = 2 * W;  // a leaf moves at most 2 * wact rows
    __shared__ unsigned long long s_key[2][NW];
    __shared__ float s_row[2][NW][W];
    __shared__ int s_count;
    __shared__ int s_mdst[MOVES], s_msrc[MOVES];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int mp = m - c0;
    float* Ap = A + static_cast<size_t>(c0) * lda + c0;

    float a[R][W];
    int pos[R];
#pragma unroll
    for (int i = 0; i < R; ++i)
    {
        const int lr = tid + i * T;
        const bool valid = lr < mp;
#pragma unroll
        for (int j = 0; j < W; ++j) a[i][j] = (valid && j < wact) ? Ap[static_cast<size_t>(j) * lda + lr] : 0.0f;
        pos[i] = valid ? lr : -1;  // padding rows: never a candidate, never below a pivot
    }
    if (tid == 0) s_count = 0;
    int zeroCol = 0;

#pragma unroll
    for (int j = 0; j < W; ++j)
    {
        if (FULL || j < wact)  // uniform
        {
            const int par = j & 1;
            // Candidate of this thread: its best row (or, without pivoting, the row at j),
            // as a 64-bit key (|a| bits + 1, then ~position, so the lowest position wins
            // magnitude ties). NaN: on the diagonal position it is the pivot (0x7fffffff),
            // elsewhere it never wins (0) -- src/utils/kernel/real_lu.hpp:59-77, :100-109.
            unsigned long long best = 0ull;
            int bi = -1;
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                const bool cand = pivoting ? pos[i] >= j : pos[i] == j;
                const unsigned ab = __float_as_uint(fabsf(a[i][j]));
                const unsigned mag = ab <= 0x7f800000u ? ab + 1u : (pos[i] == j ? 0x7fffffffu : 0u);
                const unsigned long long k =
                    cand ? (1ull << 63) | (static_cast<unsigned long long>(mag) << 32) | static_cast<unsigned>(~pos[i])
                         : 0ull;
                bi = k > best ? i : bi;
                best = k > best ? k : best;
            }
            const unsigned long long wmax = WaveMaxU64Split(best);
            if (best == wmax && bi >= 0)  // the wave's winner (keys are unique per position)
            {
                s_key[par][wid] = wmax;
                // Selected by a bit of a mask, not by i == bi: the compiler would turn the
                // equality into a[bi][k], a run-time row index that puts a[][] in scratch.
                const unsigned bsel = 1u << bi;
#pragma unroll
                for (int k = 0; k < W; ++k)
                {
                    float v = a[0][k];
#pragma unroll
                    for (int i = 1; i < R; ++i) v = (bsel >> i) & 1u ? a[i][k] : v;
                    s_row[par][wid][k] = v;
                }
            }
            if (lane == 0 && wmax == 0ull) s_key[par][wid] = 0ull;
            __syncthreads();
            unsigned long long top = s_key[par][0];
            int tw = 0;
#pragma unroll
            for (int w = 1; w < NW; ++w)
            {
                const unsigned long long kw = s_key[par][w];
                if (kw > top)
                {
                    top = kw;
                    tw = w;
                }
            }
            const int bpos = top == 0ull ? j : static_cast<int>(~static_cast<unsigned>(top & 0xffffffffull));
            float prow[W];
#pragma unroll
            for (int k = 0; k < W; ++k) prow[k] = s_row[par][tw][k];
            // Positions j and bpos are exchanged.
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                if (pos[i] == bpos)
                    pos[i] = j;
                else if (pos[i] == j)
                    pos[i] = bpos;
            }
            if (tid == 0 && ipiv != nullptr) ipiv[c0 + j] = c0 + bpos + 1;
            // The pivot is the same in every thread: made uniform, the reciprocal / division
            // choice is a scalar branch and the division sequence runs only for a subnormal
            // pivot.
            const float d = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(prow[j])));
            if (d == 0.0f && zeroCol == 0) zeroCol = c0 + j + 1;
            // Branch-free update: rows that are not below the pivot take l = 0, for which
            // the FMA leaves them unchanged (a zero pivot skips the column likewise).
            const bool recip = fabsf(d) >= FLT_MIN;  // uniform
            const float r = 1.0f / d;
            float q[R];
            if (recip)
#pragma unroll
                for (int i = 0; i < R; ++i) q[i] = a[i][j] * r;
            else
#pragma unroll
                for (int i = 0; i < R; ++i) q[i] = a[i][j] / d;
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                const bool below = pos[i] > j && d != 0.0f;
                const float l = below ? q[i] : 0.0f;
                a[i][j] = below ? q[i] : a[i][j];
#pragma unroll
                for (int k = 1; k < W; ++k)  // constant bounds: unrolled with the j loop
                    if (k > j) a[i][k] = fmaf(-l, prow[k], a[i][k]);
            }
        }
    }

    // Store each row at its final position and record where the data came from.
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
                s_mdst[slot] = c0 + dst;
                s_msrc[slot] = c0 + lr;
            }
        }
    }
    // Panel-level sources: read before any write (rows only exchange among themselves).
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
    if (tid == 0 && zeroCol != 0 && *info == 0) *info = zeroCol;

    // The same interchanges on the other panel columns [j0, c0) and [c0 + wact, panelEnd):
    // thread t takes moved entry t % MOVES of column t / MOVES of each batch of T / MOVES
    // columns; every entry of a batch is gathered before any is scattered (rows only
    // exchange among themselves).
    const int cnt = s_count;
    const int left = c0 - j0, others = left + (panelEnd - c0 - wact);
    constexpr int CPB = T / MOVES;  // columns per batch
    constexpr int SWB = 8;          // batches gathered per barrier
    const int e = tid % MOVES;
    for (int b0 = 0; b0 < others; b0 += CPB * SWB)  // uniform trip count
    {
        float v[SWB];
        int cc[SWB];
#pragma unroll
        for (int bb = 0; bb < SWB; ++bb)
        {
            const int q = b0 + bb * CPB + tid / MOVES;
            cc[bb] = -1;
            v[bb] = 0.0f;
            if (q < others && e < cnt)
            {
                cc[bb] = q < left ? j0 + q : c0 + wact + (q - left);
                v[bb] = A[static_cast<size_t>(cc[bb]) * lda + s_msrc[e]];
            }
        }
        __syncthreads();
#pragma unroll
        for (int bb = 0; bb < SWB; ++bb)
            if (cc[bb] >= 0) A[static_cast<size_t>(cc[bb]) * lda + s_mdst[e]] = v[bb];
    }
}

// ---------------------------------------------------------------------------------------
// Tall leaf over several workgroups: rows [c0, m) are split over P = gridDim.x one-wave
// workgroups of 64*R consecutive rows, so each CU does 1/P of the per-column work and no
// workgroup barrier sits on the per-column path. Per column every wave reduces its
// candidates to one 64-bit key (DPP) and publishes the key, then its candidate's W values,
// as tagged 64-bit words (relaxed agent-scope atomic stores; tag = this launch's base +
// column, so each word is valid on its own and no fence or counter is needed); then it
// polls the P keys, takes the maximum and reads the winning row. The update of columns
// beyond j+1 by pivot j is deferred to the next column, where it runs between publishing
// the key and publishing the row, i.e. while the peers' keys are in flight; per element the
// updates still come in pivot order, so the result equals the unblocked elimination.
// Slots alternate with the column parity: a workgroup publishes column j+2 only after all
// have published j+1, i.e. after all have finished reading column j. Pivot choice, scaling
// and update order are those of leaf_getf2_kernel (a NaN on the diagonal position is the
// pivot, a NaN elsewhere ranks below every number). Spins are bounded:
// on a timeout the workgroup raises *err, runs its remaining columns without waiting and
// stores NaN rows; its peers then time out in turn.
// SPREAD: the grid holds 8 workgroups per participant and only those with ids = 0 mod 8
// take part. Workgroups are dispatched to the XCDs round-robin by id, so the participants
// share one XCD (verified per device by LeafSpreadProbe when a handle is created; without
// it the leaves run unspread). The exchange words are then stored at workgroup scope --
// written through the CU's L1 into that XCD's L2 and kept there -- and read with the usual
// agent-scope loads, which return a line held dirty in the XCD's L2 and fetch clean lines
// anew: the words never make the trip to the device-coherent memory path (a round trip of
// ~0.49 us instead of ~0.74 us, measured between two workgroups). Were a participant ever
// placed on another XCD, it would not see the others' words nor they its: the bounded
// waits would run out (devInfo -1000), and no stale word passes the tag check.
// ---------------------------------------------------------------------------------------
__device__ inline unsigned long long LoadSlot(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void StoreWord(unsigned long long* p, unsigned long long v, bool local)
{
    if (local)  // uniform
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void StoreSlot(unsigned long long* p, unsigned tag, unsigned payload, bool local = false)
{
    StoreWord(p, (static_cast<unsigned long long>(tag) << 32) | payload, local);
}

// FULL: wact == W (no column of the unrolled loop is skipped at run time).
template <int R, int W, bool FULL, bool SPREAD = false>
__global__ __launch_bounds__(64) void leaf_multi_kernel(float* __restrict__ A, int lda, int m, int j0, int c0, int wact,
                                                        int panelEnd, int pivoting, int newPanel,
                                                        int* __restrict__ ipiv, int* __restrict__ info,
                                                        int* __restrict__ rowsrc, unsigned long long* slots,
                                                        unsigned tagBase, int* err, unsigned spinLimit)
{
    constexpr int T = 64;
    static_assert(W <= kRowWords && W <= T, "slot layout");
    constexpr int MOVES = 2 * W;
    __shared__ float s_row[W];
    __shared__ float s_prow[2][W];  // pivot rows of columns j (slot j % 2) and j - 1
    __shared__ int s_bpos[W];
    __shared__ int s_count;
    __shared__ int s_mdst[MOVES], s_msrc[MOVES];

    const int lane = threadIdx.x;
    if (SPREAD && (blockIdx.x & 7) != 0) return;  // not a participant
    const int P = SPREAD ? static_cast<int>(gridDim.x >> 3) : static_cast<int>(gridDim.x);
    const int g = SPREAD ? static_cast<int>(blockIdx.x >> 3) : static_cast<int>(blockIdx.x);
    const int mp = m - c0;
    float* Ap = A + static_cast<size_t>(c0) * lda + c0;
    const int rbase = g * T * R;
    // An earlier leaf of this factorization timed out: its factors are NaN already.
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    bool aborted = false;  // uniform (one wave)
    // SPREAD: the participants share one XCD (established on the host when the handle was
    // created), so the exchange words are stored at workgroup scope into its L2.
    constexpr bool local = SPREAD;

    float a[R][W];
    int pos[R], srcv[R];
#pragma unroll
    for (int i = 0; i < R; ++i)
    {
        const int lr = rbase + lane + i * T;
        const bool valid = lr < mp;
#pragma unroll
        for (int j = 0; j < W; ++j) a[i][j] = (valid && j < wact) ? Ap[static_cast<size_t>(j) * lda + lr] : 0.0f;
        pos[i] = valid ? lr : -1;
        // Panel-level sources, read before any workgroup can store (stores follow the
        // last column's exchange, which every workgroup enters after these loads).
        srcv[i] = valid ? (newPanel ? c0 + lr : rowsrc[c0 + lr]) : 0;
    }
    if (lane == 0) s_count = 0;
    int zeroCol = 0;
    float lprev[R];  // multipliers of the previous pivot (0 for rows not below it)
#pragma unroll
    for (int i = 0; i < R; ++i) lprev[i] = 0.0f;

#pragma unroll
    for (int j = 0; j < W; ++j)
    {
        if (FULL || j < wact)  // uniform
        {
            const unsigned tag = tagBase + static_cast<unsigned>(j);
            unsigned long long* keys = slots + (j & 1) * kMaxLeafGroups;
            unsigned long long* rows = slots + kKeyWords + static_cast<size_t>(j & 1) * kMaxLeafGroups * kRowWords;
            // Column j is up to date; columns beyond it still lack the previous pivot.
            // Key: magnitude class in the high word (0: not a candidate, 1: NaN off the
            // diagonal position, |a| bits + 2, 0x80000000: NaN on it), rank 0xffff - position
            // (positions are below 16384) in the low word.
            unsigned long long best = 0ull;
            int bi = -1;
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                const bool cand = pivoting ? pos[i] >= j : pos[i] == j;
                const unsigned ab = __float_as_uint(fabsf(a[i][j]));
                const unsigned mag = ab <= 0x7f800000u ? ab + 2u : (pos[i] == j ? 0x80000000u : 1u);
                const unsigned long long k =
                    cand ? (static_cast<unsigned long long>(mag) << 32) | static_cast<unsigned>(0xffff - pos[i]) : 0ull;
                bi = k > best ? i : bi;
                best = k > best ? k : best;
            }
            const unsigned long long top = WaveMaxU64Split(best);
            if (lane == 0)
                StoreWord(keys + g,
                          (static_cast<unsigned long long>(tag & 0xffffu) << 48) | ((top >> 32) << 16) |
                              (top & 0xffffull),
                          local);
            // Deferred: the previous pivot on columns j+1 .. W-1.
            if (j > 0)
#pragma unroll
                for (int k = j + 1; k < W; ++k)
                {
                    const float pk = s_prow[(j - 1) & 1][k];
#pragma unroll
                    for (int i = 0; i < R; ++i) a[i][k] = fmaf(-lprev[i], pk, a[i][k]);
                }
            // The candidate's row, now complete, through LDS to lanes 0 .. wact-1.
            if (best == top && bi >= 0)
#pragma unroll
                for (int i = 0; i < R; ++i)
                    if (i == bi)
#pragma unroll
                        for (int k = 0; k < W; ++k) s_row[k] = a[i][k];
            __syncthreads();
            if (lane < wact)
                StoreSlot(rows + g * kRowWords + lane, tag, __float_as_uint(s_row[lane < W ? lane : 0]), local);
            // The P keys: lane l reads the key words of workgroups l and l + 64.
            unsigned long long key0 = 0ull, key1 = 0ull;
            bool ok = false;
            // After a timeout the remaining columns run without waiting (no exchange, no
            // interchange): the column loop stays one straight unrolled sequence, and the
            // factors end as NaN.
            const unsigned lim = aborted ? 0u : spinLimit;
            const unsigned long long t16 = tag & 0xffffu;
            for (unsigned spins = 0; spins < lim; ++spins)
            {
                const unsigned long long w0 = lane < P ? LoadSlot(keys + lane) : t16 << 48;
                const unsigned long long w1 = lane + 64 < P ? LoadSlot(keys + lane + 64) : t16 << 48;
                // (magnitude << 32) | rank, 0 for peers without a candidate
                key0 = (((w0 >> 16) & 0xffffffffull) << 32) | (w0 & 0xffffull);
                key1 = (((w1 >> 16) & 0xffffffffull) << 32) | (w1 & 0xffffull);
                if (__all((w0 >> 48) == t16 && (w1 >> 48) == t16))
                {
                    ok = true;
                    break;
                }
            }
            const unsigned long long gtop = WaveMaxU64Split(key0 > key1 ? key0 : key1);
            const unsigned long long hit0 = __ballot(ok && lane < P && key0 == gtop);
            const unsigned long long hit1 = __ballot(ok && lane + 64 < P && key1 == gtop);
            const int winner = hit0 != 0ull   ? __ffsll(static_cast<long long>(hit0)) - 1
                               : hit1 != 0ull ? 64 + __ffsll(static_cast<long long>(hit1)) - 1
                                              : 0;
            // The winner's row: lane k holds its value in column k.
            unsigned long long rw = 0ull;
            for (unsigned spins = 0; ok; ++spins)
            {
                rw = lane < wact ? LoadSlot(rows + winner * kRowWords + lane)
                                 : static_cast<unsigned long long>(tag) << 32;
                if (__all(static_cast<unsigned>(rw >> 32) == tag)) break;
                if (spins >= spinLimit) ok = false;
            }
            if (!ok)
            {
                if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                aborted = true;
            }
            if (lane < W) s_prow[j & 1][lane] = __uint_as_float(static_cast<unsigned>(rw));
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
            // The pivot, read from lane j of the winner's row into a scalar register: the
            // reciprocal / division choice is then a scalar branch.
            const float d = __uint_as_float(__builtin_amdgcn_readlane(static_cast<unsigned>(rw), j));
            if (d == 0.0f && zeroCol == 0 && !aborted) zeroCol = c0 + j + 1;
            const bool recip = fabsf(d) >= FLT_MIN;  // uniform
            const float r = 1.0f / d;
            float q[R];
            if (recip)
#pragma unroll
                for (int i = 0; i < R; ++i) q[i] = a[i][j] * r;
            else
#pragma unroll
                for (int i = 0; i < R; ++i) q[i] = a[i][j] / d;
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                const bool below = pos[i] > j && d != 0.0f;
                lprev[i] = below ? q[i] : 0.0f;
                a[i][j] = below ? q[i] : a[i][j];
            }
            // Column j+1 now (the next candidate search needs it); the rest next column.
            if (j + 1 < W)
            {
                const float pk = s_prow[j & 1][j + 1];
#pragma unroll
                for (int i = 0; i < R; ++i) a[i][j + 1] = fmaf(-lprev[i], pk, a[i][j + 1]);
            }
        }
    }
    if (aborted)
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int j = 0; j < W; ++j) a[i][j] = __builtin_nanf("");

    // Each row to its final position (leaf columns), and its panel-level source.
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
    if (g == 0 && lane == 0 && zeroCol != 0 && *info == 0) *info = zeroCol;
    if (aborted) return;  // uniform
    __syncthreads();

    // The moved rows, from the interchange sequence (identical in every workgroup): the row
    // that ends at position p came from the position found by undoing the interchanges in
    // reverse. Candidates: positions 0..wact-1 (items 0..W-1) and the pivot positions
    // beyond them (items W..2W-1, first occurrence only).
    for (int it = lane; it < 2 * W; it += T)
    {
        int p = -1;
        if (it < W)
            p = it < wact ? it : -1;
        else if (it - W < wact)
        {
            const int jj = it - W;
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
                s_mdst[slot] = c0 + p;
                s_msrc[slot] = c0 + x;
            }
        }
    }
    __syncthreads();

    // The same interchanges on the other panel columns [j0, c0) and [c0 + wact, panelEnd),
    // columns dealt round-robin to the workgroups; each column is gathered completely
    // before it is scattered.
    const int cnt = s_count;
    const int left = c0 - j0, others = left + (panelEnd - c0 - wact);
    constexpr int SWB = 4, EPL = (MOVES + T - 1) / T;  // moved entries per lane
    for (int q0 = g; q0 < others; q0 += SWB * P)  // uniform trip count
    {
        float v[SWB][EPL];
        int cc[SWB];
#pragma unroll
        for (int bb = 0; bb < SWB; ++bb)
        {
            const int q = q0 + bb * P;
            cc[bb] = q < others ? (q < left ? j0 + q : c0 + wact + (q - left)) : -1;
#pragma unroll
            for (int u = 0; u < EPL; ++u)
            {
                const int e = lane + u * T;
                v[bb][u] = (cc[bb] >= 0 && e < cnt) ? A[static_cast<size_t>(cc[bb]) * lda + s_msrc[e]] : 0.0f;
            }
        }
        __syncthreads();
#pragma unroll
        for (int bb = 0; bb < SWB; ++bb)
#pragma unroll
            for (int u = 0; u < EPL; ++u)
            {
                const int e = lane + u * T;
                if (cc[bb] >= 0 && e < cnt) A[static_cast<size_t>(cc[bb]) * lda + s_mdst[e]] = v[bb][u];
            }
    }
}

__global__ void clear_slots_kernel(unsigned long long* slots, int words, int* err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) slots[i] = 0ull;
    if (i == 0) *err = 0;
}

// Builds the list of rows moved by the whole panel (rowsrc[r] != r) and applies it to the
// global permutation. One workgroup.
__global__ __launch_bounds__(1024) void compact_moves_kernel(int m, int j0, const int* __restrict__ rowsrc,
                                                             int* __restrict__ moved, int* __restrict__ perm)
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
    const int cnt = s_count;
    if (threadIdx.x == 0) moved[0] = cnt;
    if (perm != nullptr)
    {
        int vals[2];
        int dsts[2];
#pragma unroll
        for (int q = 0; q < 2; ++q)
        {
            const int e = threadIdx.x + q * 1024;
            dsts[q] = -1;
            if (e < cnt)
            {
                dsts[q] = moved[1 + 2 * e];
                vals[q] = perm[moved[2 + 2 * e]];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 2; ++q)
            if (dsts[q] >= 0) perm[dsts[q]] = vals[q];
    }
}

// Applies the panel's moved-row list to the columns [lo1, hi1) and [lo2, hi2): one wave64
// per column gathers every source row, then scatters.
__global__ __launch_bounds__(256) void outer_swap_kernel(float* __restrict__ A, int lda, int lo1, int hi1, int lo2,
                                                         int hi2, const int* __restrict__ moved)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int idx = blockIdx.x * 4 + wave;
    const int n1 = hi1 - lo1;
    if (idx >= n1 + (hi2 - lo2)) return;
    const int c = idx < n1 ? lo1 + idx : lo2 + (idx - n1);
    float* col = A + static_cast<size_t>(c) * lda;
    const int cnt = moved[0];
    constexpr int MAXQ = (2 * NBMAX + 63) / 64;
    float v[MAXQ];
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

// Deferred interchanges of the columns left of later panels. composed[r] <- mu(composed[r])
// for r in [lo, m), with mu the row map of one panel's moved-row list (dst -> src; every
// dst and src of that list is >= lo), built in LDS. One workgroup.
__global__ __launch_bounds__(1024) void compose_moves_kernel(int m, int lo, const int* __restrict__ moved,
                                                             int* __restrict__ composed)
{
    extern __shared__ int s_map[];  // m - lo entries
    const int len = m - lo;
    for (int r = threadIdx.x; r < len; r += blockDim.x) s_map[r] = lo + r;
    __syncthreads();
    const int cnt = moved[0];
    for (int e = threadIdx.x; e < cnt; e += blockDim.x) s_map[moved[1 + 2 * e] - lo] = moved[2 + 2 * e];
    __syncthreads();
    for (int r = threadIdx.x; r < len; r += blockDim.x) composed[lo + r] = s_map[composed[lo + r] - lo];
}

// Column c0 + blockIdx.x, rows [lo, m): row r takes the value of row composed[r], gathered
// through LDS (the whole column part first, then written back in order).
__global__ __launch_bounds__(256) void gather_rows_kernel(float* __restrict__ A, int lda, int c0, int lo, int m,
                                                          const int* __restrict__ composed)
{
    extern __shared__ float s_col[];  // m - lo values
    float* col = A + static_cast<size_t>(c0 + blockIdx.x) * lda;
    const int len = m - lo;
    for (int r = threadIdx.x; r < len; r += blockDim.x) s_col[r] = col[composed[lo + r]];
    __syncthreads();
    for (int r = threadIdx.x; r < len; r += blockDim.x) col[lo + r] = s_col[r];
}

__global__ void iota_kernel(int* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

template <int T, int R, int W>
hipError_t LaunchLeaf(float* A, int lda, int m, int j0, int c0, int wact, int panelEnd, int pivoting, int newPanel,
                      int* ipiv, int* info, const Work& w, hipStream_t stream)
{
    if (wact == W)
        hipLaunchKernelGGL((leaf_getf2_kernel<T, R, W, true>), dim3(1), dim3(T), 0, stream, A, lda, m, j0, c0, wact,
                           panelEnd, pivoting, newPanel, ipiv, info, w.rowsrc);
    else
        hipLaunchKernelGGL((leaf_getf2_kernel<T, R, W, false>), dim3(1), dim3(T), 0, stream, A, lda, m, j0, c0, wact,
                           panelEnd, pivoting, newPanel, ipiv, info, w.rowsrc);
    return hipGetLastError();
}

// Panels with more active rows than this run their leaves over several workgroups
// (leaf_multi_kernel); OPS_SOLVER_MULTI_LEAF_ROWS overrides it (0: never).
int MultiLeafRows()
{
    static const int rows = []
    {
        const char* e = std::getenv("OPS_SOLVER_MULTI_LEAF_ROWS");
        return e != nullptr ? std::atoi(e) : 2048;
    }();
    return rows;
}
inline bool UseMultiLeaf(int mp) { return MultiLeafRows() > 0 && mp > MultiLeafRows(); }

// Multi-workgroup leaves by cooperative launch (OPS_SOLVER_LEAF_COOP, default 1).
bool LeafCooperative()
{
    static const bool on = []
    {
        const char* e = std::getenv("OPS_SOLVER_LEAF_COOP");
        return e == nullptr || std::atoi(e) != 0;
    }();
    return on;
}

constexpr int kMultiT = 64, kMultiR = 2, kMultiW = 64;  // one wave of 128 rows per workgroup, 64 columns

// Per device: 1 when LeafSpreadProbe found the workgroups with ids = 0 mod 8 of a
// 1024-workgroup cooperative launch on one XCD, -1 when not, 0 before the probe.
constexpr int kMaxDevices = 64;
std::atomic<int> g_spreadOk[kMaxDevices];
__device__ int g_xccOfGroup[8 * kMaxLeafGroups];

__global__ void xcc_probe_kernel()
{
    if (threadIdx.x == 0) g_xccOfGroup[blockIdx.x] = static_cast<int>(__builtin_amdgcn_s_getreg((3 << 11) | 20) & 0xfu);
}

}  // namespace

// Full-width tall leaves run spread over 8x the grid and exchange through one XCD's L2 when
// the probe allowed it; OPS_SOLVER_LEAF_XCD=0 turns that off (read per launch).
bool LeafSpread()
{
    const char* e = std::getenv("OPS_SOLVER_LEAF_XCD");
    if (e != nullptr && std::atoi(e) == 0) return false;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) return false;
    return g_spreadOk[device].load(std::memory_order_relaxed) == 1;
}

namespace
{

hipError_t MultiLeaf(float* A, int lda, int m, int j0, int c0, int wact, int panelEnd, int pivoting, int newPanel,
                     int* ipiv, int* info, const Work& w, hipStream_t stream)
{
    const int mp = m - c0;
    const int groups = (mp + kMultiT * kMultiR - 1) / (kMultiT * kMultiR);
    if (groups > kMaxLeafGroups || wact > kMultiW || w.epoch == nullptr) return hipErrorInvalidValue;
    unsigned tagBase = (*w.epoch)++ * 64u;
    // Key words carry 16-bit tags: within one factorization (slots cleared at its start)
    // the tags must not wrap. At most n / 64 <= 256 leaves (plus tournament epochs) use them.
    if (tagBase + 64u > 0xffffu) return hipErrorInvalidValue;
    unsigned spinLimit = SpinLimit();
    // Cooperative launch: the workgroups wait for each other every column, so the runtime
    // must make all of them resident at once (it refuses a grid that cannot be).
    int* rowsrc = w.rowsrc;
    unsigned long long* slots = w.slots;
    int* err = w.err;
    void* args[] = {&A, &lda, &m, &j0, &c0, &wact, &panelEnd, &pivoting, &newPanel, &ipiv, &info, &rowsrc, &slots,
                    &tagBase, &err, &spinLimit};
    const bool full = wact == kMultiW;
    if (!LeafCooperative())
    {
        if (full)
            hipLaunchKernelGGL((leaf_multi_kernel<kMultiR, kMultiW, true>), dim3(groups), dim3(kMultiT), 0, stream, A,
                               lda, m, j0, c0, wact, panelEnd, pivoting, newPanel, ipiv, info, rowsrc, slots, tagBase,
                               err, spinLimit);
        else
            hipLaunchKernelGGL((leaf_multi_kernel<kMultiR, kMultiW, false>), dim3(groups), dim3(kMultiT), 0, stream, A,
                               lda, m, j0, c0, wact, panelEnd, pivoting, newPanel, ipiv, info, rowsrc, slots, tagBase,
                               err, spinLimit);
        return hipGetLastError();
    }
    if (full && LeafSpread())
        return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&leaf_multi_kernel<kMultiR, kMultiW, true, true>),
                                          dim3(8 * groups), dim3(kMultiT), args, 0, stream);
    const void* fn = full ? reinterpret_cast<const void*>(&leaf_multi_kernel<kMultiR, kMultiW, true>)
                          : reinterpret_cast<const void*>(&leaf_multi_kernel<kMultiR, kMultiW, false>);
    return hipLaunchCooperativeKernel(fn, dim3(groups), dim3(kMultiT), args, 0, stream);
}

// Panels with more active rows than this take tournament leaves when the caller allows
// them; OPS_SOLVER_TSLU_ROWS overrides it.
int TsluRows()
{
    static const int rows = []
    {
        const char* e = std::getenv("OPS_SOLVER_TSLU_ROWS");
        return e != nullptr ? std::atoi(e) : 1024;
    }();
    return rows;
}
inline bool UseTslu(const Work& w, int mp) { return w.tournament && mp > TsluRows() && mp <= kTsluMaxRows; }

// Leaf width for a panel of mp active rows: every shape keeps R*W <= 128 values per lane
// (512-thread workgroups leave up to 256 VGPRs per lane).
int LeafWidth(int mp, const Work& w)
{
    if (UseTslu(w, mp)) return kTsluWidth;
    if (UseMultiLeaf(mp)) return kMultiW;
    if (mp <= 512) return 64;
    if (mp <= 2048) return 32;
    if (mp <= 8192) return 8;
    return 4;
}

hipError_t Leaf(float* A, int lda, int m, int j0, int c0, int wact, int panelEnd, int pivoting, int newPanel,
                int* ipiv, int* info, const Work& w, hipStream_t stream)
{
    const int mp = m - c0;
    // Algorithmic traffic of a leaf: its panel columns read and written once.
    ProfileScope prof(ProfileKernel::PanelLeaf, 8.0 * mp * wact, stream);
    if (pivoting && UseTslu(w, mp))
        return TsluLeaf(A, lda, m, j0, c0, wact, panelEnd, newPanel, ipiv, info, w.rowsrc, w.tslu, (*w.epoch)++,
                        w.err, stream);
    if (UseMultiLeaf(mp)) return MultiLeaf(A, lda, m, j0, c0, wact, panelEnd, pivoting, newPanel, ipiv, info, w, stream);
#define LEAF(T, R, W) \
    LaunchLeaf<T, R, W>(A, lda, m, j0, c0, wact, panelEnd, pivoting, newPanel, ipiv, info, w, stream)
    if (mp <= 256) return LEAF(256, 1, 64);
    if (mp <= 512) return LEAF(512, 1, 64);
    if (mp <= 1024) return LEAF(512, 2, 32);
    if (mp <= 2048) return LEAF(512, 4, 32);
    if (mp <= 4096) return LEAF(512, 8, 8);
    if (mp <= 8192) return LEAF(512, 16, 8);
    return LEAF(1024, 16, 4);
#undef LEAF
}

#define RET_IF(e)                          \
    do                                     \
    {                                      \
        hipError_t _e = (e);               \
        if (_e != hipSuccess) return _e;   \
    } while (0)

// Factors columns [c0, c0 + w) of the outer panel [j0, panelEnd) over rows [c0, m); every
// panel column already holds the updates of the columns left of c0. The left half is
// factored first (recursively), its L11 solves the right half's top rows and an MFMA
// product updates the rest of the right half, then the right half is factored. Each leaf
// applies its row interchanges to all other panel columns itself.
hipError_t PanelRec(float* a, int lda, int m, int j0, int panelEnd, int c0, int w, int pivoting, int* ipiv, int* info,
                    const Work& wk, hipStream_t stream)
{
    const int W = LeafWidth(m - c0, wk);
    if (w <= W) return Leaf(a, lda, m, j0, c0, w, panelEnd, pivoting, c0 == j0, ipiv, info, wk, stream);
    const int h = (w / 2 + W - 1) / W * W;  // whole leaves on the left
    RET_IF(PanelRec(a, lda, m, j0, panelEnd, c0, h, pivoting, ipiv, info, wk, stream));
    const int r = w - h;
    const float* L11 = a + static_cast<size_t>(c0) * lda + c0;
    float* B12 = a + static_cast<size_t>(c0 + h) * lda + c0;
    if (h <= 64)  // one diagonal block: a single substitution launch
        RET_IF(TrsmBlock(h, r, L11, lda, true, true, false, B12, lda, stream));
    else
    {
        RET_IF(TrsmLowerUnitInverse(h, L11, lda, wk.linvNode, stream));
        RET_IF(TrsmLowerUnitApply(h, r, L11, lda, wk.linvNode, B12, lda, stream));
    }
    if (m - c0 - h > 0)
        RET_IF(GemmSub(m - c0 - h, r, h, a + static_cast<size_t>(c0) * lda + c0 + h, lda, 0,
                       a + static_cast<size_t>(c0 + h) * lda + c0, lda, a + static_cast<size_t>(c0 + h) * lda + c0 + h,
                       lda, Precision::FP32, stream));
    return PanelRec(a, lda, m, j0, panelEnd, c0 + h, r, pivoting, ipiv, info, wk, stream);
}

int OuterBlock(int m, int n, Precision prec)
{
    const int k = std::min(m, n);
    if (prec != Precision::FP32) return k >= 4096 ? 512 : (k >= 1024 ? 256 : 128);
    return k >= 2048 ? 256 : 128;
}

}  // namespace

size_t RealGetrfWorkspaceBytes(int m, int n)
{
    (void)n;
    // The tournament leaves' scratch only where one can run (a panel taller than TsluRows).
    return WorkInts(std::max(m, 1)) * sizeof(int) + 8 + kSlotBytes + 16 + 256 + 3 * kLinvPanelBytes +
           (m > TsluRows() ? TsluScratchBytes(m) : 0);
}

int* RealGetrfErrorWord(void* work, int m) { return Layout(work, std::max(m, 1)).err; }

// ipiv == nullptr selects no pivoting. perm (optional, m ints) receives the row
// permutation P with (P A)[i] = A[perm[i]]. pack (optional, BF16 / BF16X3 trailing
// precisions): GemmPackedWorkspaceBytes(m, n, 512) bytes for the packed trailing update.
hipError_t RealGetrfPerm(int m, int n, float* a, int lda, int* ipiv, int* info, void* work, Precision trailing,
                         int* permOut, hipStream_t stream, void* pack, const LookAhead* la, bool tournament)
{
    if (m > kMaxRows) return hipErrorInvalidValue;
    Work w = Layout(work, m);
    if (permOut) w.perm = permOut;
    w.tournament = tournament && ipiv != nullptr && m > TsluRows();
    if (w.tournament) RET_IF(TsluClear(w.tslu, m, stream));
    unsigned epoch = 1;
    w.epoch = &epoch;
    FillInfo(info, 1, 0, stream);
    {
        // The multi-workgroup leaves' exchange slots (when they run) and the timeout flag
        // shared by every bounded wait of the factorization.
        const int words = UseMultiLeaf(m) ? static_cast<int>(kSlotBytes / sizeof(unsigned long long)) : 0;
        hipLaunchKernelGGL(clear_slots_kernel, dim3(std::max(1, (words + 255) / 256)), dim3(256), 0, stream, w.slots,
                           words, w.err);
    }
    const int kmax = std::min(m, n);
    if (kmax == 0) return hipGetLastError();
    const int pivoting = ipiv != nullptr;
    if (permOut) hipLaunchKernelGGL(iota_kernel, dim3((m + 255) / 256), dim3(256), 0, stream, permOut, m);
    const int NB = OuterBlock(m, n, trailing);
    const int npanel = (kmax + NB - 1) / NB;
    // Deferred interchanges (default; OPS_SOLVER_DEFER_SWAPS=0 turns them off): the columns
    // left of a panel receive its interchanges not panel by panel but once at the end,
    // through the composition of all later panels' row maps (every panel keeps its moved-row
    // list for that): about n^2/2 gathers instead of ~2 n^2 scattered reads and writes,
    // which in column-major storage cost a cache line each and compete with the panel.
    const char* dz = std::getenv("OPS_SOLVER_DEFER_SWAPS");
    const bool defer = pivoting && npanel > 1 && (dz == nullptr || std::atoi(dz) != 0) &&
                       static_cast<size_t>(npanel) <= MaxPanels(m);
    auto moved = [&](int k) { return defer ? w.movedAll + static_cast<size_t>(k) * kMovedInts : w.moved[k & 1]; };

    // Look-ahead: the caller's stream s0 carries the critical path (the next block column's
    // update and its panel); the side stream s1 carries the rest of each trailing update.
    const bool two = la != nullptr && la->side != nullptr && npanel > 1;
    const bool split = two && la->panelStream != nullptr;  // panel path on its own CUs
    hipStream_t s0 = split ? la->panelStream : stream, s1 = two ? la->side : stream;
    if (two)
    {
        RET_IF(hipEventRecord(la->fork, stream));
        RET_IF(hipStreamWaitEvent(s1, la->fork, 0));
        if (split) RET_IF(hipStreamWaitEvent(s0, la->fork, 0));
    }

    // Panel k: factor, then list its interchanges and fold them into perm.
    auto panel = [&](int k) -> hipError_t
    {
        const int j0 = k * NB, nb = std::min(NB, kmax - j0);
        RET_IF(PanelRec(a, lda, m, j0, j0 + nb, j0, nb, pivoting, ipiv, info, w, s0));
        if (pivoting)
            hipLaunchKernelGGL(compact_moves_kernel, dim3(1), dim3(1024), 0, s0, m, j0, w.rowsrc, moved(k),
                               permOut);
        RET_IF(hipGetLastError());
        // The U12 solves of both streams use the block inverses of this panel's L11.
        return TrsmLowerUnitInverse(nb, a + static_cast<size_t>(j0) * lda + j0, lda, w.linvPanel[k & 1], s0);
    };
    // Applies panel k to the columns [lo, hi) (interchanges, U12 solve, trailing product)
    // and, when leftSwaps, its interchanges to the columns left of the panel.
    auto update = [&](int k, int lo, int hi, bool leftSwaps, bool packed, hipStream_t st) -> hipError_t
    {
        const int j0 = k * NB, nb = std::min(NB, kmax - j0), pe = j0 + nb;
        const int cols = hi - lo, left = leftSwaps ? j0 : 0;
        if (pivoting && cols + left > 0)
            hipLaunchKernelGGL(outer_swap_kernel, dim3((cols + left + 3) / 4), dim3(256), 0, st, a, lda, 0, left, lo,
                               hi, moved(k));
        RET_IF(hipGetLastError());
        if (cols <= 0) return hipSuccess;
        RET_IF(TrsmLowerUnitApply(nb, cols, a + static_cast<size_t>(j0) * lda + j0, lda, w.linvPanel[k & 1],
                                  a + static_cast<size_t>(lo) * lda + j0, lda, st));
        if (m - pe <= 0) return hipSuccess;
        const float* L21 = a + static_cast<size_t>(j0) * lda + pe;
        const float* U12 = a + static_cast<size_t>(lo) * lda + j0;
        float* A22 = a + static_cast<size_t>(lo) * lda + pe;
        if (packed && pack != nullptr)
            return GemmSubPacked(m - pe, cols, nb, L21, lda, U12, lda, A22, lda, trailing, pack, st);
        return GemmSub(m - pe, cols, nb, L21, lda, 0, U12, lda, A22, lda, trailing, st);
    };

    RET_IF(panel(0));
    if (two) RET_IF(hipEventRecord(la->panel[0], s0));
    for (int k = 0; k < npanel; ++k)
    {
        const int pe = std::min(kmax, (k + 1) * NB);
        const int nbNext = k + 1 < npanel ? std::min(NB, kmax - pe) : 0;
        // Bulk of the trailing update (s1): everything right of the next block column.
        if (two) RET_IF(hipStreamWaitEvent(s1, la->panel[k & 1], 0));
        RET_IF(update(k, pe + nbNext, n, !defer, true, s1));
        if (two) RET_IF(hipEventRecord(la->bulk[k & 1], s1));
        // Critical path (s0): the next block column, then its panel.
        if (nbNext > 0)
        {
            if (two && k > 0) RET_IF(hipStreamWaitEvent(s0, la->bulk[(k - 1) & 1], 0));
            RET_IF(update(k, pe, pe + nbNext, false, false, s0));
            RET_IF(panel(k + 1));
            if (two) RET_IF(hipEventRecord(la->panel[(k + 1) & 1], s0));
        }
    }
    if (two) RET_IF(hipStreamWaitEvent(s0, la->bulk[(npanel - 1) & 1], 0));
    if (defer)
    {
        // Backward over the panels: composed = mu_{p+1} o ... o mu_last, then panel p's
        // columns gather their rows [(p+1) NB, m) through it.
        hipLaunchKernelGGL(iota_kernel, dim3((m + 255) / 256), dim3(256), 0, s0, w.composed, m);
        for (int p = npanel - 2; p >= 0; --p)
        {
            const int lo = (p + 1) * NB, nb = std::min(NB, kmax - p * NB);
            if (lo >= m) continue;
            const size_t lds = static_cast<size_t>(m - lo) * sizeof(int);
            hipLaunchKernelGGL(compose_moves_kernel, dim3(1), dim3(1024), lds, s0, m, lo, moved(p + 1), w.composed);
            hipLaunchKernelGGL(gather_rows_kernel, dim3(nb), dim3(256), lds, s0, a, lda, p * NB, lo, m, w.composed);
        }
        RET_IF(hipGetLastError());
    }
    FoldWaitError(w.err, info, s0);
    if (split)
    {
        RET_IF(hipEventRecord(la->join, s0));
        RET_IF(hipStreamWaitEvent(stream, la->join, 0));
    }
    return hipGetLastError();
}

hipError_t RealGetrf(int m, int n, float* a, int lda, int* ipiv, int* info, void* work, Precision trailing,
                     hipStream_t stream)
{
    return RealGetrfPerm(m, n, a, lda, ipiv, info, work, trailing, nullptr, stream, nullptr, nullptr);
}

// ---------------------------------------------------------------------------------------
// getrs on Device column-major data.
// ---------------------------------------------------------------------------------------
namespace
{
// ipiv (LAPACK interchange sequence) -> permutation, perm[i] = source row of row i of P*B.
// With s_k the interchange of rows k and ipiv[k] - 1 (applied for k = 0, 1, ...), row i of
// P*B comes from row s_0(s_1(... s_{n-1}(i))): one lane per position traces it backward
// through the interchanges, ipiv staged in LDS (padded to a multiple of 4 with identities,
// read four at a time as broadcasts). An out-of-range entry (and, with CLAMP_BELOW, one
// pointing above its own row) is taken as no interchange.
template <bool CLAMP_BELOW>
__global__ __launch_bounds__(64) void ipiv_to_perm_kernel(int n, const int* __restrict__ ipiv, int* __restrict__ perm)
{
    extern __shared__ int s_piv[];
    const int n4 = (n + 3) & ~3;
#pragma unroll 16  // 16 loads per lane in flight
    for (int k = threadIdx.x; k < n4; k += 64)
    {
        int p = k < n ? ipiv[k] - 1 : k;
        if (p < 0 || p >= n || (CLAMP_BELOW && p < k)) p = k;
        s_piv[k] = p;
    }
    __syncthreads();
    const int i = blockIdx.x * 64 + static_cast<int>(threadIdx.x);
    int x = i;
    const int4* q = reinterpret_cast<const int4*>(s_piv);
#pragma unroll 4
    for (int k = n4 - 4; k >= 0; k -= 4)
    {
        const int4 p = q[k >> 2];
        x = x == k + 3 ? p.w : (x == p.w ? k + 3 : x);
        x = x == k + 2 ? p.z : (x == p.z ? k + 2 : x);
        x = x == k + 1 ? p.y : (x == p.y ? k + 1 : x);
        x = x == k ? p.x : (x == p.x ? k : x);
    }
    if (i < n) perm[i] = x;
}

// dst(:, c) = src(perm, c) (gather) or dst(perm, c) = src(:, c) (scatter).
template <bool GATHER>
__global__ void permute_rows_kernel(int n, int nrhs, const int* __restrict__ perm, const float* __restrict__ src,
                                    int lds, float* __restrict__ dst, int ldd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= nrhs) return;
    const int p = perm ? perm[i] : i;
    if (GATHER)
        dst[static_cast<size_t>(c) * ldd + i] = src[static_cast<size_t>(c) * lds + p];
    else
        dst[static_cast<size_t>(c) * ldd + p] = src[static_cast<size_t>(c) * lds + i];
}

int SolveBlock(int n) { return n >= 4096 ? 256 : 128; }
}  // namespace

namespace
{
size_t GetrsVectorBytes(int n, int nrhs)
{
    return ((static_cast<size_t>(n) * std::max(nrhs, 1) + static_cast<size_t>(n) + 64) * sizeof(float) + 255) / 256 *
           256;
}
// One right-hand side of at least this order uses the single-launch chained solve.
constexpr int kChainMinN = 512;
}  // namespace

size_t RealGetrsWorkspaceBytes(int n, int nrhs)
{
    return GetrsVectorBytes(n, nrhs) + TrsvChainStateBytes(std::max(n, 1));
}

// Triangular solves on a column-major n x nrhs block X (ldx) with packed LU factors.
hipError_t RealLuSolveInPlace(int trans, int n, int nrhs, const float* lu, int lda, float* x, int ldx,
                              hipStream_t stream)
{
    const int NBS = SolveBlock(n);
    if (trans == 0)
    {
        for (int k = 0; k < n; k += NBS)  // L y = b (unit lower)
        {
            const int kb = std::min(NBS, n - k);
            RET_IF(TrsmBlock(kb, nrhs, lu + static_cast<size_t>(k) * lda + k, lda, true, true, false, x + k, ldx,
                             stream));
            if (n - k - kb > 0)
                RET_IF(GemmSub(n - k - kb, nrhs, kb, lu + static_cast<size_t>(k) * lda + k + kb, lda, 0, x + k, ldx,
                               x + k + kb, ldx, Precision::FP32, stream));
        }
        for (int k = ((n - 1) / NBS) * NBS; k >= 0; k -= NBS)  // U x = y
        {
            const int kb = std::min(NBS, n - k);
            RET_IF(TrsmBlock(kb, nrhs, lu + static_cast<size_t>(k) * lda + k, lda, false, false, false, x + k, ldx,
                             stream));
            if (k > 0)
                RET_IF(GemmSub(k, nrhs, kb, lu + static_cast<size_t>(k) * lda, lda, 0, x + k, ldx, x, ldx,
                               Precision::FP32, stream));
        }
    }
    else
    {
        for (int k = 0; k < n; k += NBS)  // U^T w = b (lower, non-unit)
        {
            const int kb = std::min(NBS, n - k);
            RET_IF(TrsmBlock(kb, nrhs, lu + static_cast<size_t>(k) * lda + k, lda, false, false, true, x + k, ldx,
                             stream));
            if (n - k - kb > 0)  // x[k+kb:] -= U[k:k+kb, k+kb:]^T w[k:k+kb]
                RET_IF(GemmSub(n - k - kb, nrhs, kb, lu + static_cast<size_t>(k + kb) * lda + k, lda, 1, x + k, ldx,
                               x + k + kb, ldx, Precision::FP32, stream));
        }
        for (int k = ((n - 1) / NBS) * NBS; k >= 0; k -= NBS)  // L^T z = w (upper, unit)
        {
            const int kb = std::min(NBS, n - k);
            RET_IF(TrsmBlock(kb, nrhs, lu + static_cast<size_t>(k) * lda + k, lda, true, true, true, x + k, ldx,
                             stream));
            if (k > 0)  // x[0:k] -= L[k:k+kb, 0:k]^T z[k:k+kb]
                RET_IF(GemmSub(k, nrhs, kb, lu + k, lda, 1, x + k, ldx, x, ldx, Precision::FP32, stream));
        }
    }
    return hipGetLastError();
}

hipError_t IpivToPerm(int n, const int* ipiv, int* perm, hipStream_t stream, bool clampBelow)
{
    if (n > kMaxRows) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    const dim3 grid((n + 63) / 64);
    const size_t lds = static_cast<size_t>((n + 3) & ~3) * sizeof(int);
    if (clampBelow)
        hipLaunchKernelGGL(ipiv_to_perm_kernel<true>, grid, dim3(64), lds, stream, n, ipiv, perm);
    else
        hipLaunchKernelGGL(ipiv_to_perm_kernel<false>, grid, dim3(64), lds, stream, n, ipiv, perm);
    return hipGetLastError();
}

hipError_t PermuteRows(bool gather, int n, int nrhs, const int* perm, const float* src, int lds, float* dst, int ldd,
                       hipStream_t stream)
{
    dim3 grid((n + 255) / 256, nrhs);
    if (gather)
        hipLaunchKernelGGL(permute_rows_kernel<true>, grid, dim3(256), 0, stream, n, nrhs, perm, src, lds, dst, ldd);
    else
        hipLaunchKernelGGL(permute_rows_kernel<false>, grid, dim3(256), 0, stream, n, nrhs, perm, src, lds, dst, ldd);
    return hipGetLastError();
}

namespace
{
// dst (column-major, ldd) <- src (row-major rows x cols, lds); 32x32 tiles through LDS so
// both the read and the write are coalesced.
__global__ __launch_bounds__(256) void transpose_rm_kernel(int rows, int cols, const float* __restrict__ src,
                                                           int64_t lds, float* __restrict__ dst, int64_t ldd)
{
    __shared__ float tile[32][33];
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
}  // namespace

hipError_t TransposeRM(int rows, int cols, const float* src, int64_t lds, float* dst, int64_t ldd, hipStream_t stream)
{
    dim3 grid((cols + 31) / 32, (rows + 31) / 32);
    hipLaunchKernelGGL(transpose_rm_kernel, grid, dim3(256), 0, stream, rows, cols, src, lds, dst, ldd);
    return hipGetLastError();
}

hipError_t RealGetrs(int trans, int n, int nrhs, const float* lu, int lda, const int* ipiv, float* b, int ldb,
                     void* work, hipStream_t stream, int* info)
{
    if (n <= 0 || nrhs <= 0) return hipSuccess;
    float* x = static_cast<float*>(work);
    int* perm = reinterpret_cast<int*>(x + static_cast<size_t>(n) * nrhs);
    if (ipiv) RET_IF(IpivToPerm(n, ipiv, perm, stream));
    const int* p = ipiv ? perm : nullptr;
    if (trans == 0)
    {
        RET_IF(PermuteRows(true, n, nrhs, p, b, ldb, x, n, stream));
        if (nrhs == 1 && n >= kChainMinN)
        {
            TrsvChainState ts = TrsvChainLayout(static_cast<char*>(work) + GetrsVectorBytes(n, nrhs), n);
            RET_IF(TrsvChainReset(ts, stream));
            RET_IF(TrsvChain(ts, false, lu, lda, x, stream));
            RET_IF(TrsvChain(ts, true, lu, lda, x, stream));
            FoldWaitError(ts.err, info, stream);
        }
        else
            RET_IF(RealLuSolveInPlace(0, n, nrhs, lu, lda, x, n, stream));
        // copy back
        RET_IF(PermuteRows(true, n, nrhs, nullptr, x, n, b, ldb, stream));
    }
    else
    {
        RET_IF(PermuteRows(true, n, nrhs, nullptr, b, ldb, x, n, stream));
        RET_IF(RealLuSolveInPlace(1, n, nrhs, lu, lda, x, n, stream));
        RET_IF(PermuteRows(false, n, nrhs, p, x, n, b, ldb, stream));
    }
    return hipGetLastError();
}

// Records the XCD of every workgroup of a 1024-workgroup cooperative launch and allows the
// spread tall leaves on this device when those with ids = 0 mod 8 share one XCD. Runs once
// per device (the device current), synchronously on the null stream; two threads probing
// one device at once both store the same result.
void LeafSpreadProbe(int device)
{
    if (device < 0 || device >= kMaxDevices || g_spreadOk[device].load() != 0) return;
    constexpr int G = 8 * kMaxLeafGroups;
    int result = -1;
    if (hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&xcc_probe_kernel), dim3(G), dim3(64), nullptr, 0,
                                   nullptr) == hipSuccess &&
        hipStreamSynchronize(nullptr) == hipSuccess)
    {
        int h[G];
        if (hipMemcpyFromSymbol(h, HIP_SYMBOL(g_xccOfGroup), sizeof(h)) == hipSuccess)
        {
            result = 1;
            for (int k = 1; k < kMaxLeafGroups; ++k)
                if (h[8 * k] != h[0]) result = -1;
        }
    }
    (void)hipGetLastError();
    g_spreadOk[device].store(result);
}

}  // namespace ops
