// Uneven batched LU solve (getrf + getrs per matrix), EXTENSION: the reference has no
// batched getrf/getrs.
//
// One persistent launch: a grid of one 1024-thread workgroup per CU pulls items from a
// device queue (largest orders first: longest-processing-time order). An item is a group of
// matrices of similar order and one configuration, which splits the workgroup's 16 waves:
//   W64   n <= 64:  one wave per matrix, 16 matrices per item;
//   C2x4  n <= 128: column-cyclic, 2 rows per lane, 4 waves per matrix, 4 per item;
//   C2x8  n <= 128: column-cyclic, 2 rows per lane, 8 waves per matrix, 2 per item;
//   C3x8  n <= 192: column-cyclic, 3 rows per lane, 8 waves per matrix, 2 per item;
//   C3x16 n <= 192: column-cyclic, 3 rows per lane, 16 waves, 1 per item;
//   C4x16 n <= 256: column-cyclic, 4 rows per lane, 16 waves, 1 per item.
// The configuration of each 16-wide size bin above 64 is chosen per device by an autotuner
// that times every configuration that fits the bin on one item per CU of synthetic systems
// of the bin's largest order (at most ~80 MB of samples, generated on the device) and keeps the
// one with the least time per matrix; a plan may also be given the schedule explicitly.
// The last workgroup to leave re-arms the queue, so executing a plan is one kernel launch.
//
// Column-cyclic kernel (one matrix per workgroup): lane l of every wave holds rows
// l, l + 64, ... of the matrix; wave w holds columns w, w + NW, w + 2 NW, ... of the augmented
// matrix [A | B]. The pivot search of column k runs entirely inside its owner wave (a 64-bit
// key argmax: largest |a|, lowest position on ties, a NaN wins only on the diagonal); the owner publishes the
// pivot position and the multipliers of its rows through a ring of LDS entries guarded by
// per-entry flags, and every wave updates its own columns with the pivot row read by
// v_readlane from its own registers. No workgroup barrier runs during the elimination: a
// wave waits only for the column it needs, and the owner of column k + 1 updates that
// column first and publishes its pivot before finishing the update of column k
// (look-ahead). Rows keep their registers; the position each occupies is tracked
// identically in every wave (implicit pivoting, as in the register leaves of the dense
// factorization). The right-hand sides are extra columns, so the elimination also forward
// substitutes them; the back substitution runs from the stored U, one wave per right-hand
// side.
// Outputs: packed LU (rows at their final positions), 1-based ipiv, X, info (first zero
// pivot).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <climits>
#include <map>
#include <mutex>
#include <new>
#include <vector>

#include "internal.h"

namespace ops
{
namespace
{
constexpr int kNrhsMax = 4;  // right-hand sides per matrix carried in registers
constexpr int kBS = 32;      // back-substitution block

// Configurations (see above): rows per lane, waves per matrix, matrices per item, largest n.
enum Cfg : int
{
    kW64 = 0,
    kC2x4 = 1,
    kC2x8 = 2,
    kC3x8 = 3,
    kC3x16 = 4,
    kC4x16 = 5,
    kNumCfg = 6
};
constexpr int kCfgCap[kNumCfg] = {16, 4, 2, 2, 1, 1};
constexpr int kCfgMaxN[kNumCfg] = {64, 128, 128, 192, 192, 256};
constexpr int kBinW = 16;                          // size-bin width above 64
constexpr int kNumBins = (256 - 64) / kBinW;       // bins 65..80, ..., 241..256
inline int BinOf(int n) { return (n - 65) / kBinW; }  // n > 64

struct Item
{
    int first;  // index into the plan's order
    int count;  // matrices in this item (<= kCfgCap[cfg])
    int cfg;
    int pad;
};

// NaN on the diagonal position (pos == k) wins, NaN elsewhere never does: the reference's
// running best starts at the diagonal and only a strictly larger |a| replaces it
// (src/utils/kernel/real_lu.hpp:59-77, :100-109).
__device__ inline unsigned long long KeyOf(float v, int pos, int k)
{
    const float a = fabsf(v);
    const unsigned hi = (a == a) ? __float_as_uint(a) + 1u : (pos == k ? 0x7fffffffu : 0u);
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
            (active && ownK && pos >= k && pos != INT_MAX) ? KeyOf(cur, pos, k) : 0ull;
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

// ---------------------------------------------------------------------------------------
// Column-cyclic kernel for 65 <= n <= 256.

// LDS ring entries of published pivot columns (a wave may run that many columns ahead of
// the slowest one before it waits for an entry).
template <int NW>
constexpr int RingSize() { return NW >= 8 ? 32 : 16; }

template <int RPL, int NW, int CPW>
struct CycLds
{
    float4 l[RingSize<NW>()][64];            // multipliers of the published column, by lane (row slots .x..)
    unsigned long long tag[RingSize<NW>()];  // (column index + 1) | pivot position, lane, slot << 32
    int done[NW];                   // columns whose pivot information each wave has taken
    int zmin;                       // first zero pivot (1-based) over the owners
    int abort;                      // set when a wait exceeded its bound: every wave then runs out
    float4 prow[NW][2][CPW / 4];    // per wave: pivot row of the current / the deferred step
    float y[kNrhsMax][RPL * 64];
};

constexpr int kCycSpin = 1 << 20;  // bound on the sleeps of one wait (~30 ms)

// True when the workgroup has given up (another wave's wait ran out) or this wait has run
// out, in which case it raises the flag; info then reports -1 for the matrix.
__device__ inline bool Aborted(int& flag, int it)
{
    if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(&flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)))
        return true;
    if (it < kCycSpin) return false;
    __hip_atomic_store(&flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return true;
}

__device__ inline float Rl(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// The pivot row's values in this wave's column slots, written to the wave's LDS row buffer
// by the lane that holds the row (lane lp, slot sp; call under lane == lp) for the chunks of
// four slots below nv; every lane then reads them back as a broadcast. This replaces one
// v_readlane per slot and keeps the branch over sp away from the update of the slot array.
template <int S, int RPL, int CPW>
__device__ __attribute__((always_inline)) inline void PutPivotRow(const float (&a)[RPL][CPW], int sp, int nv,
                                                                  float4* __restrict__ row)
{
    if constexpr (S < RPL)
    {
        if (S == RPL - 1 || sp == S)
        {
#pragma unroll
            for (int K = 0; K < CPW; K += 4)
                if (K < nv) row[K / 4] = make_float4(a[S][K], a[S][K + 1], a[S][K + 2], a[S][K + 3]);
        }
        else
            PutPivotRow<S + 1>(a, sp, nv, row);
    }
}

// Rank-1 update of the unfinished column slots [0, nv) of a wave, a[i][kk] -= l[i] * u[kk],
// in chunks of four slots (chunks at or beyond nv are skipped). Slots >= nv hold no live
// column (finished columns are already stored), so the chunk that straddles nv needs no
// mask.
template <int RPL, int CPW>
__device__ __attribute__((always_inline)) inline void UpdateSlots(float (&a)[RPL][CPW], const float (&l)[RPL], int nv,
                                                                  const float4* __restrict__ row)
{
#pragma unroll
    for (int K = 0; K < CPW; K += 4)
    {
        if (K < nv)
        {
            const float4 u4 = row[K / 4];
            const float u[4] = {u4.x, u4.y, u4.z, u4.w};
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int i = 0; i < RPL; ++i) a[i][K + t] = fmaf(-l[i], u[t], a[i][K + t]);
        }
    }
}

// The look-ahead step's update of slots 1.. with two rank-1 updates, in order: the deferred
// one of the previous column (ld, pivot row in rowD; zero when nothing was deferred) and the
// current one (lc, pivot row in rowC), while the slots rotate down by one (slot 0, the
// column just pivoted, leaves the array), so the wave's next pivot column is always slot 0
// and no slot is addressed by a run-time index. The current pivot row has not received the
// deferred update yet, so its values are corrected first with the same fused multiply-add
// the row itself would have seen (ldp = its deferred multiplier): the result is
// bit-identical to applying the two updates one after the other.
template <int RPL, int CPW>
__device__ __attribute__((always_inline)) inline void RotSlots(float (&a)[RPL][CPW], const float (&ld)[RPL], float ldp,
                                                               const float (&lc)[RPL], int nv,
                                                               const float4* __restrict__ rowD,
                                                               const float4* __restrict__ rowC)
{
#pragma unroll
    for (int K = 0; K < CPW; K += 4)
    {
        if (K < nv)
        {
            const float4 d4 = rowD[K / 4], c4 = rowC[K / 4];
            const float ud[4] = {d4.x, d4.y, d4.z, d4.w};
            const float uc[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                const int kk = K + t;
                if (kk == 0) continue;
                const float ck = fmaf(-ldp, ud[t], uc[t]);
#pragma unroll
                for (int i = 0; i < RPL; ++i) a[i][kk - 1] = fmaf(-lc[i], ck, fmaf(-ld[i], ud[t], a[i][kk]));
            }
        }
    }
}

// v[sp] of lane lp (sp, lp uniform): every slot is read by v_readlane and the choice is made
// on the scalar side (a select over the array itself would be turned into an indexed
// load from scratch memory).
template <int RPL>
__device__ inline float RlSel(const float (&v)[RPL], int sp, int lp)
{
    float r = Rl(v[0], lp);
#pragma unroll
    for (int i = 1; i < RPL; ++i)
    {
        const float t = Rl(v[i], lp);
        r = __int_as_float(sp == i ? __float_as_int(t) : __float_as_int(r));
    }
    return r;
}

template <int CTRL>
__device__ inline unsigned DppU32(unsigned v)
{
    return static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, 0xF, 0xF, false));
}
// Wave-wide max (MAX = true) or min of a 32-bit unsigned value, returned uniform.
template <bool MAX>
__device__ inline unsigned WaveReduceU32(unsigned v)
{
    auto op = [](unsigned a, unsigned b) { return MAX ? (a > b ? a : b) : (a < b ? a : b); };
    v = op(v, DppU32<0xB1>(v));   // quad_perm [1,0,3,2]
    v = op(v, DppU32<0x4E>(v));   // quad_perm [2,3,0,1]
    v = op(v, DppU32<0x124>(v));  // row_ror:4
    v = op(v, DppU32<0x128>(v));  // row_ror:8
    const unsigned r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const unsigned r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return op(op(r0, r1), op(r2, r3));
}

// Pivot of column k from the owner's values col (rows by slot): among the rows not yet
// pivoted, the largest |a| and, on ties, the lowest position; a NaN on the diagonal
// position k is the pivot, a NaN elsewhere never wins (src/utils/kernel/real_lu.hpp:100-109).
// A 32-bit key orders the magnitudes (0: not a candidate, 1: NaN below the diagonal,
// 0xffffffff: NaN on the diagonal, else bits(|a|) + 2) and is reduced across the wave in one
// pass; ties across lanes are rare and resolved by a second (min position) pass. Then the
// multipliers of the rows below the pivot after the interchange (reciprocal scaling when
// |d| >= FLT_MIN, division otherwise) replace their values in col.
template <int RPL>
__device__ __attribute__((always_inline)) inline void PivotColumn(float (&col)[RPL], const int (&pos)[RPL], int k,
                                                                  float (&l)[RPL], int& bpos, int& lp, int& sp,
                                                                  float& d)
{
    unsigned key[RPL];
    unsigned bkey = 0u;
#pragma unroll
    for (int i = 0; i < RPL; ++i)  // branch-free: bitwise logic and selects only
    {
        const bool cand = (pos[i] >= k) & (pos[i] != INT_MAX);
        const float v = col[i];
        unsigned kv = (__float_as_uint(v) & 0x7fffffffu) + 2u;
        kv = v != v ? (pos[i] == k ? 0xffffffffu : 1u) : kv;
        key[i] = cand ? kv : 0u;
        bkey = bkey > key[i] ? bkey : key[i];
    }
    int bp = INT_MAX;  // lowest position among this lane's rows holding bkey
#pragma unroll
    for (int i = 0; i < RPL; ++i) bp = min(bp, key[i] == bkey ? pos[i] : INT_MAX);
    int bi = 0;
    float bv = col[0];
#pragma unroll
    for (int i = 1; i < RPL; ++i)
    {
        bi = pos[i] == bp ? i : bi;
        bv = pos[i] == bp ? col[i] : bv;
    }
    const unsigned top = WaveReduceU32<true>(bkey);
    const unsigned long long hit = __builtin_amdgcn_ballot_w64(bkey == top);
    if (__builtin_popcountll(hit) == 1)
        lp = __builtin_ctzll(hit);
    else  // several lanes hold the largest magnitude: the lowest position wins
    {
        const unsigned mp = WaveReduceU32<false>(bkey == top ? static_cast<unsigned>(bp) : 0xffffffffu);
        lp = __builtin_ctzll(__builtin_amdgcn_ballot_w64(bkey == top && static_cast<unsigned>(bp) == mp));
    }
    bpos = __builtin_amdgcn_readlane(bp, lp);
    d = Rl(bv, lp);
    sp = __builtin_amdgcn_readlane(bi, lp);
    int np[RPL];
    bool below[RPL];
#pragma unroll
    for (int i = 0; i < RPL; ++i)
    {
        np[i] = pos[i] == bpos ? k : (pos[i] == k ? bpos : pos[i]);
        below[i] = (np[i] > k) & (np[i] != INT_MAX) & (d != 0.0f);
    }
    if (fabsf(d) >= FLT_MIN)  // uniform
    {
        const float rcp = 1.0f / d;
#pragma unroll
        for (int i = 0; i < RPL; ++i) l[i] = below[i] ? col[i] * rcp : 0.0f;
    }
    else
    {
#pragma unroll
        for (int i = 0; i < RPL; ++i) l[i] = below[i] ? col[i] / d : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < RPL; ++i) col[i] = (below[i] & (l[i] != 0.0f)) ? l[i] : col[i];
}

// Back substitution over the columns k of position segment S (64 S <= k < 64 S + 64, k < n),
// last to first, then the segments below. Lane l holds y at positions l + 64 i; column k of
// U (rows 0 .. 64 S + 63, slots 0..S) is read from the stored LU with a prefetch of D columns
// in registers. x_k = y_k * (1 / U_kk): the reciprocal depends only on U, so it stays off
// the dependent chain that runs through the y values.
template <int S, int RPL, int Q>
__device__ __attribute__((always_inline)) inline void BackSubSeg(const float* __restrict__ A, int n, int lane,
                                                                 float (&y)[Q][RPL])
{
    constexpr int D = 8;
    const int kLo = 64 * S;
    const int kHi = min(n, 64 * S + 64) - 1;
    float u[D][S + 1];
#pragma unroll
    for (int t = 0; t < D; ++t)
    {
        const int k = kHi - t;
#pragma unroll
        for (int i = 0; i <= S; ++i)
            u[t][i] = (k >= kLo && lane + 64 * i < n) ? A[static_cast<size_t>(k) * n + lane + 64 * i] : 0.0f;
    }
    for (int k0 = kHi; k0 >= kLo; k0 -= D)
    {
#pragma unroll
        for (int t = 0; t < D; ++t)
        {
            const int k = k0 - t;
            if (k >= kLo)
            {
                const int kl = k - kLo;
                const float r = 1.0f / Rl(u[t][S], kl);
#pragma unroll
                for (int q = 0; q < Q; ++q)
                {
                    const float x = Rl(y[q][S], kl) * r;
#pragma unroll
                    for (int i = 0; i < S; ++i) y[q][i] = fmaf(-u[t][i], x, y[q][i]);
                    y[q][S] = lane < kl ? fmaf(-u[t][S], x, y[q][S]) : (lane == kl ? x : y[q][S]);
                }
            }
            const int kn = k - D;
#pragma unroll
            for (int i = 0; i <= S; ++i)
                u[t][i] = (kn >= kLo && lane + 64 * i < n) ? A[static_cast<size_t>(kn) * n + lane + 64 * i] : 0.0f;
        }
    }
    if constexpr (S > 0) BackSubSeg<S - 1, RPL, Q>(A, n, lane, y);
}

// Column slots per wave of the column-cyclic solve (a multiple of 4).
template <int RPL, int NW>
constexpr int CyclicCpw() { return ((64 * RPL + kNrhsMax + NW - 1) / NW + 3) / 4 * 4; }
template <int RPL, int NW>
using CyclicLds = CycLds<RPL, NW, CyclicCpw<RPL, NW>()>;

// One matrix (mi; -1: none) by the NW waves with local thread indices tid = 0 .. 64 NW - 1.
// Every call runs exactly two workgroup barriers, so groups of waves solving different
// matrices of one item (and groups without a matrix) stay in step.
template <int RPL, int NW>
__device__ __attribute__((always_inline)) void CyclicSolve(int mi, int tid, CyclicLds<RPL, NW>& L,
                                                           const int* __restrict__ nArr,
                                                           const long long* __restrict__ pivOff, int nrhs,
                                                           float* const* __restrict__ Aarr,
                                                           float* const* __restrict__ Barr, int* __restrict__ ipiv,
                                                           int* __restrict__ info)
{
    constexpr int CPW = CyclicCpw<RPL, NW>();  // column slots per wave
    constexpr int kRing = RingSize<NW>();
    static_assert(NW >= kNrhsMax, "the back substitution takes one wave per right-hand side");
    if (mi < 0)
    {
        __syncthreads();
        __syncthreads();
        return;
    }
    const int n = nArr[mi];
    float* __restrict__ A = Aarr[mi];
    float* __restrict__ B = Barr[mi];
    int* __restrict__ piv = ipiv + pivOff[mi];
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    if (tid < kRing) L.tag[tid] = 0ull;
    if (tid < NW) L.done[tid] = 0;
    if (tid == 0)
    {
        L.zmin = INT_MAX;
        L.abort = 0;
    }
    for (int t = tid; t < NW * (CPW / 4); t += 64 * NW)  // no deferred pivot row yet
        L.prow[t / (CPW / 4)][1][t % (CPW / 4)] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);

    float a[RPL][CPW];
    int pos[RPL];
#pragma unroll
    for (int i = 0; i < RPL; ++i) pos[i] = lane + 64 * i < n ? lane + 64 * i : INT_MAX;
#pragma unroll
    for (int kk = 0; kk < CPW; ++kk)
    {
        const int j = w + NW * kk;
        const float* src = j < n ? A + static_cast<size_t>(j) * n
                                 : (j < n + nrhs ? B + static_cast<size_t>(j - n) * n : nullptr);
#pragma unroll
        for (int i = 0; i < RPL; ++i) a[i][kk] = (src != nullptr && pos[i] != INT_MAX) ? src[lane + 64 * i] : 0.0f;
    }
    __syncthreads();

    int zero = 0;
    float lc[RPL];  // multipliers of the current column (this lane's rows)
    int bp = 0, lp = 0, sp = 0;

    int ringFree = kRing - 1;  // largest column whose ring entry is known to be free
    auto publish = [&](int k, const float (&l)[RPL], int bpos, int plane, int pslot)
    {
        for (int it = 0; k > ringFree; ++it)  // the entry's previous column must have been taken
        {
            int mn = INT_MAX;
#pragma unroll
            for (int v = 0; v < NW; ++v)
                mn = min(mn, __hip_atomic_load(&L.done[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
            ringFree = __builtin_amdgcn_readfirstlane(mn) + kRing - 1;
            if (k <= ringFree) break;
            if (Aborted(L.abort, it)) break;
            __builtin_amdgcn_s_sleep(1);
        }
        const int e = k % kRing;
        float4 lv = make_float4(l[0], RPL > 1 ? l[RPL > 1 ? 1 : 0] : 0.0f, RPL > 2 ? l[RPL > 2 ? 2 : 0] : 0.0f,
                                RPL > 3 ? l[RPL > 3 ? 3 : 0] : 0.0f);
        L.l[e][lane] = lv;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        const unsigned meta = static_cast<unsigned>(bpos) | (static_cast<unsigned>(plane) << 10) |
                              (static_cast<unsigned>(pslot) << 16);
        if (lane == 0)
        {
            __hip_atomic_store(&L.tag[e], static_cast<unsigned long long>(k + 1) | (static_cast<unsigned long long>(meta) << 32),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            piv[k] = bpos + 1;
        }
    };

    auto take = [&](int c)  // pivot information of column c from its owner
    {
        const int e = c % kRing;
        unsigned long long tg = 0ull;
        float4 lv;
        for (int it = 0;; ++it)
        {
            // The multipliers are read right behind the tag: LDS serves one wave's reads in
            // order, so when the tag matches, the entry's data (written before the tag by
            // its owner, behind a release fence) is what the second read returned.
            tg = __hip_atomic_load(&L.tag[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);  // keep the data read behind the tag read
            lv = L.l[e][lane];
            if (__builtin_amdgcn_readfirstlane(static_cast<unsigned>(tg)) == static_cast<unsigned>(c + 1)) break;
            if (Aborted(L.abort, it)) break;
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const unsigned meta = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(tg >> 32));
        bp = meta & 1023u;
        lp = (meta >> 10) & 63u;
        sp = meta >> 16;
        lc[0] = lv.x;
        if constexpr (RPL > 1) lc[RPL > 1 ? 1 : 0] = lv.y;
        if constexpr (RPL > 2) lc[RPL > 2 ? 2 : 0] = lv.z;
        if constexpr (RPL > 3) lc[RPL > 3 ? 3 : 0] = lv.w;
    };
    auto taken = [&](int c)  // the ring entry of column c may be reused
    {
        if (lane == 0) __hip_atomic_store(&L.done[w], c + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    auto interchange = [&](int c)
    {
#pragma unroll
        for (int i = 0; i < RPL; ++i) pos[i] = pos[i] == bp ? c : (pos[i] == c ? bp : pos[i]);
    };
    float4* __restrict__ rowC = L.prow[w][0];  // this wave's copy of the current pivot row
    float4* __restrict__ rowD = L.prow[w][1];  // ... and of the deferred step's pivot row
    // A finished (pivoted) column goes to memory at once, by original row index; a pass at
    // the end moves every row to its final position.
    auto storeCol = [&](int j, const float (&col)[RPL])
    {
        float* dst = A + static_cast<size_t>(j) * n;
#pragma unroll
        for (int i = 0; i < RPL; ++i)
            if (pos[i] != INT_MAX) dst[lane + 64 * i] = col[i];
    };

    // Slot kk of this wave holds column w + NW * (kk + rot); the live ones are slots [0, nv).
    int rot = 0;
    int nv = (n + nrhs - w + NW - 1) / NW;
    if (w == 0)  // column 0
    {
        float col[RPL];
#pragma unroll
        for (int i = 0; i < RPL; ++i) col[i] = a[i][0];
        float d;
        PivotColumn(col, pos, 0, lc, bp, lp, sp, d);
        if (d == 0.0f) zero = 1;
        publish(0, lc, bp, lp, sp);
        storeCol(0, col);
#pragma unroll
        for (int kk = 1; kk < CPW; ++kk)
#pragma unroll
            for (int i = 0; i < RPL; ++i) a[i][kk - 1] = a[i][kk];
        ++rot;
        --nv;
    }
    // Each pass: plain steps, then (when there is room) a step that updates only slot 0 (this
    // wave's next pivot column) and defers the rest of its update, then the look-ahead step:
    // it brings slot 0 up to date, pivots it and publishes it at raised wave priority, and
    // only then applies the deferred and the current update to the other slots. The pivot
    // chain from column to column thus never waits for a wave's bulk update.
    int c = 0;
    for (;;)
    {
        const int cla = c + ((w - 1 - c) % NW + NW) % NW;  // next c with (c + 1) % NW == w
        const int stop = cla < n - 1 ? cla - 1 : n;
        for (; c < stop; ++c)
        {
            if (c % NW != w) take(c);  // else: this wave published column c itself
            taken(c);
            interchange(c);
            if (lane == lp) PutPivotRow<0>(a, sp, nv, rowC);
            UpdateSlots(a, lc, nv, rowC);
        }
        if (c >= n) break;
        float ld[RPL];
        if (c < cla)  // c == cla - 1 (never this wave's own column, NW >= 4): slot 0 only
        {
            take(c);
            taken(c);
            interchange(c);
            if (lane == lp) PutPivotRow<0>(a, sp, nv, rowD);
            float col[RPL];
#pragma unroll
            for (int i = 0; i < RPL; ++i) col[i] = a[i][0];
            const float u = RlSel(col, sp, lp);
#pragma unroll
            for (int i = 0; i < RPL; ++i) a[i][0] = fmaf(-lc[i], u, a[i][0]);
#pragma unroll
            for (int i = 0; i < RPL; ++i) ld[i] = lc[i];
            ++c;
        }
        else  // nothing deferred: rowD still holds its initial zeros (first pass of wave 1)
        {
#pragma unroll
            for (int i = 0; i < RPL; ++i) ld[i] = 0.0f;
        }
        // c == cla < n - 1: column c + 1 is this wave's slot 0
        __builtin_amdgcn_s_setprio(2);
        take(c);
        interchange(c);
        float col[RPL];
#pragma unroll
        for (int i = 0; i < RPL; ++i) col[i] = a[i][0];
        const float u = RlSel(col, sp, lp);
#pragma unroll
        for (int i = 0; i < RPL; ++i) col[i] = fmaf(-lc[i], u, col[i]);
        float ln[RPL];
        int bpn, lpn, spn;
        float d;
        PivotColumn(col, pos, c + 1, ln, bpn, lpn, spn, d);
        if (d == 0.0f && zero == 0) zero = c + 2;
        publish(c + 1, ln, bpn, lpn, spn);
        __builtin_amdgcn_s_setprio(0);
        taken(c);
        if (lane == lp) PutPivotRow<0>(a, sp, nv, rowC);
        const float ldp = RlSel(ld, sp, lp);
        RotSlots(a, ld, ldp, lc, nv, rowD, rowC);
        storeCol(c + 1, col);
        ++rot;
        --nv;
#pragma unroll
        for (int i = 0; i < RPL; ++i) lc[i] = ln[i];
        bp = bpn;
        lp = lpn;
        sp = spn;
        ++c;
    }

    // The right-hand sides (the live slots left) to LDS by position; this wave's finished
    // columns from original row order to final row order (all loads of a column complete
    // before its stores).
#pragma unroll
    for (int kk = 0; kk < CPW; ++kk)
    {
        const int j = w + NW * (kk + rot);
        if (kk < nv && j >= n && j < n + nrhs)
#pragma unroll
            for (int i = 0; i < RPL; ++i)
                if (pos[i] != INT_MAX) L.y[j - n][pos[i]] = a[i][kk];
    }
    for (int j = w; j < n; j += NW)
    {
        float* col = A + static_cast<size_t>(j) * n;
        float v[RPL];
#pragma unroll
        for (int i = 0; i < RPL; ++i) v[i] = pos[i] != INT_MAX ? col[lane + 64 * i] : 0.0f;
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): every lane's loads are done
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
        for (int i = 0; i < RPL; ++i)
            if (pos[i] != INT_MAX) col[pos[i]] = v[i];
    }
    if (zero != 0 && lane == 0) atomicMin(&L.zmin, zero);
    __syncthreads();
    if (tid == 0) info[mi] = L.abort ? -1 : (L.zmin == INT_MAX ? 0 : L.zmin);
    if (w >= nrhs) return;  // NW >= kNrhsMax: one wave per right-hand side

    // Back substitution U x = y of right-hand side w by wave w: lane l holds positions
    // l + 64 s. Column k of U comes from the stored LU; x_k = y_k * (1 / U_kk); rows above k
    // fold x_k in.
    float y[1][RPL];
#pragma unroll
    for (int i = 0; i < RPL; ++i) y[0][i] = lane + 64 * i < n ? L.y[w][lane + 64 * i] : 0.0f;
    BackSubSeg<RPL - 1, RPL, 1>(A, n, lane, y);
#pragma unroll
    for (int i = 0; i < RPL; ++i)
        if (lane + 64 * i < n) B[static_cast<size_t>(w) * n + lane + 64 * i] = y[0][i];
}

template <int A_, int B_>
constexpr int CMax() { return A_ > B_ ? A_ : B_; }

// LDS of one item: the largest over the configurations of (matrices per item) x (LDS per matrix).
constexpr int kItemLds =
    CMax<CMax<CMax<16 * static_cast<int>(sizeof(GroupLds<64, 1>)), 4 * static_cast<int>(sizeof(CyclicLds<2, 4>))>(),
              CMax<2 * static_cast<int>(sizeof(CyclicLds<2, 8>)), 2 * static_cast<int>(sizeof(CyclicLds<3, 8>))>()>(),
         CMax<static_cast<int>(sizeof(CyclicLds<3, 16>)), static_cast<int>(sizeof(CyclicLds<4, 16>))>()>();

// The waves of an item in groups of NW, one matrix each (order[first + group], when there).
template <int RPL, int NW>
__device__ __attribute__((always_inline)) void CyclicItem(const Item& item, int tid, unsigned char* smem,
                                                          const int* __restrict__ order, const int* __restrict__ nArr,
                                                          const long long* __restrict__ pivOff, int nrhs,
                                                          float* const* __restrict__ Aarr,
                                                          float* const* __restrict__ Barr, int* __restrict__ ipiv,
                                                          int* __restrict__ info)
{
    const int grp = __builtin_amdgcn_readfirstlane(tid / (64 * NW));
    const int mi = grp < item.count ? order[item.first + grp] : -1;
    CyclicLds<RPL, NW>* L = reinterpret_cast<CyclicLds<RPL, NW>*>(smem) + grp;
    CyclicSolve<RPL, NW>(mi, tid - grp * 64 * NW, *L, nArr, pivOff, nrhs, Aarr, Barr, ipiv, info);
}

// The persistent solve: one 1024-thread workgroup per CU takes items until the queue
// (counters[0]) runs dry; the last workgroup out (counters[1]) resets both counters, so the
// next launch of the plan starts from an empty queue with no separate reset.
__global__ __launch_bounds__(1024) void vb_persistent_kernel(const Item* __restrict__ items, int numItems,
                                                             const int* __restrict__ order,
                                                             const int* __restrict__ nArr,
                                                             const long long* __restrict__ pivOff, int nrhs,
                                                             float* const* __restrict__ Aarr,
                                                             float* const* __restrict__ Barr, int* __restrict__ ipiv,
                                                             int* __restrict__ info, int* counters)
{
    __shared__ __attribute__((aligned(16))) unsigned char smem[kItemLds];
    __shared__ int s_item[2];  // by iteration parity: read after one barrier, rewritten after the next
    const int tid = threadIdx.x;
    int next = 0;              // thread 0: the item fetched ahead
    if (tid == 0) next = atomicAdd(&counters[0], 1);
    for (int iter = 0;; ++iter)
    {
        if (tid == 0) s_item[iter & 1] = next;
        __syncthreads();
        const int it = s_item[iter & 1];
        if (it >= numItems) break;  // every workgroup reaches this exit once the queue drains
        // The following item's index, fetched while this one runs.
        if (tid == 0) next = atomicAdd(&counters[0], 1);
        const Item item = items[it];
        switch (item.cfg)  // uniform
        {
        case kW64:
        {
            const int g = __builtin_amdgcn_readfirstlane(tid >> 6);
            const bool live = g < item.count;
            const int mi = live ? order[item.first + g] : 0;
            const int n = live ? nArr[mi] : 0;
            solve_group<64, 1, true>(n, n, nrhs, live ? Aarr[mi] : nullptr, live ? Barr[mi] : nullptr,
                                     ipiv + (live ? pivOff[mi] : 0), info + mi, tid & 63, live,
                                     reinterpret_cast<GroupLds<64, 1>*>(smem)[g]);
            break;
        }
        case kC2x4: CyclicItem<2, 4>(item, tid, smem, order, nArr, pivOff, nrhs, Aarr, Barr, ipiv, info); break;
        case kC2x8: CyclicItem<2, 8>(item, tid, smem, order, nArr, pivOff, nrhs, Aarr, Barr, ipiv, info); break;
        case kC3x8: CyclicItem<3, 8>(item, tid, smem, order, nArr, pivOff, nrhs, Aarr, Barr, ipiv, info); break;
        case kC3x16: CyclicItem<3, 16>(item, tid, smem, order, nArr, pivOff, nrhs, Aarr, Barr, ipiv, info); break;
        default: CyclicItem<4, 16>(item, tid, smem, order, nArr, pivOff, nrhs, Aarr, Barr, ipiv, info); break;
        }
    }
    if (tid == 0 && atomicAdd(&counters[1], 1) == static_cast<int>(gridDim.x) - 1)
    {
        __hip_atomic_store(&counters[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&counters[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Synthetic systems for the autotuner: count matrices of order n at A + i * n * n (diagonally
// dominant, deterministic), right-hand sides of ones, pointer arrays.
__global__ void tune_fill_kernel(int n, int count, int nrhs, float* A, float* B, float** Ap, float** Bp)
{
    const size_t per = static_cast<size_t>(n) * n, total = per * count;
    for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * blockDim.x)
    {
        const int i = static_cast<int>(e / per);
        const int c = static_cast<int>((e % per) / n), r = static_cast<int>(e % n);
        const float off = 0.25f * static_cast<float>(((r * 7 + c * 13 + i) % 17) - 8) / (8.0f * n);
        A[e] = (r == c ? 2.0f : 0.0f) + off;
    }
    for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         e < static_cast<size_t>(count) * n * (nrhs > 0 ? nrhs : 1); e += static_cast<size_t>(gridDim.x) * blockDim.x)
        B[e] = 1.0f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
    {
        Ap[i] = A + static_cast<size_t>(i) * n * n;
        Bp[i] = B + static_cast<size_t>(i) * n * (nrhs > 0 ? nrhs : 1);
    }
}

}  // namespace

struct VbatchedPlan
{
    int device = 0;
    int batch = 0;
    int nrhs = 0;
    int numItems = 0;
    int grid = 0;                // one workgroup per CU
    int schedule[kNumBins] = {}; // configuration of each size bin above 64
    Item* dItems = nullptr;
    int* dOrder = nullptr;
    int* dN = nullptr;
    long long* dPivOff = nullptr;
    int* dCounter = nullptr;     // queue head and exit count, zero between launches
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

static hipError_t LaunchItems(const Item* items, int numItems, int grid, const int* order, const int* nArr,
                              const long long* pivOff, int nrhs, float* const* dA, float* const* dB, int* dIpiv,
                              int* dInfo, int* counters, hipStream_t s)
{
    hipLaunchKernelGGL(vb_persistent_kernel, dim3(grid), dim3(1024), 0, s, items, numItems, order, nArr, pivOff, nrhs,
                       dA, dB, dIpiv, dInfo, counters);
    return hipGetLastError();
}

// One launch; capturable into a hipGraph from the caller's stream.
hipError_t VbatchedLaunch(VbatchedPlan* p, float* const* dA, float* const* dB, int* dIpiv, int* dInfo, hipStream_t s)
{
    return LaunchItems(p->dItems, p->numItems, p->grid, p->dOrder, p->dN, p->dPivOff, p->nrhs, dA, dB, dIpiv, dInfo,
                       p->dCounter, s);
}

// Items of the plan: matrices in decreasing order of n, each group of up to kCfgCap[cfg]
// consecutive matrices of one configuration an item (longest-processing-time order).
static std::vector<Item> BuildItems(const std::vector<int>& ns, const std::vector<int>& order, const int* schedule)
{
    std::vector<Item> items;
    Item open[kNumCfg];
    for (int c = 0; c < kNumCfg; ++c) open[c] = {0, 0, c, 0};
    for (int i = 0; i < static_cast<int>(order.size()); ++i)
    {
        const int n = ns[order[i]];
        const int c = n <= 64 ? kW64 : schedule[BinOf(n)];
        // Items hold consecutive entries of the order: start a new one when this
        // configuration's open item is full or is not the latest item of its kind.
        if (open[c].count > 0 && (open[c].count == kCfgCap[c] || open[c].first + open[c].count != i))
        {
            items.push_back(open[c]);
            open[c].count = 0;
        }
        if (open[c].count == 0) open[c].first = i;
        ++open[c].count;
    }
    for (int c = 0; c < kNumCfg; ++c)
        if (open[c].count > 0) items.push_back(open[c]);
    std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.first < y.first; });
    return items;
}

// Autotuner (per device, once per process): for each bin above 64 and each configuration
// that fits it, one launch of numCUs items (one per CU) of synthetic systems of the bin's
// largest order; the configuration with the least time per matrix wins. Bins whose only
// fitting configuration is C4x16 are not timed.
static std::mutex g_tuneMutex;
static std::map<int, std::array<int, kNumBins>> g_tuned;

static hipError_t Autotune(int device, int numCUs, int nrhs, int* schedule, hipStream_t s)
{
    std::lock_guard<std::mutex> lock(g_tuneMutex);
    auto hit = g_tuned.find(device);
    if (hit != g_tuned.end())
    {
        std::copy(hit->second.begin(), hit->second.end(), schedule);
        return hipSuccess;
    }
    size_t maxMats = 0, maxElems = 0;
    for (int c = 1; c < kNumCfg; ++c)
    {
        maxMats = std::max<size_t>(maxMats, static_cast<size_t>(numCUs) * kCfgCap[c]);
        maxElems = std::max<size_t>(maxElems, static_cast<size_t>(numCUs) * kCfgCap[c] * kCfgMaxN[c] * kCfgMaxN[c]);
    }
    float *dA = nullptr, *dB = nullptr;
    float **dAp = nullptr, **dBp = nullptr;
    int *dPiv = nullptr, *dInfo = nullptr, *dOrder = nullptr, *dN = nullptr, *dCnt = nullptr;
    long long* dOff = nullptr;
    Item* dItems = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&dA, maxElems * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dB, maxMats * 256 * kNrhsMax * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dAp, maxMats * sizeof(float*));
    if (e == hipSuccess) e = hipMalloc(&dBp, maxMats * sizeof(float*));
    if (e == hipSuccess) e = hipMalloc(&dPiv, maxMats * 256 * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&dInfo, maxMats * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&dOrder, maxMats * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&dN, maxMats * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&dOff, maxMats * sizeof(long long));
    if (e == hipSuccess) e = hipMalloc(&dItems, static_cast<size_t>(numCUs) * sizeof(Item));
    if (e == hipSuccess) e = hipMalloc(&dCnt, 256);
    if (e == hipSuccess) e = hipMemset(dCnt, 0, 256);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    std::vector<int> iota(maxMats);
    for (size_t i = 0; i < maxMats; ++i) iota[i] = static_cast<int>(i);
    if (e == hipSuccess) e = hipMemcpy(dOrder, iota.data(), maxMats * sizeof(int), hipMemcpyHostToDevice);
    std::array<int, kNumBins> best{};
    for (int b = 0; b < kNumBins && e == hipSuccess; ++b)
    {
        const int n = 64 + kBinW * (b + 1);
        double bestT = 1e30;
        best[b] = kC4x16;
        for (int c = 1; c < kNumCfg && e == hipSuccess; ++c)
        {
            if (kCfgMaxN[c] < n || (n > 192 && c == kC4x16)) continue;  // does not fit / nothing to compare
            const int cap = kCfgCap[c], mats = numCUs * cap;
            std::vector<Item> items(numCUs);
            for (int i = 0; i < numCUs; ++i) items[i] = {i * cap, cap, c, 0};
            std::vector<int> nv(mats, n);
            std::vector<long long> off(mats);
            for (int i = 0; i < mats; ++i) off[i] = static_cast<long long>(i) * n;
            e = hipMemcpy(dItems, items.data(), items.size() * sizeof(Item), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(dN, nv.data(), mats * sizeof(int), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(dOff, off.data(), mats * sizeof(long long), hipMemcpyHostToDevice);
            float t = 1e30f;
            for (int rep = 0; rep < 2 && e == hipSuccess; ++rep)
            {
                hipLaunchKernelGGL(tune_fill_kernel, dim3(1024), dim3(256), 0, s, n, mats, nrhs, dA, dB, dAp, dBp);
                e = hipEventRecord(e0, s);
                if (e == hipSuccess)
                    e = LaunchItems(dItems, numCUs, numCUs, dOrder, dN, dOff, nrhs, dAp, dBp, dPiv, dInfo, dCnt, s);
                if (e == hipSuccess) e = hipEventRecord(e1, s);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                t = std::min(t, ms);
            }
            const double perMatrix = static_cast<double>(t) / cap;
            if (e == hipSuccess && perMatrix < bestT)
            {
                bestT = perMatrix;
                best[b] = c;
            }
        }
    }
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dAp);
    (void)hipFree(dBp);
    (void)hipFree(dPiv);
    (void)hipFree(dInfo);
    (void)hipFree(dOrder);
    (void)hipFree(dN);
    (void)hipFree(dOff);
    (void)hipFree(dItems);
    (void)hipFree(dCnt);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return e;
    g_tuned[device] = best;
    std::copy(best.begin(), best.end(), schedule);
    return hipSuccess;
}

}  // namespace ops

using namespace ops;

namespace
{
// schedule: kNumBins configuration ids (1..5) for the bins 65..80, ..., 241..256, or null
// to take the autotuned schedule.
aclsolverStatus_t CreatePlan(aclsolverHandle_t handle, int batchSize, const int* nArray, int nrhs,
                             const int* schedule, aclsolverVbatchedPlan_t* plan)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    if (plan == nullptr || batchSize < 0 || (batchSize > 0 && nArray == nullptr) || nrhs < 0 || nrhs > kNrhsMax)
        return ACLSOLVER_STATUS_INVALID_VALUE;
    std::vector<int> ns(nArray, nArray + batchSize);
    for (int v : ns)
        if (v < 1 || v > 256) return ACLSOLVER_STATUS_INVALID_VALUE;
    if (schedule != nullptr)
        for (int b = 0; b < kNumBins; ++b)
            if (schedule[b] < 1 || schedule[b] >= kNumCfg || kCfgMaxN[schedule[b]] < 64 + kBinW * (b + 1))
                return ACLSOLVER_STATUS_INVALID_VALUE;
    VbatchedPlan* p = new (std::nothrow) VbatchedPlan();
    if (p == nullptr) return ACLSOLVER_STATUS_ALLOC_FAILED;
    p->batch = batchSize;
    p->nrhs = nrhs;
    (void)hipGetDevice(&p->device);
    p->grid = H(handle)->numCUs;
    hipStream_t s = Stream(handle);
    hipError_t e = hipSuccess;
    if (schedule != nullptr)
        std::copy(schedule, schedule + kNumBins, p->schedule);
    else
        e = Autotune(p->device, p->grid, kNrhsMax, p->schedule, s);
    std::vector<int> order(batchSize);
    for (int i = 0; i < batchSize; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ns[x] > ns[y]; });
    const std::vector<Item> items = BuildItems(ns, order, p->schedule);
    std::vector<long long> pivOff(batchSize + 1, 0);
    for (int i = 0; i < batchSize; ++i)
    {
        pivOff[i + 1] = pivOff[i] + ns[i];
        const double n = ns[i];
        p->flops += 2.0 / 3.0 * n * n * n + 2.0 * n * n * nrhs;
        p->bytes += (2.0 * n * n + 2.0 * n * nrhs) * 4.0 + n * 4.0 + 4.0;
    }
    p->numItems = static_cast<int>(items.size());
    if (batchSize > 0 && e == hipSuccess)
    {
        e = hipMalloc(&p->dItems, std::max<size_t>(1, items.size()) * sizeof(Item));
        if (e == hipSuccess) e = hipMalloc(&p->dOrder, batchSize * sizeof(int));
        if (e == hipSuccess) e = hipMalloc(&p->dN, batchSize * sizeof(int));
        if (e == hipSuccess) e = hipMalloc(&p->dPivOff, batchSize * sizeof(long long));
        if (e == hipSuccess) e = hipMalloc(&p->dCounter, 256);
        if (e == hipSuccess) e = hipMemset(p->dCounter, 0, 256);
        if (e == hipSuccess)
            e = hipMemcpy(p->dItems, items.data(), items.size() * sizeof(Item), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(p->dOrder, order.data(), batchSize * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(p->dN, ns.data(), batchSize * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(p->dPivOff, pivOff.data(), batchSize * sizeof(long long), hipMemcpyHostToDevice);
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
}  // namespace

OPS_API aclsolverStatus_t aclsolverSgesvVbatched_createPlan(aclsolverHandle_t handle, int batchSize,
                                                            const int* nArray, int nrhs,
                                                            aclsolverVbatchedPlan_t* plan)
{
    return CreatePlan(handle, batchSize, nArray, nrhs, nullptr, plan);
}

OPS_API aclsolverStatus_t aclsolverSgesvVbatched_createPlanWithSchedule(aclsolverHandle_t handle, int batchSize,
                                                                        const int* nArray, int nrhs,
                                                                        const int* schedule,
                                                                        aclsolverVbatchedPlan_t* plan)
{
    if (schedule == nullptr) return ACLSOLVER_STATUS_INVALID_VALUE;
    return CreatePlan(handle, batchSize, nArray, nrhs, schedule, plan);
}

OPS_API aclsolverStatus_t aclsolverSgesvVbatched_getSchedule(aclsolverVbatchedPlan_t plan, int* schedule)
{
    if (plan == nullptr || schedule == nullptr) return ACLSOLVER_STATUS_INVALID_VALUE;
    const VbatchedPlan* p = static_cast<const VbatchedPlan*>(plan);
    std::copy(p->schedule, p->schedule + kNumBins, schedule);
    return ACLSOLVER_STATUS_SUCCESS;
}

OPS_API aclsolverStatus_t aclsolverSgesvVbatched_execute(aclsolverHandle_t handle, aclsolverVbatchedPlan_t plan,
                                                         float* const dAarray[], float* const dBarray[], int* dIpiv,
                                                         int* dInfo)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
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
