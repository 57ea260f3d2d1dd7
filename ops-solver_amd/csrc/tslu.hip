// Tournament-pivoting panel leaf (TSLU, the panel step of communication-avoiding LU:
// Grigori, Demmel and Xiang, SIAM J. Matrix Anal. Appl. 32(4), 2011) for the mixed-precision
// solve family. Its row order is free (the refinement, not the pivot sequence, is what the
// caller relies on), so the tall panel leaf need not exchange a pivot between CUs at every
// column as partial pivoting does: the pivots of W columns are chosen by a reduction tree.
//
//   tslu_tree_kernel, one wave per node, 256 rows per node (4 per lane) in registers:
//     level 0: node b factors rows [256 b, 256 b + 256) of the W leaf columns with partial
//     pivoting (wave-wide key argmax, pivot row through LDS, rank-1 updates) and publishes
//     the ORIGINAL values of the W rows it chose as candidates;
//     level L > 0: the last of a node's 8 children to finish (an arrival counter; nobody
//     waits, so no co-residency is assumed) stacks the 8 x W candidates and repeats; the
//     root's factorization of its candidates is the factorization of the top W x W block:
//     its pivot order fixes the interchange sequence, its factored rows are L11 \ U11.
//   tslu_apply_kernel: every other row solves l U11 = a (row-wise substitution, U11 in
//     LDS) and lands at its final position; the other panel columns take the interchanges.
// Pivot keys and scaling are those of the partial-pivoting leaves: first maximum of |a|
// (lowest position on ties, a NaN on the diagonal position is the pivot, a NaN elsewhere
// ranks below every number), reciprocal scaling when |d| >= FLT_MIN, division below it, a
// zero pivot skips its column and is reported in info.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "internal.h"
#include "tslu.h"
#include "wave.h"

namespace ops
{
namespace
{
constexpr int W = kTsluWidth;  // leaf columns
constexpr int R = 4;           // rows per lane in every node
constexpr int kNodeRows = 64 * R;
constexpr int G = kNodeRows / W;  // children per node (a node stacks G x W candidates)
constexpr int kLevels = 3;  // 64 x 256 rows -> 8 nodes -> 1
constexpr int kNodes = kTsluMaxRows / kNodeRows;  // nodes of level 0 at most
static_assert(kNodeRows % W == 0 && 2 * W <= 64, "node layout");
static_assert(kNodes <= G * G, "three levels reach a single root");

struct Scratch
{
    unsigned* counters;  // kLevels x kNodes arrival counters, zero between launches
    // A node's candidates as tagged words (leaf tag << 32 | payload), written and read with
    // agent-scope atomics: no cache maintenance, each word valid on its own.
    unsigned long long* cand;    // kLevels x kNodes x (W x W): values, column-major
    unsigned long long* candId;  // kLevels x kNodes x W: panel-relative rows (~0u: none)
    float* top;          // W x W: the factored top block, column-major
    float* disp;         // W x W: original leaf values of the displaced top rows
    int* moves;          // [0] count, then (destination, source) global row pairs
    int* flags;          // m ints: (index in disp) + 1 at a displaced row's destination
};

constexpr size_t Align(size_t b) { return (b + 255) / 256 * 256; }

Scratch Layout(void* base)
{
    char* p = static_cast<char*>(base);
    Scratch s;
    s.counters = reinterpret_cast<unsigned*>(p);
    p += Align(sizeof(unsigned) * kLevels * kNodes);
    s.cand = reinterpret_cast<unsigned long long*>(p);
    p += Align(sizeof(unsigned long long) * kLevels * kNodes * W * W);
    s.candId = reinterpret_cast<unsigned long long*>(p);
    p += Align(sizeof(unsigned long long) * kLevels * kNodes * W);
    s.top = reinterpret_cast<float*>(p);
    p += Align(sizeof(float) * W * W);
    s.disp = reinterpret_cast<float*>(p);
    p += Align(sizeof(float) * W * W);
    s.moves = reinterpret_cast<int*>(p);
    p += Align(sizeof(int) * (1 + 4 * W));
    s.flags = reinterpret_cast<int*>(p);
    return s;
}

size_t FixedBytes()
{
    return Align(sizeof(unsigned) * kLevels * kNodes) + Align(sizeof(unsigned long long) * kLevels * kNodes * W * W) +
           Align(sizeof(unsigned long long) * kLevels * kNodes * W) + 2 * Align(sizeof(float) * W * W) +
           Align(sizeof(int) * (1 + 4 * W));
}

// Coherent word access between workgroups on different XCDs (relaxed, agent scope).
__device__ inline unsigned long long LoadWord(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void StoreWord(unsigned long long* p, unsigned tag, unsigned payload)
{
    __hip_atomic_store(p, (static_cast<unsigned long long>(tag) << 32) | payload, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Partial pivoting on the node's rows (row lane + 64 i holds node position pos[i], -1 for
// padding). Leaves s_sel[j] = the node row of pivot j (-1: no row left) and returns the
// first zero-pivot column + 1 (0: none).
__device__ __attribute__((always_inline)) inline int NodeGepp(float (&a)[R][W], int (&pos)[R], int wact, float (*s_prow)[W], int* s_sel)
{
    const int lane = threadIdx.x;
    int zeroCol = 0;
    if (lane < W) s_sel[lane] = -1;
#pragma unroll
    for (int j = 0; j < W; ++j)
    {
        if (j < wact)  // uniform
        {
            const int par = j & 1;
            unsigned long long best = 0ull;
            int bi = -1;
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                const bool cand = pos[i] >= j;
                const unsigned ab = __float_as_uint(fabsf(a[i][j]));
                const unsigned mag = ab <= 0x7f800000u ? ab + 1u : (pos[i] == j ? 0x7fffffffu : 0u);
                const unsigned long long k =
                    cand ? (1ull << 63) | (static_cast<unsigned long long>(mag) << 32) | static_cast<unsigned>(~pos[i])
                         : 0ull;
                bi = k > best ? i : bi;
                best = k > best ? k : best;
            }
            const unsigned long long top = WaveMaxU64Split(best);
            // top == 0: no candidate rows left in this node; s_sel[j] stays -1
            if (top != 0ull && best == top && bi >= 0)
#pragma unroll
                for (int i = 0; i < R; ++i)
                    if (i == bi)
#pragma unroll
                        for (int k = 0; k < W; ++k) s_prow[par][k] = a[i][k];
            __syncthreads();
            const int bpos = top != 0ull ? static_cast<int>(~static_cast<unsigned>(top & 0xffffffffull)) : j;
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                if (pos[i] == bpos)
                    pos[i] = j;
                else if (pos[i] == j)
                    pos[i] = bpos;
            }
            const float d = top != 0ull ? s_prow[par][j] : 0.0f;
            if (top != 0ull && d == 0.0f && zeroCol == 0) zeroCol = j + 1;
            const bool recip = fabsf(d) >= FLT_MIN;  // uniform
            const float r = 1.0f / d;
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                const bool below = pos[i] > j && d != 0.0f;
                float q = a[i][j] * r;
                if (!recip) q = a[i][j] / d;  // uniform branch
                const float l = below ? q : 0.0f;
                a[i][j] = below ? q : a[i][j];
#pragma unroll
                for (int k = 0; k < W; ++k)
                    if (k > j) a[i][k] = fmaf(-l, s_prow[par][k], a[i][k]);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i)
        if (pos[i] >= 0 && pos[i] < wact) s_sel[pos[i]] = lane + 64 * i;
    __syncthreads();
    return zeroCol;
}

__global__ __launch_bounds__(64) void tslu_tree_kernel(const float* __restrict__ A, int lda, int m, int c0, int wact,
                                                       int newPanel, int* __restrict__ ipiv, int* __restrict__ info,
                                                       int* __restrict__ rowsrc, Scratch s, int nb0, unsigned tag,
                                                       int* err, unsigned spinLimit)
{
    __shared__ float s_prow[2][W];
    __shared__ int s_sel[W];   // node row of pivot j
    __shared__ int s_id[W];    // root: panel row of pivot j
    __shared__ int s_ipiv[W];  // root: interchange sequence, panel-relative
    const int lane = threadIdx.x;
    const int mp = m - c0;
    const float* Ap = A + static_cast<size_t>(c0) * lda + c0;
    // An earlier leaf of this factorization gave up waiting: its factors are invalid already.
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;

    float a[R][W];
    int pos[R], id[R];
    int node = blockIdx.x, nodes = nb0, nch = 0;
    int zeroCol = 0;  // the root's: first zero pivot of the top block, + 1
    // Levels unrolled (at most kLevels: 64 level-0 nodes, then 8, then the root), so the
    // node's rows stay in registers throughout.
#pragma unroll
    for (int level = 0; level < kLevels; ++level)
    {
        if (level == 0)
        {
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                const int r = node * kNodeRows + lane + 64 * i;
                const bool valid = r < mp;
                id[i] = valid ? r : -1;
                pos[i] = valid ? lane + 64 * i : -1;
#pragma unroll
                for (int j = 0; j < W; ++j)
                    a[i][j] = (valid && j < wact) ? Ap[static_cast<size_t>(j) * lda + r] : 0.0f;
            }
        }
        else
        {
            // The children's words: their stores were issued before their arrival, so every
            // word carries this leaf's tag within a few polls (bounded; on a timeout the
            // leaf gives up and raises *err).
            const size_t cb0 = static_cast<size_t>((level - 1) * kNodes + node * G);
            bool good = false;
            for (unsigned spins = 0; spins < spinLimit && !good; ++spins)
            {
                bool mine = true;
#pragma unroll
                for (int i = 0; i < R; ++i)
                {
                    const int r = lane + 64 * i, child = r / W, jj = r % W;
                    int rid = -1;
                    if (child < nch)
                    {
                        const unsigned long long wid = LoadWord(s.candId + (cb0 + child) * W + jj);
                        mine = mine && static_cast<unsigned>(wid >> 32) == tag;
                        rid = static_cast<int>(static_cast<unsigned>(wid));
                    }
                    id[i] = rid;
                    pos[i] = rid >= 0 ? r : -1;
#pragma unroll
                    for (int j = 0; j < W; ++j)
                    {
                        float v = 0.0f;
                        if (child < nch && j < wact)
                        {
                            const unsigned long long wv = LoadWord(s.cand + (cb0 + child) * W * W + j * W + jj);
                            mine = mine && static_cast<unsigned>(wv >> 32) == tag;
                            v = __uint_as_float(static_cast<unsigned>(wv));
                        }
                        a[i][j] = rid >= 0 ? v : 0.0f;
                    }
                }
                good = __all(mine);
            }
            if (!good)
            {
                if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return;
            }
        }
        zeroCol = NodeGepp(a, pos, wact, s_prow, s_sel);
        if (nodes == 1) break;  // the root
        // Publish the original values of the chosen rows (lane j: pivot j; every word of the
        // node is written, 0 / ~0u for missing rows, so the parent's tag check completes).
        unsigned long long* out = s.cand + static_cast<size_t>(level * kNodes + node) * W * W;
        if (lane < W)
        {
            const int sel = lane < wact ? s_sel[lane] : -1;
            int rid = -1;
            if (sel >= 0 && level > 0)
            {
                const size_t cb = static_cast<size_t>((level - 1) * kNodes + node * G + sel / W);
                rid = static_cast<int>(static_cast<unsigned>(LoadWord(s.candId + cb * W + sel % W)));
            }
            else if (sel >= 0)
                rid = node * kNodeRows + sel;
#pragma unroll
            for (int k = 0; k < W; ++k)
                if (k < wact)
                {
                    float v = 0.0f;
                    if (sel >= 0 && level == 0)
                        v = Ap[static_cast<size_t>(k) * lda + rid];
                    else if (sel >= 0)  // already seen with this leaf's tag above
                    {
                        const size_t cb = static_cast<size_t>((level - 1) * kNodes + node * G + sel / W);
                        v = __uint_as_float(static_cast<unsigned>(LoadWord(s.cand + cb * W * W + k * W + sel % W)));
                    }
                    StoreWord(out + k * W + lane, tag, __float_as_uint(v));
                }
            StoreWord(s.candId + static_cast<size_t>(level * kNodes + node) * W + lane, tag, static_cast<unsigned>(rid));
        }
        // The last child of the parent to arrive factors the parent (no fence: the parent
        // checks the tag of every word it reads).
        const int parent = node / G;
        nch = min(G, nodes - parent * G);
        unsigned* ctr = s.counters + (level + 1) * kNodes + parent;
        unsigned old = 0u;
        if (lane == 0) old = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        old = __shfl(old, 0);
        if (old != static_cast<unsigned>(nch - 1)) return;
        // every child has arrived: ready for the next launch
        if (lane == 0) __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        node = parent;
        nodes = (nodes + G - 1) / G;
    }

    // ---- root: the pivots' factored rows are the top block ----
#pragma unroll
    for (int i = 0; i < R; ++i)
        if (pos[i] >= 0 && pos[i] < wact)
        {
            s_id[pos[i]] = id[i];
#pragma unroll
            for (int k = 0; k < W; ++k)
                if (k < wact) s.top[k * W + pos[i]] = a[i][k];
        }
    __syncthreads();
    // Interchange sequence: lane l tracks the current position of pivot l's row.
    int cpos = lane < wact ? s_id[lane] : -1;
#pragma unroll
    for (int j = 0; j < W; ++j)
        if (j < wact)
        {
            const int x = __builtin_amdgcn_readlane(cpos, j);
            if (lane > j && cpos == j) cpos = x;
            if (lane == 0) s_ipiv[j] = x;
        }
    __syncthreads();
    if (lane < wact) ipiv[c0 + lane] = c0 + s_ipiv[lane] + 1;
    if (lane == 0 && zeroCol != 0 && *info == 0) *info = c0 + zeroCol;
    // Moved rows: lane e < W ends at position e; lane W + t ends at pivot t's old position
    // when that lies beyond the top block (first occurrence). Sources by undoing the
    // interchanges in reverse.
    int p = -1;
    if (lane < W)
        p = lane < wact ? lane : -1;
    else
    {
        const int t = lane - W;
        if (t < wact && s_ipiv[t] >= wact)
        {
            p = s_ipiv[t];
            for (int tt = 0; tt < t; ++tt)
                if (s_ipiv[tt] == p) p = -1;
        }
    }
    int src = p;
    if (p >= 0)
        for (int jj = wact - 1; jj >= 0; --jj)
        {
            const int b = s_ipiv[jj];
            src = src == jj ? b : (src == b ? jj : src);
        }
    const bool moved = p >= 0 && src != p;
    const bool setSrc = moved || (newPanel && p >= 0 && p < wact);
    const int newSrc = setSrc ? (newPanel ? c0 + src : rowsrc[c0 + src]) : 0;
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned long long mv = __ballot(moved);
    const unsigned long long dv = __ballot(moved && p >= wact);
    if (moved)
    {
        const int slot = __popcll(mv & below);
        s.moves[1 + 2 * slot] = c0 + p;
        s.moves[2 + 2 * slot] = c0 + src;
    }
    if (lane == 0) s.moves[0] = __popcll(mv);
    if (moved && p >= wact)  // a displaced top row (src < wact): stage its original values
    {
        const int e = __popcll(dv & below);
        s.flags[c0 + p] = e + 1;
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (k < wact) s.disp[k * W + e] = Ap[static_cast<size_t>(k) * lda + src];
    }
    if (setSrc) rowsrc[c0 + p] = newSrc;  // after every lane's read (the store depends on it)
}

// Row part (blocks < rowBlocks, one row per thread): rows of the top block take the root's
// factored values; every other row solves l U11 = a column by column (the rank-1 update
// order of partial pivoting) and is stored in place, or at a pivot's old position for a
// displaced top row. Column part: one wave per other panel column applies the moves.
__global__ __launch_bounds__(256) void tslu_apply_kernel(float* __restrict__ A, int lda, int m, int j0, int c0,
                                                         int wact, int panelEnd, int newPanel,
                                                         int* __restrict__ rowsrc, int* __restrict__ ipiv, Scratch s,
                                                         int rowBlocks, const int* err)
{
    const int tid = threadIdx.x;
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
    {
        // The tree gave up (this leaf or an earlier one): no pivots were chosen. Leave a
        // valid identity interchange sequence and panel row sources, so that the callers'
        // row bookkeeping stays in range, and NaN in the leaf columns.
        if (static_cast<int>(blockIdx.x) < rowBlocks)
        {
            const int q = blockIdx.x * 256 + tid, mp = m - c0;
            if (q < mp)
            {
                float* Ap = A + static_cast<size_t>(c0) * lda + c0;
#pragma unroll
                for (int k = 0; k < W; ++k)
                    if (k < wact) Ap[static_cast<size_t>(k) * lda + q] = __builtin_nanf("");
                if (newPanel) rowsrc[c0 + q] = c0 + q;
                if (q < wact) ipiv[c0 + q] = c0 + q + 1;
            }
        }
        return;
    }
    if (static_cast<int>(blockIdx.x) < rowBlocks)
    {
        __shared__ float s_u[W][W];  // s_u[j][k] = U(k, j)
        for (int e = tid; e < W * W; e += 256) s_u[e / W][e % W] = s.top[e];
        __syncthreads();
        const int mp = m - c0;
        float* Ap = A + static_cast<size_t>(c0) * lda + c0;
        const int q = blockIdx.x * 256 + tid;
        if (q < wact)
        {
#pragma unroll
            for (int k = 0; k < W; ++k)
                if (k < wact) Ap[static_cast<size_t>(k) * lda + q] = s.top[k * W + q];
        }
        else if (q < mp)
        {
            const int f = s.flags[c0 + q];
            float a[W];
#pragma unroll
            for (int k = 0; k < W; ++k)
                a[k] = k < wact ? (f ? s.disp[k * W + f - 1] : Ap[static_cast<size_t>(k) * lda + q]) : 0.0f;
            if (f) s.flags[c0 + q] = 0;
#pragma unroll
            for (int j = 0; j < W; ++j)
                if (j < wact)  // uniform
                {
                    float v = a[j];
#pragma unroll
                    for (int k = 0; k < W; ++k)
                        if (k < j) v = fmaf(-a[k], s_u[j][k], v);
                    const float d = s_u[j][j];
                    float l = 0.0f;
                    if (d != 0.0f)  // uniform
                        l = fabsf(d) >= FLT_MIN ? v * (1.0f / d) : v / d;
                    Ap[static_cast<size_t>(j) * lda + q] = d != 0.0f ? l : v;
                    a[j] = l;
                }
            if (newPanel && !f) rowsrc[c0 + q] = c0 + q;
        }
        return;
    }
    const int wave = tid >> 6, lane = tid & 63;
    const int idx = (blockIdx.x - rowBlocks) * 4 + wave;
    const int left = c0 - j0, others = left + (panelEnd - c0 - wact);
    if (idx >= others) return;
    const int c = idx < left ? j0 + idx : c0 + wact + (idx - left);
    float* col = A + static_cast<size_t>(c) * lda;
    const int cnt = s.moves[0];
    int d = -1;
    float v = 0.0f;
    if (lane < cnt)
    {
        d = s.moves[1 + 2 * lane];
        v = col[s.moves[2 + 2 * lane]];
    }
    if (d >= 0) col[d] = v;  // after every lane's gather (the store depends on it)
}

__global__ void tslu_clear_kernel(unsigned* counters, int ncounters, int* flags, int m, unsigned long long* words,
                                  int nwords)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncounters) counters[i] = 0u;
    if (i < m) flags[i] = 0;
    if (i < nwords) words[i] = 0ull;  // tag 0: never a leaf's tag
}
}  // namespace

size_t TsluScratchBytes(int m) { return FixedBytes() + Align(sizeof(int) * static_cast<size_t>(m > 0 ? m : 1)); }

hipError_t TsluClear(void* scratch, int m, hipStream_t stream)
{
    Scratch s = Layout(scratch);
    // cand and candId are contiguous (both 256-byte multiples): cleared as one word range.
    const int nwords = kLevels * kNodes * W * W + kLevels * kNodes * W;
    int n = m > kLevels * kNodes ? m : kLevels * kNodes;
    n = n > nwords ? n : nwords;
    hipLaunchKernelGGL(tslu_clear_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, s.counters, kLevels * kNodes,
                       s.flags, m, s.cand, nwords);
    return hipGetLastError();
}

hipError_t TsluLeaf(float* A, int lda, int m, int j0, int c0, int wact, int panelEnd, int newPanel, int* ipiv,
                    int* info, int* rowsrc, void* scratch, unsigned tag, int* err, hipStream_t stream)
{
    const int mp = m - c0;
    if (wact < 1 || wact > W || mp < wact || mp > kTsluMaxRows || ipiv == nullptr) return hipErrorInvalidValue;
    Scratch s = Layout(scratch);
    const int nb0 = (mp + kNodeRows - 1) / kNodeRows;
    hipLaunchKernelGGL(tslu_tree_kernel, dim3(nb0), dim3(64), 0, stream, A, lda, m, c0, wact, newPanel, ipiv, info,
                       rowsrc, s, nb0, tag, err, SpinLimit());
    const int rowBlocks = (mp + 255) / 256;
    const int others = (c0 - j0) + (panelEnd - c0 - wact);
    const int colBlocks = (others + 3) / 4;
    hipLaunchKernelGGL(tslu_apply_kernel, dim3(rowBlocks + colBlocks), dim3(256), 0, stream, A, lda, m, j0, c0, wact,
                       panelEnd, newPanel, rowsrc, ipiv, s, rowBlocks, err);
    return hipGetLastError();
}
}  // namespace ops
