// Real LU factorization (sgetrf) and the Device column-major solve, gfx950.
//
// Algorithm: right-looking blocked LU with partial pivoting, outer block width NB.
// Reference: src/sgetrf/real_lu_blocked_kernel.cpp:1469-1499 (BlockedFactorLaunch::Run:
// panel, swap+TRSM, trailing product per block). The MI355X design:
//   * panel: a sequence of register-resident leaf factorizations (leaf_getf2_kernel), one
//     workgroup of 256/1024 threads holding every active row of W panel columns in VGPRs.
//     Per column the pivot is a wave64 argmax (DPP/cross-lane shuffles) followed by a
//     16-entry LDS reduction; the row swap, the scaling and the rank-1 update never leave
//     the CU. Pivot rule, zero-pivot info and reciprocal scaling follow
//     src/sgetrf/real_lu_blocked_kernel.cpp:342-389, :467-482 and
//     src/utils/kernel/real_lu.hpp:419-437.
//   * within the panel, each leaf's row interchanges are applied to the other panel columns
//     as a gather/scatter of the (<= 2W) moved rows, followed by the W x W unit-lower solve
//     and an MFMA update of the rest of the panel.
//   * after the panel, the composed interchanges of all NB pivots are applied to every
//     column outside the panel (one coalesced gather/scatter per column), then the NB x NB
//     unit-lower TRSM on the U12 block and the MFMA trailing update A22 -= L21 * U12.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <type_traits>

#include "lu_kernels.h"
#include "trsv.h"

namespace ops
{
namespace
{

constexpr int kMaxRows = 16384;  // largest m the register-resident leaf supports

// Workspace layout (ints): [0, m) panel row sources; [m, m + 4*NBMAX + 8) moved-row list;
// then the leaf map (1 + 2*128 ints); then the global permutation (m ints).
constexpr int NBMAX = 512;
struct Work
{
    int* rowsrc;
    int* moved;
    int* leafmap;
    int* perm;
};
inline Work Layout(void* base, int m)
{
    int* p = static_cast<int*>(base);
    Work w;
    w.rowsrc = p;
    w.moved = p + m;
    w.leafmap = w.moved + 4 * NBMAX + 8;
    w.perm = w.leafmap + 1 + 2 * 128 + 7;
    return w;
}
inline size_t WorkInts(int m) { return static_cast<size_t>(m) * 2 + 4 * NBMAX + 8 + 1 + 2 * 128 + 7; }

// ---------------------------------------------------------------------------------------
// Leaf: factor columns [c0, c0+wact) over rows [c0, m) held in registers.
// Thread t owns local rows t, t+T, t+2T, ... (R of them). Rows never move between
// registers: each row tracks the position it occupies under the LAPACK interchange
// sequence, the pivot of column j is the candidate (position >= j) with the largest |a|
// and, on ties, the lowest position -- exactly the first-row rule of a physical swap.
// At the end each row is stored at its final position, which applies the interchanges
// to the leaf columns.
// ---------------------------------------------------------------------------------------
template <int T, int R, int W>
__global__ __launch_bounds__(T) void leaf_getf2_kernel(float* __restrict__ A, int lda, int m, int j0, int c0, int wact,
                                                       int pivoting, int newPanel, int* __restrict__ ipiv,
                                                       int* __restrict__ info, int* __restrict__ rowsrc,
                                                       int* __restrict__ leafmap)
{
    constexpr int NW = T / 64;
    __shared__ float s_prow[W];
    __shared__ float s_best[NW];
    __shared__ int s_bpos[NW];
    __shared__ int s_count;

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
        pos[i] = valid ? lr : INT_MAX;
    }
    if (tid == 0) s_count = 0;
    int zeroCol = 0;

#pragma unroll
    for (int j = 0; j < W; ++j)
    {
        if (j < wact)  // uniform
        {
            int bpos = j;
            if (pivoting)
            {
                float best = -1.0f;
                int bp = INT_MAX;
#pragma unroll
                for (int i = 0; i < R; ++i)
                {
                    if (pos[i] >= j && pos[i] != INT_MAX)
                    {
                        const float v = fabsf(a[i][j]);
                        if (v > best || (v == best && pos[i] < bp))
                        {
                            best = v;
                            bp = pos[i];
                        }
                    }
                }
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
                    const float ob = s_best[w];
                    const int op = s_bpos[w];
                    if (ob > best || (ob == best && op < bp))
                    {
                        best = ob;
                        bp = op;
                    }
                }
                bpos = bp == INT_MAX ? j : bp;
            }
            // The pivot row publishes itself; positions j and bpos are exchanged.
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                if (pos[i] == bpos)
                {
#pragma unroll
                    for (int k = 0; k < W; ++k) s_prow[k] = a[i][k];
                }
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
            const float d = s_prow[j];
            if (d == 0.0f && zeroCol == 0) zeroCol = c0 + j + 1;
            // Branch-free update: rows that are not below the pivot take l = 0, for which
            // the FMA leaves them unchanged (a zero pivot skips the column likewise).
            const bool recip = fabsf(d) >= FLT_MIN;
            const float r = 1.0f / d;
#pragma unroll
            for (int i = 0; i < R; ++i)
            {
                const bool below = pos[i] > j && pos[i] != INT_MAX && d != 0.0f;
                const float q = recip ? a[i][j] * r : a[i][j] / d;
                const float l = below ? q : 0.0f;
                a[i][j] = below ? q : a[i][j];
#pragma unroll
                for (int k = j + 1; k < W; ++k) a[i][k] = fmaf(-l, s_prow[k], a[i][k]);
            }
            __syncthreads();
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
                leafmap[1 + 2 * slot] = c0 + dst;
                leafmap[2 + 2 * slot] = c0 + lr;
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
    if (tid == 0)
    {
        leafmap[0] = s_count;
        if (zeroCol != 0 && *info == 0) *info = zeroCol;
    }
}

// Applies the leaf's moved-row map to panel columns outside the leaf and, for the panel
// columns right of the leaf, solves the leaf's W x W unit-lower block on the leaf rows.
// One workgroup (128 threads) per column.
__global__ __launch_bounds__(128) void panel_swap_trsm_kernel(float* __restrict__ A, int lda, int j0, int c0,
                                                              int wact, int panelEnd, const int* __restrict__ leafmap)
{
    __shared__ float s_x[64];
    const int left = c0 - j0;
    const int idx = blockIdx.x;
    const int c = idx < left ? j0 + idx : c0 + wact + (idx - left);
    float* col = A + static_cast<size_t>(c) * lda;
    const int cnt = leafmap[0];
    const int t = threadIdx.x;
    float v = 0.0f;
    int dst = -1;
    if (t < cnt)
    {
        dst = leafmap[1 + 2 * t];
        v = col[leafmap[2 + 2 * t]];
    }
    __syncthreads();
    if (dst >= 0) col[dst] = v;
    __syncthreads();
    if (c < c0) return;  // left columns only take the interchanges
    if (t < wact) s_x[t] = col[c0 + t];
    __syncthreads();
    const float* L = A + static_cast<size_t>(c0) * lda + c0;
    for (int s = 0; s < wact - 1; ++s)
    {
        if (t > s && t < wact) s_x[t] = fmaf(-L[static_cast<size_t>(s) * lda + t], s_x[s], s_x[t]);
        __syncthreads();
    }
    if (t < wact) col[c0 + t] = s_x[t];
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

// Applies the panel's moved-row list to columns [0, j0) and [j0+nb, n): one wave64 per
// column gathers every source row, then scatters.
__global__ __launch_bounds__(256) void outer_swap_kernel(float* __restrict__ A, int lda, int j0, int nb, int n,
                                                         const int* __restrict__ moved)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int idx = blockIdx.x * 4 + wave;
    const int outside = j0 + (n - j0 - nb);
    if (idx >= outside) return;
    const int c = idx < j0 ? idx : idx + nb;
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

__global__ void iota_kernel(int* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

template <int T, int R, int W>
hipError_t LaunchLeaf(float* A, int lda, int m, int j0, int c0, int wact, int pivoting, int newPanel, int* ipiv,
                      int* info, const Work& w, hipStream_t stream)
{
    hipLaunchKernelGGL((leaf_getf2_kernel<T, R, W>), dim3(1), dim3(T), 0, stream, A, lda, m, j0, c0, wact, pivoting,
                       newPanel, ipiv, info, w.rowsrc, w.leafmap);
    return hipGetLastError();
}

// Leaf width for a panel of mp active rows: every shape keeps R*W <= 128 values per lane
// (512-thread workgroups leave up to 256 VGPRs per lane).
int LeafWidth(int mp)
{
    if (mp <= 512) return 32;
    if (mp <= 1024) return 16;
    if (mp <= 8192) return 8;
    return 4;
}

hipError_t Leaf(float* A, int lda, int m, int j0, int c0, int wact, int pivoting, int newPanel, int* ipiv, int* info,
                const Work& w, hipStream_t stream)
{
    const int mp = m - c0;
#define LEAF(T, R, W) LaunchLeaf<T, R, W>(A, lda, m, j0, c0, wact, pivoting, newPanel, ipiv, info, w, stream)
    if (mp <= 256) return LEAF(256, 1, 32);
    if (mp <= 512) return LEAF(512, 1, 32);
    if (mp <= 1024) return LEAF(512, 2, 16);
    if (mp <= 2048) return LEAF(512, 4, 8);
    if (mp <= 4096) return LEAF(512, 8, 8);
    if (mp <= 8192) return LEAF(512, 16, 8);
    return LEAF(512, 32, 4);
#undef LEAF
}

int OuterBlock(int m, int n, Precision prec)
{
    const int k = std::min(m, n);
    if (prec != Precision::FP32) return k >= 4096 ? 512 : (k >= 1024 ? 256 : 128);
    return k >= 2048 ? 256 : 128;
}

#define RET_IF(e)                          \
    do                                     \
    {                                      \
        hipError_t _e = (e);               \
        if (_e != hipSuccess) return _e;   \
    } while (0)

}  // namespace

size_t RealGetrfWorkspaceBytes(int m, int n)
{
    (void)n;
    return WorkInts(std::max(m, 1)) * sizeof(int);
}

// ipiv == nullptr selects no pivoting. perm (optional, m ints) receives the row
// permutation P with (P A)[i] = A[perm[i]].
hipError_t RealGetrfPerm(int m, int n, float* a, int lda, int* ipiv, int* info, void* work, Precision trailing,
                         int* permOut, hipStream_t stream)
{
    if (m > kMaxRows) return hipErrorInvalidValue;
    Work w = Layout(work, m);
    if (permOut) w.perm = permOut;
    FillInfo(info, 1, 0, stream);
    const int kmax = std::min(m, n);
    if (kmax == 0) return hipGetLastError();
    const int pivoting = ipiv != nullptr;
    if (permOut) hipLaunchKernelGGL(iota_kernel, dim3((m + 255) / 256), dim3(256), 0, stream, permOut, m);
    const int NB = OuterBlock(m, n, trailing);
    for (int j0 = 0; j0 < kmax; j0 += NB)
    {
        const int nb = std::min(NB, kmax - j0);
        const int panelEnd = j0 + nb;
        // Panel: leaves of width W with in-panel interchange, TRSM and update.
        for (int c0 = j0; c0 < panelEnd;)
        {
            const int W = LeafWidth(m - c0);
            const int wact = std::min(W, panelEnd - c0);
            RET_IF(Leaf(a, lda, m, j0, c0, wact, pivoting, c0 == j0, ipiv, info, w, stream));
            const int others = nb - wact;
            if (others > 0)
            {
                if (pivoting)
                    hipLaunchKernelGGL(panel_swap_trsm_kernel, dim3(others), dim3(128), 0, stream, a, lda, j0, c0,
                                       wact, panelEnd, w.leafmap);
                else if (panelEnd - c0 - wact > 0)
                    RET_IF(TrsmBlock(wact, panelEnd - c0 - wact, a + static_cast<size_t>(c0) * lda + c0, lda, true,
                                     true, false, a + static_cast<size_t>(c0 + wact) * lda + c0, lda, stream));
                RET_IF(hipGetLastError());
                const int right = panelEnd - c0 - wact;
                if (right > 0 && m - c0 - wact > 0)
                    RET_IF(GemmSub(m - c0 - wact, right, wact, a + static_cast<size_t>(c0) * lda + c0 + wact, lda, 0,
                                   a + static_cast<size_t>(c0 + wact) * lda + c0, lda,
                                   a + static_cast<size_t>(c0 + wact) * lda + c0 + wact, lda, Precision::FP32,
                                   stream));
            }
            c0 += wact;
        }
        // Interchanges outside the panel, U12 solve, trailing update.
        if (pivoting)
        {
            hipLaunchKernelGGL(compact_moves_kernel, dim3(1), dim3(1024), 0, stream, m, j0, w.rowsrc, w.moved,
                               permOut);
            const int outside = j0 + (n - panelEnd);
            if (outside > 0)
                hipLaunchKernelGGL(outer_swap_kernel, dim3((outside + 3) / 4), dim3(256), 0, stream, a, lda, j0, nb,
                                   n, w.moved);
            RET_IF(hipGetLastError());
        }
        const int right = n - panelEnd;
        if (right > 0)
        {
            RET_IF(TrsmBlock(nb, right, a + static_cast<size_t>(j0) * lda + j0, lda, true, true, false,
                             a + static_cast<size_t>(panelEnd) * lda + j0, lda, stream));
            if (m - panelEnd > 0)
            {
                ProfileScope prof(ProfileKernel::TrailingGemm, 2.0 * (m - panelEnd) * right * nb, stream);
                RET_IF(GemmSub(m - panelEnd, right, nb, a + static_cast<size_t>(j0) * lda + panelEnd, lda, 0,
                               a + static_cast<size_t>(panelEnd) * lda + j0, lda,
                               a + static_cast<size_t>(panelEnd) * lda + panelEnd, lda, trailing, stream));
            }
        }
    }
    return hipGetLastError();
}

hipError_t RealGetrf(int m, int n, float* a, int lda, int* ipiv, int* info, void* work, Precision trailing,
                     hipStream_t stream)
{
    return RealGetrfPerm(m, n, a, lda, ipiv, info, work, trailing, nullptr, stream);
}

// ---------------------------------------------------------------------------------------
// getrs on Device column-major data.
// ---------------------------------------------------------------------------------------
namespace
{
// ipiv (LAPACK interchange sequence) -> permutation, perm[i] = source row of row i of P*B.
// One workgroup; the sequential interchanges run on an LDS index array.
__global__ __launch_bounds__(1024) void ipiv_to_perm_kernel(int n, const int* __restrict__ ipiv, int* __restrict__ perm)
{
    extern __shared__ int s_idx[];
    for (int i = threadIdx.x; i < n; i += blockDim.x) s_idx[i] = i;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < n; ++k)
        {
            const int p = ipiv[k] - 1;
            if (p != k)
            {
                const int t = s_idx[k];
                s_idx[k] = s_idx[p];
                s_idx[p] = t;
            }
        }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = s_idx[i];
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

hipError_t IpivToPerm(int n, const int* ipiv, int* perm, hipStream_t stream)
{
    if (n > kMaxRows) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ipiv_to_perm_kernel, dim3(1), dim3(1024), static_cast<size_t>(n) * sizeof(int), stream, n, ipiv,
                       perm);
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
                     void* work, hipStream_t stream)
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

}  // namespace ops
