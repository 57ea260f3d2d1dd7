// Batched small-matrix inverses (smatinv_batched, sgetri_batched), gfx950.
//
// Reference: src/smatinv_batched/smatinv_batched_host.cpp:11-61 -> real_lu_small_do
// (src/sgetri/sgetri_kernel.cpp:168-198, SmallInverse::Run,
// src/utils/kernel/real_lu_small_inverse.hpp:426-468): per matrix an LU with partial
// pivoting followed by the inverse; info = first zero pivot.
//
// MI355X layout: NL lanes of a wave64 own one matrix (NL = 8, 16, 32 or 64 >= n), lane r
// holding row r in VGPRs, so a wave carries 64/NL matrices and every lane does useful FMAs.
// Column-major rows are loaded with one coalesced 4*NL-byte segment per column.
//
// Direct inverse: in-place Gauss-Jordan with implicit partial pivoting. Row interchanges
// are never performed; each lane tracks the position its row would occupy under the
// LAPACK interchange sequence, so the pivot is the candidate of largest |a| with the
// lowest position (the first-row tie rule of src/utils/kernel/real_lu.hpp:59-77) and
// ipiv is reproduced exactly. Rows that were already pivots keep a scale factor, which
// makes every update a single FMA with the broadcast pivot row; the rows that are still
// candidates are updated as fma(-(a_ik * (1/d)), u_j, a_ij), bit for bit the Schur
// complement of the reciprocal-scaled LU (src/sgetrf/real_lu_blocked_kernel.cpp:467-482),
// so pivot decisions match the LU. The pivot row is broadcast through a per-matrix LDS
// slot. At the end row i of the result is scaled and scattered to row pos(i) and its
// columns to pos^-1, which undoes the implicit interchanges.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cstdlib>

#include "internal.h"

namespace ops
{
namespace
{

// Pivot key: larger |a| first, then the lower (position, lane). NaN follows the
// reference's scan, whose running best starts at the diagonal candidate and is replaced
// only by a strictly larger magnitude (src/utils/kernel/real_lu_small_dispatch.hpp:270-286):
// a NaN on the diagonal position wins (high word 0x7fffffff), a NaN anywhere else never
// does (high word 0); numbers have |a| bits + 1. Non-candidates are 0.
__device__ inline unsigned long long GjKey(float v, int posLane, bool diagonal)
{
    const float a = fabsf(v);
    const unsigned hi = (a == a) ? __float_as_uint(a) + 1u : (diagonal ? 0x7fffffffu : 0u);
    return (1ull << 63) | (static_cast<unsigned long long>(hi) << 32) | static_cast<unsigned>(~posLane);
}

template <int CTRL>
__device__ inline unsigned long long GjDpp(unsigned long long v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(v & 0xffffffffull), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(v >> 32), CTRL, 0xF, 0xF, false);
    return (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) | static_cast<unsigned>(lo);
}

// Maximum key over each aligned NL-lane group (NL = 8, 16, 32 or 64), left in every lane
// of the group: DPP quad swaps and (half-)row mirrors within 16-lane rows, a swizzle across
// the two rows of a 32-lane half, lane reads across the halves.
template <int NL>
__device__ inline unsigned long long GroupMaxKey(unsigned long long v)
{
    auto mx = [](unsigned long long a, unsigned long long b) { return a > b ? a : b; };
    v = mx(v, GjDpp<0xB1>(v));                  // quad_perm [1,0,3,2]
    v = mx(v, GjDpp<0x4E>(v));                  // quad_perm [2,3,0,1]
    v = mx(v, GjDpp<0x141>(v));                 // row_half_mirror -> groups of 8
    if (NL >= 16) v = mx(v, GjDpp<0x140>(v));   // row_mirror -> rows of 16
    if (NL >= 32)
    {
        // xor 16 inside each 32-lane half (swizzle bit-mask mode: and 0x1f, xor 0x10)
        const int lo = __builtin_amdgcn_ds_swizzle(static_cast<int>(v & 0xffffffffull), 0x401F);
        const int hi = __builtin_amdgcn_ds_swizzle(static_cast<int>(v >> 32), 0x401F);
        v = mx(v, (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) | static_cast<unsigned>(lo));
    }
    if (NL >= 64)
    {
        auto rl = [](unsigned long long x, int lane)
        {
            const unsigned l = __builtin_amdgcn_readlane(static_cast<int>(x & 0xffffffffull), lane);
            const unsigned h = __builtin_amdgcn_readlane(static_cast<int>(x >> 32), lane);
            return (static_cast<unsigned long long>(h) << 32) | l;
        };
        v = mx(rl(v, 0), rl(v, 32));
    }
    return v;
}

// MATS matrices per workgroup of 256 threads (4 waves); each NL-lane group one matrix.
template <int NL>
__global__ __launch_bounds__(256) void matinv_gj_kernel(int n, const float* const* __restrict__ Aarr,
                                                        const float* __restrict__ Abase, int lda, int64_t strideA,
                                                        float* const* __restrict__ Carr, float* __restrict__ Cbase,
                                                        int ldc, int64_t strideC, int* __restrict__ ipivOut,
                                                        int* __restrict__ info, int batch)
{
    constexpr int GROUPS = 256 / NL;
    __shared__ __attribute__((aligned(16))) float s_row[GROUPS][NL];
    __shared__ int s_inv[GROUPS][NL];
    const int tid = threadIdx.x;
    const int g = tid / NL, lane = tid % NL;
    const int mat = blockIdx.x * GROUPS + g;
    const bool live = mat < batch;
    const float* A = nullptr;
    float* C = nullptr;
    if (live)
    {
        A = Aarr ? Aarr[mat] : Abase + static_cast<int64_t>(mat) * strideA;
        C = Carr ? Carr[mat] : Cbase + static_cast<int64_t>(mat) * strideC;
    }
    const bool rowLive = live && lane < n;

    float S[NL];
#pragma unroll
    for (int j = 0; j < NL; ++j) S[j] = (rowLive && j < n) ? A[static_cast<int64_t>(j) * lda + lane] : 0.0f;
    float sigma = 1.0f;
    int pos = lane;       // position under the interchange sequence
    bool used = false;    // already a pivot row
    int zero = 0;

#pragma unroll
    for (int k = 0; k < NL; ++k)
    {
        if (k >= n) continue;
        const unsigned long long myKey = (rowLive && !used) ? GjKey(S[k], (pos << 8) | lane, pos == k) : 0ull;
        const unsigned long long top = GroupMaxKey<NL>(myKey);
        // No candidate at all cannot happen for k < n (position k's row is a candidate).
        const int key = static_cast<int>(~static_cast<unsigned>(top & 0xffffffffull));
        const int p = key & 0xFF;
        const int ppos = key >> 8;
        const bool isP = lane == p;
        if (live && lane == 0 && ipivOut != nullptr) ipivOut[static_cast<int64_t>(mat) * n + k] = ppos + 1;
        // LAPACK interchange of positions k and ppos.
        if (pos == k) pos = ppos;
        if (isP) pos = k;
        // Broadcast the pivot row.
        if (isP)
        {
#pragma unroll
            for (int j = 0; j < NL; ++j) s_row[g][j] = S[j];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        float u[NL];
#pragma unroll
        for (int j = 0; j < NL; ++j) u[j] = s_row[g][j];
        const float d = u[k];
        if (d == 0.0f)
        {
            if (zero == 0) zero = k + 1;
            if (isP) used = true;
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        const float r = 1.0f / d;
        const float c = isP ? 0.0f : S[k] * r;
#pragma unroll
        for (int j = 0; j < NL; ++j)
            if (j != k) S[j] = fmaf(-c, u[j], S[j]);
        S[k] = isP ? 1.0f : -c;
        if (isP)
        {
            sigma = r;
            used = true;
        }
        __builtin_amdgcn_wave_barrier();
    }

    // Result: row `lane` (true values sigma*S) goes to row pos; column j goes to pos^-1(j).
    if (rowLive) s_inv[g][pos] = lane;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (rowLive)
    {
#pragma unroll
        for (int j = 0; j < NL; ++j)
            if (j < n)
            {
                const int col = s_inv[g][j];
                C[static_cast<int64_t>(col) * ldc + pos] = sigma * S[j];
            }
    }
    if (live && lane == 0) info[mat] = zero;
}

// The same Gauss-Jordan inverse with R rows per lane: NL lanes own one matrix (n <= NL*R),
// lane l holding rows l*R .. l*R+R-1, so a wave carries 64/NL matrices and each pivot step's
// search, broadcast and reciprocal are shared by R rows of multiply-adds. Pivot keys,
// scaling and the final scatter are those of matinv_gj_kernel.
template <int NL, int R>
__global__ __launch_bounds__(256) void matinv_gj_rows_kernel(int n, const float* const* __restrict__ Aarr,
                                                             const float* __restrict__ Abase, int lda, int64_t strideA,
                                                             float* const* __restrict__ Carr, float* __restrict__ Cbase,
                                                             int ldc, int64_t strideC, int* __restrict__ ipivOut,
                                                             int* __restrict__ info, int batch)
{
    constexpr int N = NL * R;  // columns held per row
    constexpr int GROUPS = 256 / NL;
    __shared__ __attribute__((aligned(16))) float s_row[GROUPS][N];
    __shared__ int s_inv[GROUPS][N];
    const int tid = threadIdx.x;
    const int g = tid / NL, lane = tid % NL;
    const int mat = blockIdx.x * GROUPS + g;
    const bool live = mat < batch;
    const float* A = nullptr;
    float* C = nullptr;
    if (live)
    {
        A = Aarr ? Aarr[mat] : Abase + static_cast<int64_t>(mat) * strideA;
        C = Carr ? Carr[mat] : Cbase + static_cast<int64_t>(mat) * strideC;
    }
    float S[R][N];
    float sigma[R];
    int pos[R];
    bool used[R], rowLive[R];
#pragma unroll
    for (int i = 0; i < R; ++i)
    {
        const int r = lane * R + i;
        rowLive[i] = live && r < n;
#pragma unroll
        for (int j = 0; j < N; ++j) S[i][j] = (rowLive[i] && j < n) ? A[static_cast<int64_t>(j) * lda + r] : 0.0f;
        sigma[i] = 1.0f;
        pos[i] = r;
        used[i] = false;
    }
    int zero = 0;

#pragma unroll
    for (int k = 0; k < N; ++k)
    {
        if (k >= n) continue;
        unsigned long long best = 0ull;
        int bi = -1;
#pragma unroll
        for (int i = 0; i < R; ++i)
        {
            const unsigned long long key =
                (rowLive[i] && !used[i]) ? GjKey(S[i][k], (pos[i] << 8) | (lane * R + i), pos[i] == k) : 0ull;
            bi = key > best ? i : bi;
            best = key > best ? key : best;
        }
        const unsigned long long top = GroupMaxKey<NL>(best);
        const int key = static_cast<int>(~static_cast<unsigned>(top & 0xffffffffull));
        const int p = key & 0xFF;  // pivot row
        const int ppos = key >> 8;
        if (live && lane == 0 && ipivOut != nullptr) ipivOut[static_cast<int64_t>(mat) * n + k] = ppos + 1;
        // LAPACK interchange of positions k and ppos.
#pragma unroll
        for (int i = 0; i < R; ++i)
        {
            const bool isP = lane * R + i == p;
            if (pos[i] == k) pos[i] = ppos;
            if (isP) pos[i] = k;
        }
        // Broadcast the pivot row.
        if (best == top && bi >= 0)
#pragma unroll
            for (int i = 0; i < R; ++i)
                if (i == bi)
#pragma unroll
                    for (int j = 0; j < N; ++j) s_row[g][j] = S[i][j];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        float u[N];
#pragma unroll
        for (int j = 0; j < N; ++j) u[j] = s_row[g][j];
        const float d = u[k];
        if (d == 0.0f)
        {
            if (zero == 0) zero = k + 1;
#pragma unroll
            for (int i = 0; i < R; ++i)
                if (lane * R + i == p) used[i] = true;
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        const float rc = 1.0f / d;
#pragma unroll
        for (int i = 0; i < R; ++i)
        {
            const bool isP = lane * R + i == p;
            const float c = isP ? 0.0f : S[i][k] * rc;
#pragma unroll
            for (int j = 0; j < N; ++j)
                if (j != k) S[i][j] = fmaf(-c, u[j], S[i][j]);
            S[i][k] = isP ? 1.0f : -c;
            if (isP)
            {
                sigma[i] = rc;
                used[i] = true;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }

    // Result: row r (true values sigma*S) goes to row pos; column j goes to pos^-1(j).
#pragma unroll
    for (int i = 0; i < R; ++i)
        if (rowLive[i]) s_inv[g][pos[i]] = lane * R + i;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
    for (int j = 0; j < N; ++j)
        if (j < n)
        {
            const int col = live ? s_inv[g][j] : 0;
#pragma unroll
            for (int i = 0; i < R; ++i)
                if (rowLive[i]) C[static_cast<int64_t>(col) * ldc + pos[i]] = sigma[i] * S[i][j];
        }
    if (live && lane == 0) info[mat] = zero;
}

// Inverse from LU factors and pivots: X = U^-1 L^-1 P, as two substitution sweeps over the
// rows of P*I with the factor rows held by their lanes.
// Reference: src/sgetri_batched/sgetri_batched_host.cpp:41-83 (RunInverseBatch) and the
// LAPACK-order small inverse src/utils/kernel/real_lu_small_inverse.hpp:220-242.
template <int NL>
__global__ __launch_bounds__(256) void getri_small_kernel(int n, const float* const* __restrict__ Aarr, int lda,
                                                          const int* __restrict__ piv, float* const* __restrict__ Carr,
                                                          int ldc, int* __restrict__ info, int batch)
{
    constexpr int GROUPS = 256 / NL;
    __shared__ __attribute__((aligned(16))) float s_row[GROUPS][NL];
    __shared__ int s_perm[GROUPS][NL];
    const int tid = threadIdx.x;
    const int g = tid / NL, lane = tid % NL;
    const int mat = blockIdx.x * GROUPS + g;
    const bool live = mat < batch;
    const bool rowLive = live && lane < n;
    const float* A = live ? Aarr[mat] : nullptr;
    float* C = live ? Carr[mat] : nullptr;
    float LU[NL];
#pragma unroll
    for (int j = 0; j < NL; ++j) LU[j] = (rowLive && j < n) ? A[static_cast<int64_t>(j) * lda + lane] : 0.0f;
    // Permutation from the interchange sequence (one lane per matrix, n <= 64 steps).
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
    // B = P*I: row r is e_{perm[r]}.
    float X[NL];
    const int pr = rowLive ? s_perm[g][lane] : -1;
#pragma unroll
    for (int j = 0; j < NL; ++j) X[j] = (j == pr) ? 1.0f : 0.0f;
    // Forward substitution L Y = P I (unit lower).
#pragma unroll
    for (int q = 0; q < NL; ++q)
    {
        if (q >= n) continue;
        if (lane == q)
        {
#pragma unroll
            for (int j = 0; j < NL; ++j) s_row[g][j] = X[j];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane > q && rowLive)
        {
            const float l = LU[q];
#pragma unroll
            for (int j = 0; j < NL; ++j) X[j] = fmaf(-l, s_row[g][j], X[j]);
        }
        __builtin_amdgcn_wave_barrier();
    }
    // Backward substitution U X = Y.
#pragma unroll
    for (int q = NL - 1; q >= 0; --q)
    {
        if (q >= n) continue;
        if (lane == q)
        {
            const float d = LU[q];
            if (d == 0.0f)
            {
#pragma unroll
                for (int j = 0; j < NL; ++j) s_row[g][j] = X[j];
            }
            else
            {
#pragma unroll
                for (int j = 0; j < NL; ++j)
                {
                    X[j] = X[j] / d;
                    s_row[g][j] = X[j];
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane < q && rowLive)
        {
            const float uq = LU[q];
#pragma unroll
            for (int j = 0; j < NL; ++j) X[j] = fmaf(-uq, s_row[g][j], X[j]);
        }
        __builtin_amdgcn_wave_barrier();
    }
    // info: first zero on U's diagonal (lane r holds U(r, r) in register r).
    float diag = 0.0f;
#pragma unroll
    for (int j = 0; j < NL; ++j)
        if (j == lane) diag = LU[j];
    const unsigned long long zmask = __ballot(rowLive && diag == 0.0f);
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
hipError_t LaunchMatinv(int n, const float* const* Aarr, const float* Abase, int lda, int64_t strideA,
                        float* const* Carr, float* Cbase, int ldc, int64_t strideC, int* ipiv, int* info, int batch,
                        hipStream_t stream)
{
    constexpr int GROUPS = 256 / NL;
    const int blocks = (batch + GROUPS - 1) / GROUPS;
    hipLaunchKernelGGL(matinv_gj_kernel<NL>, dim3(blocks), dim3(256), 0, stream, n, Aarr, Abase, lda, strideA, Carr,
                       Cbase, ldc, strideC, ipiv, info, batch);
    return hipGetLastError();
}

template <int NL, int R>
hipError_t LaunchMatinvRows(int n, const float* const* Aarr, const float* Abase, int lda, int64_t strideA,
                            float* const* Carr, float* Cbase, int ldc, int64_t strideC, int* ipiv, int* info, int batch,
                            hipStream_t stream)
{
    constexpr int GROUPS = 256 / NL;
    const int blocks = (batch + GROUPS - 1) / GROUPS;
    hipLaunchKernelGGL((matinv_gj_rows_kernel<NL, R>), dim3(blocks), dim3(256), 0, stream, n, Aarr, Abase, lda, strideA,
                       Carr, Cbase, ldc, strideC, ipiv, info, batch);
    return hipGetLastError();
}

// Rows per lane for 17..32 (OPS_SOLVER_MATINV_ROWS = 1, 2 or 4; default 2).
int MatinvRows()
{
    static const int r = []
    {
        const char* e = std::getenv("OPS_SOLVER_MATINV_ROWS");
        const int v = e != nullptr ? std::atoi(e) : 2;
        return v == 1 || v == 4 ? v : 2;
    }();
    return r;
}

template <int NL>
hipError_t LaunchGetri(int n, const float* const* Aarr, int lda, const int* piv, float* const* Carr, int ldc,
                       int* info, int batch, hipStream_t stream)
{
    constexpr int GROUPS = 256 / NL;
    const int blocks = (batch + GROUPS - 1) / GROUPS;
    hipLaunchKernelGGL(getri_small_kernel<NL>, dim3(blocks), dim3(256), 0, stream, n, Aarr, lda, piv, Carr, ldc, info,
                       batch);
    return hipGetLastError();
}

}  // namespace

// Pointer-array or strided batched direct inverse, n <= 64. ipiv (nullable) receives
// n*batch pivots. Exactly one of (Aarr, Abase) and one of (Carr, Cbase) is non-null.
hipError_t MatinvBatched(int n, const float* const* Aarr, const float* Abase, int lda, int64_t strideA,
                         float* const* Carr, float* Cbase, int ldc, int64_t strideC, int* ipiv, int* info, int batch,
                         hipStream_t stream)
{
    if (batch <= 0) return hipSuccess;
    if (n <= 8) return LaunchMatinv<8>(n, Aarr, Abase, lda, strideA, Carr, Cbase, ldc, strideC, ipiv, info, batch, stream);
    if (n <= 16)
        return LaunchMatinv<16>(n, Aarr, Abase, lda, strideA, Carr, Cbase, ldc, strideC, ipiv, info, batch, stream);
    if (n <= 32)
    {
        const int rows = MatinvRows();
        if (rows == 4)
            return LaunchMatinvRows<8, 4>(n, Aarr, Abase, lda, strideA, Carr, Cbase, ldc, strideC, ipiv, info, batch,
                                          stream);
        if (rows == 2)
            return LaunchMatinvRows<16, 2>(n, Aarr, Abase, lda, strideA, Carr, Cbase, ldc, strideC, ipiv, info, batch,
                                           stream);
        return LaunchMatinv<32>(n, Aarr, Abase, lda, strideA, Carr, Cbase, ldc, strideC, ipiv, info, batch, stream);
    }
    if (n <= 64)
        return LaunchMatinv<64>(n, Aarr, Abase, lda, strideA, Carr, Cbase, ldc, strideC, ipiv, info, batch, stream);
    return hipErrorInvalidValue;
}

hipError_t GetriSmallBatched(int n, const float* const* Aarr, int lda, const int* piv, float* const* Carr, int ldc,
                             int* info, int batch, hipStream_t stream)
{
    if (batch <= 0) return hipSuccess;
    if (n <= 8) return LaunchGetri<8>(n, Aarr, lda, piv, Carr, ldc, info, batch, stream);
    if (n <= 16) return LaunchGetri<16>(n, Aarr, lda, piv, Carr, ldc, info, batch, stream);
    if (n <= 32) return LaunchGetri<32>(n, Aarr, lda, piv, Carr, ldc, info, batch, stream);
    if (n <= 64) return LaunchGetri<64>(n, Aarr, lda, piv, Carr, ldc, info, batch, stream);
    return hipErrorInvalidValue;
}

}  // namespace ops
