// Randomized butterfly transform + LU without pivoting: the factorization behind the
// mixed-precision gesv family (EXTENSION; the reference's LU always pivots).
//
// A_r = U^T A V with U, V depth-2 recursive butterflies (Parker's RBT: B = 1/sqrt(2) *
// [[R0, R1], [R0, -R1]], R diagonal with entries exp(r/10), r uniform in [-1/2, 1/2]).
// With high probability A_r can be factored without pivoting, so the panel has no
// column-by-column argmax: every step is a small diagonal-block LU plus parallel
// triangular solves and an MFMA GEMM. The solution of A x = b is x = V (L U)^-1 U^T b; the
// GMRES refinement (gesv_ir.hip) uses this as its preconditioner and certifies the backward
// error in FP64 against the original A.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "lu_kernels.h"

namespace ops
{
namespace
{
constexpr int IB = 64;   // inner (diagonal) block
constexpr int OB = 512;  // outer block (trailing GEMM depth)

__device__ inline float ButterflyValue(unsigned long long seed, unsigned long long idx)
{
    unsigned long long z = seed + idx * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double r = static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    return static_cast<float>(exp(r / 10.0));
}

// d[0:n) = level-1 diagonal (R0 on the first half, R1 on the second), d[n:2n) = level-2
// (per half of the matrix, again R0/R1 halves), for one side.
__global__ void butterfly_diag_kernel(int n, unsigned long long seed, float* __restrict__ d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * n) d[i] = ButterflyValue(seed, static_cast<unsigned long long>(i));
}

// Applies W^T (left, rows) or W (right, columns) of a depth-2 butterfly, fused over the two
// levels. For the left side, output rows {i, i+q, i+2q, i+3q} (q = n/4) depend on the same
// four input rows; for the right side the same holds for columns. src -> dst (may alias).
// Level 1 (size n, pairs i, i+h):   top' = r0 (top + bot)/sqrt2, bot' = r1 (top - bot)/sqrt2
// Level 2 (size n/2 on each half): the same with the level-2 diagonal.
// For the right side (A V): cols' = (cols_top s0 + cols_bot s0)... see RightCombine.
__device__ inline void LeftCombine(const float* d1, const float* d2, int n, int i, float x0, float x1, float x2,
                                   float x3, float* y)
{
    // x0 = row i, x1 = row i+q, x2 = row i+h, x3 = row i+h+q ; h = n/2, q = n/4
    const int h = n / 2, q = n / 4;
    const float s = 0.70710678118654752f;
    // level 1: rows (i, i+h) and (i+q, i+q+h)
    const float a0 = d1[i] * (x0 + x2) * s;          // row i
    const float a1 = d1[i + q] * (x1 + x3) * s;      // row i+q
    const float a2 = d1[i + h] * (x0 - x2) * s;      // row i+h
    const float a3 = d1[i + h + q] * (x1 - x3) * s;  // row i+h+q
    // level 2 on each half (size h, pairs (i, i+q) within the top half and (i+h, i+h+q))
    y[0] = d2[i] * (a0 + a1) * s;
    y[1] = d2[i + q] * (a0 - a1) * s;
    y[2] = d2[i + h] * (a2 + a3) * s;
    y[3] = d2[i + h + q] * (a2 - a3) * s;
}

// A V: column j of the result combines columns of A with V's entries. For V = B (one
// level): (A B)[:, j] = s0_j (A[:, j] + A[:, j+h]) / sqrt2 for j < h, and
// (A B)[:, j+h] = s1_j (A[:, j] - A[:, j+h]) / sqrt2. Two levels: V = V1 * V2 (V1 full,
// V2 block diagonal), A V = (A V1) V2.
__device__ inline void RightCombine(const float* d1, const float* d2, int n, int j, float x0, float x1, float x2,
                                    float x3, float* y)
{
    const int h = n / 2, q = n / 4;
    const float s = 0.70710678118654752f;
    const float a0 = d1[j] * (x0 + x2) * s;
    const float a1 = d1[j + q] * (x1 + x3) * s;
    const float a2 = d1[j + h] * (x0 - x2) * s;
    const float a3 = d1[j + h + q] * (x1 - x3) * s;
    y[0] = d2[j] * (a0 + a1) * s;
    y[1] = d2[j + q] * (a0 - a1) * s;
    y[2] = d2[j + h] * (a2 + a3) * s;
    y[3] = d2[j + h + q] * (a2 - a3) * s;
}

// dst = U^T src (left side): one thread per (row i < n/4, column c).
__global__ void rbt_left_kernel(int n, const float* __restrict__ src, int lds, float* __restrict__ dst, int ldd,
                                const float* __restrict__ du)
{
    const int q = n / 4;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= q) return;
    const float* s = src + static_cast<size_t>(c) * lds;
    float y[4];
    LeftCombine(du, du + n, n, i, s[i], s[i + q], s[i + 2 * q], s[i + 3 * q], y);
    float* d = dst + static_cast<size_t>(c) * ldd;
    d[i] = y[0];
    d[i + q] = y[1];
    d[i + 2 * q] = y[2];
    d[i + 3 * q] = y[3];
}

// dst = src V (right side): one thread per (row r, column j < n/4).
__global__ void rbt_right_kernel(int n, int rows, const float* __restrict__ src, int lds, float* __restrict__ dst,
                                 int ldd, const float* __restrict__ dv)
{
    const int q = n / 4;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (r >= rows) return;
    auto at = [&](int col) { return src[static_cast<size_t>(col) * lds + r]; };
    float y[4];
    RightCombine(dv, dv + n, n, j, at(j), at(j + q), at(j + 2 * q), at(j + 3 * q), y);
    dst[static_cast<size_t>(j) * ldd + r] = y[0];
    dst[static_cast<size_t>(j + q) * ldd + r] = y[1];
    dst[static_cast<size_t>(j + 2 * q) * ldd + r] = y[2];
    dst[static_cast<size_t>(j + 3 * q) * ldd + r] = y[3];
}

// Vector versions for the solve: y = U^T x (left) and y = V x (V applied to a vector:
// V = V1 V2, so V x = V1 (V2 x)); both in FP64.
__global__ void rbt_vec_left_kernel(int n, const double* __restrict__ x, double* __restrict__ y,
                                    const float* __restrict__ du)
{
    const int q = n / 4, h = n / 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q) return;
    const double s = 0.70710678118654752440;
    const float* d1 = du;
    const float* d2 = du + n;
    const double x0 = x[i], x1 = x[i + q], x2 = x[i + h], x3 = x[i + h + q];
    const double a0 = d1[i] * (x0 + x2) * s, a1 = d1[i + q] * (x1 + x3) * s;
    const double a2 = d1[i + h] * (x0 - x2) * s, a3 = d1[i + h + q] * (x1 - x3) * s;
    y[i] = d2[i] * (a0 + a1) * s;
    y[i + q] = d2[i + q] * (a0 - a1) * s;
    y[i + h] = d2[i + h] * (a2 + a3) * s;
    y[i + h + q] = d2[i + h + q] * (a2 - a3) * s;
}

// y = V x. V1 (level 1, size n): (V1 z)[j] for j < h: (s0_j z_j + s1_j z_{j+h}) / sqrt2 ...
// V1 = 1/sqrt2 [[S0, S1],[S0, -S1]] so (V1 z)_top = (S0 z_top + S1 z_bot)/sqrt2,
// (V1 z)_bot = (S0 z_top - S1 z_bot)/sqrt2; V2 likewise on each half with the level-2
// diagonal. V x = V1 (V2 x).
__global__ void rbt_vec_right_kernel(int n, const double* __restrict__ x, double* __restrict__ y,
                                     const float* __restrict__ dv)
{
    const int q = n / 4, h = n / 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q) return;
    const double s = 0.70710678118654752440;
    const float* d1 = dv;
    const float* d2 = dv + n;
    // V2 on halves: top half pairs (i, i+q), bottom half pairs (i+h, i+h+q)
    const double x0 = x[i], x1 = x[i + q], x2 = x[i + h], x3 = x[i + h + q];
    const double b0 = (d2[i] * x0 + d2[i + q] * x1) * s;
    const double b1 = (d2[i] * x0 - d2[i + q] * x1) * s;
    const double b2 = (d2[i + h] * x2 + d2[i + h + q] * x3) * s;
    const double b3 = (d2[i + h] * x2 - d2[i + h + q] * x3) * s;
    // V1 on the full vector: pairs (i, i+h) and (i+q, i+q+h)
    y[i] = (d1[i] * b0 + d1[i + h] * b2) * s;
    y[i + h] = (d1[i] * b0 - d1[i + h] * b2) * s;
    y[i + q] = (d1[i + q] * b1 + d1[i + h + q] * b3) * s;
    y[i + h + q] = (d1[i + q] * b1 - d1[i + h + q] * b3) * s;
}

// LU without pivoting of the IB x IB diagonal block at (k, k), in LDS, one workgroup.
// info: first zero pivot (1-based, global index), the column is skipped like getrf.
__global__ __launch_bounds__(256) void diag_lu_kernel(float* __restrict__ A, int lda, int k, int kb,
                                                      int* __restrict__ info)
{
    __shared__ float s[IB][IB + 1];
    float* blk = A + static_cast<size_t>(k) * lda + k;
    for (int e = threadIdx.x; e < kb * kb; e += 256)
    {
        const int r = e % kb, c = e / kb;
        s[r][c] = blk[static_cast<size_t>(c) * lda + r];
    }
    __syncthreads();
    int zero = 0;
    for (int j = 0; j < kb; ++j)
    {
        const float d = s[j][j];
        if (d == 0.0f)
        {
            if (zero == 0) zero = k + j + 1;
            __syncthreads();
            continue;
        }
        const float r = 1.0f / d;
        const bool recip = fabsf(d) >= FLT_MIN;
        for (int i = j + 1 + threadIdx.x; i < kb; i += 256) s[i][j] = recip ? s[i][j] * r : s[i][j] / d;
        __syncthreads();
        const int rem = kb - j - 1;
        for (int e = threadIdx.x; e < rem * rem; e += 256)
        {
            const int i = j + 1 + e % rem, c = j + 1 + e / rem;
            s[i][c] = fmaf(-s[i][j], s[j][c], s[i][c]);
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < kb * kb; e += 256)
    {
        const int r = e % kb, c = e / kb;
        blk[static_cast<size_t>(c) * lda + r] = s[r][c];
    }
    if (threadIdx.x == 0 && zero != 0 && *info == 0) *info = zero;
}

// Right-side solve X U = B for the rows below the diagonal block: B (rows x kb) at
// A[k+kb.., k..], U = upper triangle of the diagonal block. One thread per row; the row
// lives in registers, U in LDS.
__global__ __launch_bounds__(256) void trsm_right_upper_kernel(float* __restrict__ A, int lda, int k, int kb,
                                                               int rowsBelow)
{
    __shared__ float u[IB][IB + 1];
    const float* blk = A + static_cast<size_t>(k) * lda + k;
    for (int e = threadIdx.x; e < kb * kb; e += 256)
    {
        const int r = e % kb, c = e / kb;
        u[r][c] = blk[static_cast<size_t>(c) * lda + r];
    }
    __syncthreads();
    const int row = blockIdx.x * 256 + threadIdx.x;
    if (row >= rowsBelow) return;
    float* b = A + static_cast<size_t>(k) * lda + k + kb + row;
    float x[IB];
#pragma unroll
    for (int c = 0; c < IB; ++c) x[c] = c < kb ? b[static_cast<size_t>(c) * lda] : 0.0f;
#pragma unroll
    for (int i = 0; i < IB; ++i)
    {
        if (i < kb)
        {
            const float d = u[i][i];
            x[i] = (d != 0.0f) ? (fabsf(d) >= FLT_MIN ? x[i] * (1.0f / d) : x[i] / d) : x[i];
#pragma unroll
            for (int j = i + 1; j < IB; ++j) x[j] = fmaf(-x[i], u[i][j], x[j]);
        }
    }
#pragma unroll
    for (int c = 0; c < IB; ++c)
        if (c < kb) b[static_cast<size_t>(c) * lda] = x[c];
}

#define GRET(e)                          \
    do                                   \
    {                                    \
        hipError_t _e = (e);             \
        if (_e != hipSuccess) return _e; \
    } while (0)

}  // namespace

// Random butterfly diagonals for U (du, 2n floats) and V (dv, 2n floats).
hipError_t RbtMakeDiagonals(int n, unsigned long long seed, float* du, float* dv, hipStream_t s)
{
    hipLaunchKernelGGL(butterfly_diag_kernel, dim3((2 * n + 255) / 256), dim3(256), 0, s, n, seed, du);
    hipLaunchKernelGGL(butterfly_diag_kernel, dim3((2 * n + 255) / 256), dim3(256), 0, s, n,
                       seed ^ 0xA5A5A5A5DEADBEEFull, dv);
    return hipGetLastError();
}

// dst = U^T A V (n % 4 == 0); tmp may equal dst.
hipError_t RbtTransform(int n, const float* A, int lda, float* dst, int ldd, const float* du, const float* dv,
                        hipStream_t s)
{
    const int q = n / 4;
    hipLaunchKernelGGL(rbt_left_kernel, dim3((q + 255) / 256, n), dim3(256), 0, s, n, A, lda, dst, ldd, du);
    hipLaunchKernelGGL(rbt_right_kernel, dim3((n + 255) / 256, q), dim3(256), 0, s, n, n, dst, ldd, dst, ldd, dv);
    return hipGetLastError();
}

hipError_t RbtApplyVecLeft(int n, const double* x, double* y, const float* du, hipStream_t s)
{
    hipLaunchKernelGGL(rbt_vec_left_kernel, dim3((n / 4 + 255) / 256), dim3(256), 0, s, n, x, y, du);
    return hipGetLastError();
}

int GenpOuterBlock() { return OB; }

hipError_t RbtApplyVecRight(int n, const double* x, double* y, const float* dv, hipStream_t s)
{
    hipLaunchKernelGGL(rbt_vec_right_kernel, dim3((n / 4 + 255) / 256), dim3(256), 0, s, n, x, y, dv);
    return hipGetLastError();
}

// LU without pivoting of an n x n matrix in place. Trailing updates of depth OB run with
// the requested MFMA input precision (on packed operands when `pack` holds
// GemmPackedWorkspaceBytes(n, n, OB)); the in-panel updates run in FP32.
hipError_t GetrfNoPivot(int n, float* a, int lda, int* info, Precision trailing, void* pack, hipStream_t s)
{
    FillInfo(info, 1, 0, s);
    for (int J = 0; J < n; J += OB)
    {
        const int jb = std::min(OB, n - J);
        const int Jend = J + jb;
        for (int k = J; k < Jend; k += IB)
        {
            const int kb = std::min(IB, Jend - k);
            hipLaunchKernelGGL(diag_lu_kernel, dim3(1), dim3(256), 0, s, a, lda, k, kb, info);
            const int below = n - k - kb;
            if (below > 0)
                hipLaunchKernelGGL(trsm_right_upper_kernel, dim3((below + 255) / 256), dim3(256), 0, s, a, lda, k,
                                   kb, below);
            const int right = Jend - k - kb;
            if (right > 0)
            {
                GRET(TrsmBlock(kb, right, a + static_cast<size_t>(k) * lda + k, lda, true, true, false,
                               a + static_cast<size_t>(k + kb) * lda + k, lda, s));
                if (below > 0)
                    GRET(GemmSub(below, right, kb, a + static_cast<size_t>(k) * lda + k + kb, lda, 0,
                                 a + static_cast<size_t>(k + kb) * lda + k, lda,
                                 a + static_cast<size_t>(k + kb) * lda + k + kb, lda, Precision::FP32, s));
            }
            GRET(hipGetLastError());
        }
        const int right = n - Jend;
        if (right > 0)
        {
            GRET(TrsmBlock(jb, right, a + static_cast<size_t>(J) * lda + J, lda, true, true, false,
                           a + static_cast<size_t>(Jend) * lda + J, lda, s));
            if (pack != nullptr)
                GRET(GemmSubPacked(right, right, jb, a + static_cast<size_t>(J) * lda + Jend, lda,
                                   a + static_cast<size_t>(Jend) * lda + J, lda,
                                   a + static_cast<size_t>(Jend) * lda + Jend, lda, trailing, pack, s));
            else
                GRET(GemmSub(right, right, jb, a + static_cast<size_t>(J) * lda + Jend, lda, 0,
                             a + static_cast<size_t>(Jend) * lda + J, lda, a + static_cast<size_t>(Jend) * lda + Jend,
                             lda, trailing, s));
        }
    }
    return hipGetLastError();
}

}  // namespace ops
