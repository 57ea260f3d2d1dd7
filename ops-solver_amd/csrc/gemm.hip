// MFMA GEMM update C -= op(A) * B and the LDS-resident diagonal-block triangular solve.
//
// The GEMM is the trailing-matrix update of the blocked LU (reference
// src/sgetrf/real_lu_blocked_kernel.cpp:1119-1158, real_lu_factor_product_kernel: FP32 with
// HF32 off, product rounded once before it is added). Here it runs on gfx950 matrix cores:
//   FP32: v_mfma_f32_32x32x2_f32, BF16: v_mfma_f32_32x32x16_bf16, FP16: v_mfma_f32_32x32x16_f16,
// always with FP32 accumulation. A workgroup of 4 wave64s owns a 128x128 tile of C; each
// wave owns a 64x64 quadrant (2x2 blocks of 32x32 accumulators, 64 accumulator registers).
// Operand tiles are staged through LDS; the FP16 variant scales each staged operand by a
// power of two chosen from the tile's largest magnitude (exact, undone after the sum).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>

#include "lu_kernels.h"

namespace ops
{
namespace
{
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int TM = 128;
constexpr int TN = 128;

// Maps the linear workgroup id onto (tile row, tile column). Consecutive workgroups are
// dealt round-robin to the 8 XCDs; regroup them so that each XCD gets a contiguous run of
// tiles (column-major tile order), which then share operand panels in that XCD's L2.
__device__ inline void TileOf(int tilesM, int tilesN, int* tm, int* tn)
{
    int total = tilesM * tilesN;
    int id = blockIdx.x;
    if (total % 8 == 0)
    {
        int per = total / 8;
        id = (id % 8) * per + id / 8;
    }
    *tm = id % tilesM;
    *tn = id / tilesM;
}

__device__ inline float OpA(const float* A, int lda, int transA, int r, int k)
{
    return transA ? A[static_cast<size_t>(r) * lda + k] : A[static_cast<size_t>(k) * lda + r];
}

// FP32 update: 128 x 128 tile per workgroup, k chunks of 16 staged through LDS with the
// next chunk prefetched into registers while the current one runs on the matrix cores
// (one barrier per chunk). The product is formed as C^T (src0 = B, src1 = op(A)) so that
// each accumulator register holds 32 consecutive rows of one column of C and the
// read-modify-write of C is coalesced.
template <int TRANSA>
__global__ __launch_bounds__(256) void gemm_sub_f32_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                           const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                           int ldc, int tilesM, int tilesN)
{
    constexpr int KC = 16;
    __shared__ float As[2][KC][TM + 4];
    __shared__ float Bs[2][KC][TN + 4];
    int tm, tn;
    TileOf(tilesM, tilesN, &tm, &tn);
    const int row0 = tm * TM, col0 = tn * TN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

    // Per chunk each thread moves 8 elements of op(A) and 8 of B.
    float ra[8], rb[8];
    auto gload = [&](int k0)
    {
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            int r, k;
            if (!TRANSA)
            {
                r = tid & 127;
                k = (tid >> 7) * 8 + q;
            }
            else
            {
                r = tid >> 1;
                k = (tid & 1) * 8 + q;
            }
            float v = 0.0f;
            if (row0 + r < M && k0 + k < K)
                v = TRANSA ? A[static_cast<size_t>(row0 + r) * lda + k0 + k] : A[static_cast<size_t>(k0 + k) * lda + row0 + r];
            ra[q] = v;
            const int j = tid >> 1, kb = (tid & 1) * 8 + q;
            float u = 0.0f;
            if (col0 + j < N && k0 + kb < K) u = B[static_cast<size_t>(col0 + j) * ldb + k0 + kb];
            rb[q] = u;
        }
    };
    auto lstore = [&](int buf)
    {
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            if (!TRANSA)
                As[buf][(tid >> 7) * 8 + q][tid & 127] = ra[q];
            else
                As[buf][(tid & 1) * 8 + q][tid >> 1] = ra[q];
            Bs[buf][(tid & 1) * 8 + q][tid >> 1] = rb[q];
        }
    };
    const int chunks = (K + KC - 1) / KC;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int c = 0; c < chunks; ++c)
    {
        const int buf = c & 1;
        if (c + 1 < chunks) gload((c + 1) * KC);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 2)
        {
            const int k = kk + (lane >> 5);
            const float a0 = As[buf][k][wm * 64 + (lane & 31)];
            const float a1 = As[buf][k][wm * 64 + 32 + (lane & 31)];
            const float b0 = Bs[buf][k][wn * 64 + (lane & 31)];
            const float b1 = Bs[buf][k][wn * 64 + 32 + (lane & 31)];
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(b0, a0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(b1, a0, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(b0, a1, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(b1, a1, acc[1][1], 0, 0, 0);
        }
        if (c + 1 < chunks) lstore(buf ^ 1);
        __syncthreads();
    }
    // acc[bi][bj] register v of lane l: product row (v%4) + 8(v/4) + 4(l/32) -> column
    // col0 + wn*64 + bj*32 + that of C; product column l%32 -> row row0 + wm*64 + bi*32 + l%32.
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
    {
        const int i = row0 + wm * 64 + bi * 32 + (lane & 31);
        if (i >= M) continue;
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
        {
#pragma unroll
            for (int v = 0; v < 16; ++v)
            {
                const int j = col0 + wn * 64 + bj * 32 + (v & 3) + 8 * (v >> 2) + 4 * (lane >> 5);
                if (j < N) C[static_cast<size_t>(j) * ldc + i] -= acc[bi][bj][v];
            }
        }
    }
}

// FP32 update on 64 x 64 tiles (no transposes) for the skinny products of the panel
// recursion, where N and K are 32 .. 256 and 128 x 128 tiles would leave most of the chip
// idle: four waves hold one 32 x 32 accumulator each, k chunks of 16 double-buffered
// through LDS as in gemm_sub_f32_kernel.
__global__ __launch_bounds__(256) void gemm_sub_f32_64_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                              const float* __restrict__ B, int ldb,
                                                              float* __restrict__ C, int ldc, int tilesM, int tilesN)
{
    constexpr int KC = 16, T64 = 64;
    __shared__ float As[2][KC][T64 + 4];
    __shared__ float Bs[2][KC][T64 + 4];
    int tm, tn;
    TileOf(tilesM, tilesN, &tm, &tn);
    const int row0 = tm * T64, col0 = tn * T64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    // Per chunk each thread moves 4 elements of A (column k, row tid % 64) and 4 of B.
    float ra[4], rb[4];
    auto gload = [&](int k0)
    {
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            const int r = tid & 63, k = (tid >> 6) * 4 + q;
            ra[q] = (row0 + r < M && k0 + k < K) ? A[static_cast<size_t>(k0 + k) * lda + row0 + r] : 0.0f;
            const int j = tid >> 2, kb = (tid & 3) * 4 + q;
            rb[q] = (col0 + j < N && k0 + kb < K) ? B[static_cast<size_t>(col0 + j) * ldb + k0 + kb] : 0.0f;
        }
    };
    auto lstore = [&](int buf)
    {
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            As[buf][(tid >> 6) * 4 + q][tid & 63] = ra[q];
            Bs[buf][(tid & 3) * 4 + q][tid >> 2] = rb[q];
        }
    };
    const int chunks = (K + KC - 1) / KC;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int c = 0; c < chunks; ++c)
    {
        const int buf = c & 1;
        if (c + 1 < chunks) gload((c + 1) * KC);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 2)
        {
            const int k = kk + (lane >> 5);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(Bs[buf][k][wn * 32 + (lane & 31)], As[buf][k][wm * 32 + (lane & 31)],
                                                       acc, 0, 0, 0);
        }
        if (c + 1 < chunks) lstore(buf ^ 1);
        __syncthreads();
    }
    const int i = row0 + wm * 32 + (lane & 31);
    if (i >= M) return;
#pragma unroll
    for (int v = 0; v < 16; ++v)
    {
        const int j = col0 + wn * 32 + (v & 3) + 8 * (v >> 2) + 4 * (lane >> 5);
        if (j < N) C[static_cast<size_t>(j) * ldc + i] -= acc[v];
    }
}

template <typename T>
__device__ inline T ToHalf(float v);
template <>
__device__ inline _Float16 ToHalf<_Float16>(float v)
{
    return static_cast<_Float16>(v);
}
template <>
__device__ inline __bf16 ToHalf<__bf16>(float v)
{
    return static_cast<__bf16>(v);
}

__device__ inline float BlockMax(float v, float* red)
{
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    float r = red[0];
    for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) r = fmaxf(r, red[w]);
    return r;
}

// Power-of-two factor that brings the largest magnitude into [2^14, 2^15) (FP16 keeps
// 2^15 < 65504); clamped so that the factor and its inverse stay finite normal floats.
__device__ inline float ScaleFor(float maxAbs)
{
    if (!(maxAbs > 0.0f) || !(maxAbs <= FLT_MAX)) return 1.0f;
    int e = ilogbf(maxAbs);
    int s = 14 - e;
    s = s > 100 ? 100 : (s < -100 ? -100 : s);
    return ldexpf(1.0f, s);
}

template <typename HT, int TRANSA, bool SCALE>
__global__ __launch_bounds__(256) void gemm_sub_half_kernel(int M, int N, int K, const float* __restrict__ A,
                                                            int lda, const float* __restrict__ B, int ldb,
                                                            float* __restrict__ C, int ldc, int tilesM, int tilesN)
{
    typedef HT hx8 __attribute__((ext_vector_type(8)));
    __shared__ __attribute__((aligned(16))) HT As[TM][32 + 8];
    __shared__ __attribute__((aligned(16))) HT Bs[TN][32 + 8];
    __shared__ float red[8];
    int tm, tn;
    TileOf(tilesM, tilesN, &tm, &tn);
    const int row0 = tm * TM, col0 = tn * TN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    float sa = 1.0f, sb = 1.0f;
    if (SCALE)
    {
        float ma = 0.0f, mb = 0.0f;
        for (int e = tid; e < TM * K; e += 256)
        {
            int r = TRANSA ? e / K : e % TM, k = TRANSA ? e % K : e / TM;
            if (row0 + r < M) ma = fmaxf(ma, fabsf(OpA(A, lda, TRANSA, row0 + r, k)));
        }
        for (int e = tid; e < TN * K; e += 256)
        {
            int j = e / K, k = e % K;
            if (col0 + j < N) mb = fmaxf(mb, fabsf(B[static_cast<size_t>(col0 + j) * ldb + k]));
        }
        sa = ScaleFor(BlockMax(ma, red));
        sb = ScaleFor(BlockMax(mb, red));
    }

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

    for (int k0 = 0; k0 < K; k0 += 32)
    {
        if (!TRANSA)
        {
            const int r = tid & 127, kq = (tid >> 7) * 16;
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
            {
                int k = kq + kk;
                float v = 0.0f;
                if (row0 + r < M && k0 + k < K) v = A[static_cast<size_t>(k0 + k) * lda + row0 + r];
                As[r][k] = ToHalf<HT>(v * sa);
            }
        }
        else
        {
            const int r = tid >> 1, kh = (tid & 1) * 16;
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
            {
                int k = kh + kk;
                float v = 0.0f;
                if (row0 + r < M && k0 + k < K) v = A[static_cast<size_t>(row0 + r) * lda + k0 + k];
                As[r][k] = ToHalf<HT>(v * sa);
            }
        }
        {
            const int j = tid >> 1, kh = (tid & 1) * 16;
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
            {
                int k = kh + kk;
                float v = 0.0f;
                if (col0 + j < N && k0 + k < K) v = B[static_cast<size_t>(col0 + j) * ldb + k0 + k];
                Bs[j][k] = ToHalf<HT>(v * sb);
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 32; ks += 16)
        {
            const int kq = ks + (lane >> 5) * 8;
            const hx8 a0 = *reinterpret_cast<const hx8*>(&As[wm * 64 + (lane & 31)][kq]);
            const hx8 a1 = *reinterpret_cast<const hx8*>(&As[wm * 64 + 32 + (lane & 31)][kq]);
            const hx8 b0 = *reinterpret_cast<const hx8*>(&Bs[wn * 64 + (lane & 31)][kq]);
            const hx8 b1 = *reinterpret_cast<const hx8*>(&Bs[wn * 64 + 32 + (lane & 31)][kq]);
            if constexpr (sizeof(HT) == 2 && __is_same(HT, _Float16))
            {
                acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0, b0, acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0, b1, acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, b0, acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, b1, acc[1][1], 0, 0, 0);
            }
            else
            {
                acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc[1][1], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const float ia = 1.0f / sa, ib = 1.0f / sb;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
        {
            const int j = col0 + wn * 64 + bj * 32 + (lane & 31);
            if (j >= N) continue;
            float* cj = C + static_cast<size_t>(j) * ldc;
#pragma unroll
            for (int v = 0; v < 16; ++v)
            {
                const int i = row0 + wm * 64 + bi * 32 + (v & 3) + 8 * (v >> 2) + 4 * (lane >> 5);
                if (i < M) cj[i] -= (acc[bi][bj][v] * ia) * ib;
            }
        }
}

// Split BF16 form: every FP32 operand x is held as hi = bf16(x), lo = bf16(x - hi) and the
// product is accumulated as hi*hi + hi*lo + lo*hi on bf16 MFMA with FP32 accumulation
// (about 16 significant bits per product instead of 8). Three MFMAs per block pair; the
// trailing update stays memory-bound, so the extra matrix-core work is mostly hidden.
template <int TRANSA>
__global__ __launch_bounds__(256) void gemm_sub_bf16x3_kernel(int M, int N, int K, const float* __restrict__ A,
                                                              int lda, const float* __restrict__ B, int ldb,
                                                              float* __restrict__ C, int ldc, int tilesM, int tilesN)
{
    typedef __bf16 hx8 __attribute__((ext_vector_type(8)));
    __shared__ __attribute__((aligned(16))) __bf16 Ah[TM][32 + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Al[TM][32 + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[TN][32 + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Bl[TN][32 + 8];
    int tm, tn;
    TileOf(tilesM, tilesN, &tm, &tn);
    const int row0 = tm * TM, col0 = tn * TN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;
    auto split = [](float v, __bf16& h, __bf16& l)
    {
        h = static_cast<__bf16>(v);
        l = static_cast<__bf16>(v - static_cast<float>(h));
    };
    for (int k0 = 0; k0 < K; k0 += 32)
    {
        if (!TRANSA)
        {
            const int r = tid & 127, kq = (tid >> 7) * 16;
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
            {
                const int k = kq + kk;
                float v = 0.0f;
                if (row0 + r < M && k0 + k < K) v = A[static_cast<size_t>(k0 + k) * lda + row0 + r];
                split(v, Ah[r][k], Al[r][k]);
            }
        }
        else
        {
            const int r = tid >> 1, kh = (tid & 1) * 16;
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
            {
                const int k = kh + kk;
                float v = 0.0f;
                if (row0 + r < M && k0 + k < K) v = A[static_cast<size_t>(row0 + r) * lda + k0 + k];
                split(v, Ah[r][k], Al[r][k]);
            }
        }
        {
            const int j = tid >> 1, kh = (tid & 1) * 16;
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
            {
                const int k = kh + kk;
                float v = 0.0f;
                if (col0 + j < N && k0 + k < K) v = B[static_cast<size_t>(col0 + j) * ldb + k0 + k];
                split(v, Bh[j][k], Bl[j][k]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 32; ks += 16)
        {
            const int kq = ks + (lane >> 5) * 8;
            const int ra = wm * 64 + (lane & 31), rb = wn * 64 + (lane & 31);
            const hx8 a0h = *reinterpret_cast<const hx8*>(&Ah[ra][kq]);
            const hx8 a1h = *reinterpret_cast<const hx8*>(&Ah[ra + 32][kq]);
            const hx8 a0l = *reinterpret_cast<const hx8*>(&Al[ra][kq]);
            const hx8 a1l = *reinterpret_cast<const hx8*>(&Al[ra + 32][kq]);
            const hx8 b0h = *reinterpret_cast<const hx8*>(&Bh[rb][kq]);
            const hx8 b1h = *reinterpret_cast<const hx8*>(&Bh[rb + 32][kq]);
            const hx8 b0l = *reinterpret_cast<const hx8*>(&Bl[rb][kq]);
            const hx8 b1l = *reinterpret_cast<const hx8*>(&Bl[rb + 32][kq]);
#define BF3(ACC, AH, AL, BH, BL)                                                \
    ACC = __builtin_amdgcn_mfma_f32_32x32x16_bf16(AL, BH, ACC, 0, 0, 0);        \
    ACC = __builtin_amdgcn_mfma_f32_32x32x16_bf16(AH, BL, ACC, 0, 0, 0);        \
    ACC = __builtin_amdgcn_mfma_f32_32x32x16_bf16(AH, BH, ACC, 0, 0, 0)
            BF3(acc[0][0], a0h, a0l, b0h, b0l);
            BF3(acc[0][1], a0h, a0l, b1h, b1l);
            BF3(acc[1][0], a1h, a1l, b0h, b0l);
            BF3(acc[1][1], a1h, a1l, b1h, b1l);
#undef BF3
        }
        __syncthreads();
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
        {
            const int j = col0 + wn * 64 + bj * 32 + (lane & 31);
            if (j >= N) continue;
            float* cj = C + static_cast<size_t>(j) * ldc;
#pragma unroll
            for (int v = 0; v < 16; ++v)
            {
                const int i = row0 + wm * 64 + bi * 32 + (v & 3) + 8 * (v >> 2) + 4 * (lane >> 5);
                if (i < M) cj[i] -= acc[bi][bj][v];
            }
        }
}

template <int CTRL>
__device__ inline float QuadBroadcast(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// B <- op(T)^{-1} B on one diagonal block of kb <= 64 rows. Four lanes (a quad) share one
// column of B, lane q holding rows 16q .. 16q+15 in registers; the unknown solved at step s
// is broadcast inside the quad with DPP, so each lane's serial chain is 64 steps of 16
// multiply-adds instead of 64 x 63 for a column per lane. op(T) sits in LDS and is read by
// all lanes of a step at the same column. Per element the update order is the LAPACK
// strsm order: ascending (forward) or descending (backward) pivot index, one fused
// multiply-add per term, one division per pivot; rows outside the triangle are left
// untouched (a select, so a non-finite unknown cannot leak into finished rows).
template <bool FWD, bool UNIT>
__global__ __launch_bounds__(256) void trsm_diag64_kernel(int kb, int N, const float* __restrict__ T, int ldt,
                                                          int transT, float* __restrict__ B, int ldb)
{
    __shared__ __attribute__((aligned(16))) float Ts[64][64];  // Ts[s][i] = op(T)(i, s)
    for (int e = threadIdx.x; e < 64 * 64; e += 256)
    {
        const int s2 = e >> 6, i = e & 63;
        float v = 0.0f;
        if (i < kb && s2 < kb) v = transT ? T[static_cast<size_t>(i) * ldt + s2] : T[static_cast<size_t>(s2) * ldt + i];
        Ts[s2][i] = v;
    }
    __syncthreads();
    const int col = blockIdx.x * 64 + (threadIdx.x >> 2), q = threadIdx.x & 3;
    const bool live = col < N;
    float* b = B + static_cast<size_t>(live ? col : 0) * ldb;
    float x[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) x[t] = (live && 16 * q + t < kb) ? b[16 * q + t] : 0.0f;
#pragma unroll
    for (int u = 0; u < 64; ++u)
    {
        const int s2 = FWD ? u : 63 - u;
        if (s2 >= kb) continue;  // uniform
        const int o = s2 >> 4, so = s2 & 15;
        float xs;
        if (o == 0)  // quad_perm [o,o,o,o]: every lane of the quad reads lane o
            xs = QuadBroadcast<0x00>(x[so]);
        else if (o == 1)
            xs = QuadBroadcast<0x55>(x[so]);
        else if (o == 2)
            xs = QuadBroadcast<0xAA>(x[so]);
        else
            xs = QuadBroadcast<0xFF>(x[so]);
        if (!UNIT) xs = xs / Ts[s2][s2];
        if (q == o) x[so] = xs;
#pragma unroll
        for (int t = 0; t < 16; ++t)
        {
            const int row = 16 * q + t;
            const bool upd = FWD ? row > s2 : row < s2;
            const float nx = fmaf(-Ts[s2][row], xs, x[t]);
            x[t] = upd ? nx : x[t];
        }
    }
#pragma unroll
    for (int t = 0; t < 16; ++t)
        if (live && 16 * q + t < kb) b[16 * q + t] = x[t];
}

// Inverses of the unit-lower 64 x 64 diagonal blocks of T (K x K): workgroup b writes
// inv(T[64b.., 64b..]) to Linv + 4096 b (column-major, ld 64; rows and columns beyond the
// block are the identity). Column c of the inverse is the forward substitution of e_c,
// run like trsm_diag64_kernel: a quad of lanes per column, 16 rows each.
__global__ __launch_bounds__(256) void trtri_unit_lower64_kernel(int K, const float* __restrict__ T, int ldt,
                                                                 float* __restrict__ Linv)
{
    __shared__ __attribute__((aligned(16))) float Ts[64][64];  // Ts[s][i] = T(i, s) of the block
    const int k0 = blockIdx.x * 64, kb = min(64, K - k0);
    const float* Tb = T + static_cast<size_t>(k0) * ldt + k0;
    for (int e = threadIdx.x; e < 64 * 64; e += 256)
    {
        const int s2 = e >> 6, i = e & 63;
        Ts[s2][i] = (i < kb && s2 < kb) ? Tb[static_cast<size_t>(s2) * ldt + i] : 0.0f;
    }
    __syncthreads();
    const int col = threadIdx.x >> 2, q = threadIdx.x & 3;
    float x[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) x[t] = (16 * q + t == col) ? 1.0f : 0.0f;
#pragma unroll
    for (int s2 = 0; s2 < 64; ++s2)
    {
        if (s2 >= kb) continue;  // uniform
        const int o = s2 >> 4, so = s2 & 15;
        float xs;
        if (o == 0)
            xs = QuadBroadcast<0x00>(x[so]);
        else if (o == 1)
            xs = QuadBroadcast<0x55>(x[so]);
        else if (o == 2)
            xs = QuadBroadcast<0xAA>(x[so]);
        else
            xs = QuadBroadcast<0xFF>(x[so]);
#pragma unroll
        for (int t = 0; t < 16; ++t)
        {
            const int row = 16 * q + t;
            const float nx = fmaf(-Ts[s2][row], xs, x[t]);
            x[t] = (row > s2 && row < kb) ? nx : x[t];
        }
    }
    float* out = Linv + static_cast<size_t>(blockIdx.x) * 4096 + col * 64 + 16 * q;
#pragma unroll
    for (int t = 0; t < 16; t += 4)
        *reinterpret_cast<float4*>(out + t) = make_float4(x[t], x[t + 1], x[t + 2], x[t + 3]);
}

// B <- T^{-1} B for unit-lower T (K x K, K <= 512) from the block inverses of
// trtri_unit_lower64_kernel: one workgroup per 32-column strip of B, held in LDS; per
// 64-row block b: X_b = inv(T_bb) B_b, then B_below -= T(below, b) X_b, both on FP32 MFMA
// (32x32x2 tiles formed as products^T so that each accumulator lane holds one row).
constexpr int kTrsmLuMaxK = 512;
__global__ __launch_bounds__(256) void trsm_lu_kernel(int K, int N, const float* __restrict__ T, int ldt,
                                                      const float* __restrict__ Linv, float* __restrict__ B, int ldb)
{
    __shared__ float Xs[kTrsmLuMaxK][33];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c0 = blockIdx.x * 32, nc = min(32, N - c0);
    const int Kp = (K + 63) & ~63;
    for (int e = tid; e < 32 * Kp; e += 256)
    {
        const int c = e / Kp, r = e - c * Kp;
        Xs[r][c] = (c < nc && r < K) ? B[static_cast<size_t>(c0 + c) * ldb + r] : 0.0f;
    }
    __syncthreads();
    const int li = lane & 31, lk = lane >> 5;
    const int nblk = Kp / 64;
    // Operand fragments are loaded for a whole 32-row tile before its 32 MFMAs, so each tile
    // waits for global memory once; the first tile below the block is fetched before the
    // block's own product.
    float lf[32];
    auto loadTile = [&](int k0, int k1, int r0)
    {
        const int row = r0 + li;
        const float* Lc = T + static_cast<size_t>(k0) * ldt + row;
#pragma unroll
        for (int t = 0; t < 32; ++t)
        {
            const int k = 2 * t + lk;
            lf[t] = (row < K && k0 + k < k1) ? Lc[static_cast<size_t>(k) * ldt] : 0.0f;
        }
    };
    for (int b = 0; b < nblk; ++b)
    {
        const int k0 = b * 64, k1 = min(K, k0 + 64);
        float lb[32];
        if (wave < 2)
        {
            const float* Lb = Linv + static_cast<size_t>(b) * 4096 + wave * 32 + li;
#pragma unroll
            for (int t = 0; t < 32; ++t) lb[t] = Lb[(2 * t + lk) * 64];
        }
        int r0 = k1 + wave * 32;
        if (r0 < K) loadTile(k0, k1, r0);
        // X_b = inv(T_bb) B_b: waves 0 and 1 take 32 rows each.
        f32x16 acc;
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
        if (wave < 2)
#pragma unroll
            for (int t = 0; t < 32; ++t)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(Xs[k0 + 2 * t + lk][li], lb[t], acc, 0, 0, 0);
        __syncthreads();
        if (wave < 2)
#pragma unroll
            for (int v = 0; v < 16; ++v) Xs[k0 + wave * 32 + li][(v & 3) + 8 * (v >> 2) + 4 * lk] = acc[v];
        __syncthreads();
        // Rows below the block: 32-row tiles dealt to the four waves.
        for (; r0 < K; r0 += 128)
        {
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
#pragma unroll
            for (int t = 0; t < 32; ++t)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(Xs[k0 + 2 * t + lk][li], lf[t], acc, 0, 0, 0);
            const int row = r0 + li;
            if (r0 + 128 < K) loadTile(k0, k1, r0 + 128);
            if (row < K)
#pragma unroll
                for (int v = 0; v < 16; ++v) Xs[row][(v & 3) + 8 * (v >> 2) + 4 * lk] -= acc[v];
        }
        __syncthreads();
    }
    for (int e = tid; e < 32 * Kp; e += 256)
    {
        const int c = e / Kp, r = e - c * Kp;
        if (c < nc && r < K) B[static_cast<size_t>(c0 + c) * ldb + r] = Xs[r][c];
    }
}

// ---------------------------------------------------------------------------------------
// Batched inverse from LU factors (getri_batched, orders 65..512): C = U^{-1} L^{-1} P,
// reference src/sgetri/real_lu_blocked_inverse_kernel.cpp:1115-1301 computes the same
// inverse blockwise per matrix; here every matrix of the batch moves through three launches
// together: the row permutation P into C, the inverses of the 64 x 64 diagonal blocks of L
// (unit lower) and U (upper), then the two block substitutions on FP32 MFMA.
// ---------------------------------------------------------------------------------------

// C(:, :) = P, P(i, perm(i)) = 1 where perm is ipiv's interchange sequence applied to the
// identity (row i of P A is row perm(i) of A); no pivots: the identity. One workgroup per
// matrix: the interchanges run on an LDS index array, then the columns are written.
__global__ __launch_bounds__(256) void getri_perm_identity_kernel(int n, const int* __restrict__ piv,
                                                                  float* const* __restrict__ Carr, int ldc)
{
    extern __shared__ int s_perm[];
    const int b = blockIdx.x;
    for (int i = threadIdx.x; i < n; i += 256) s_perm[i] = i;
    __syncthreads();
    if (threadIdx.x == 0 && piv != nullptr)
    {
        const int* p = piv + static_cast<size_t>(b) * n;
        for (int k = 0; k < n; ++k)
        {
            const int q = p[k] - 1;
            if (q != k && q >= 0 && q < n)
            {
                const int t = s_perm[k];
                s_perm[k] = s_perm[q];
                s_perm[q] = t;
            }
        }
    }
    __syncthreads();
    float* C = Carr[b];
    for (int c = 0; c < n; ++c)
        for (int i = threadIdx.x; i < n; i += 256) C[static_cast<size_t>(c) * ldc + i] = s_perm[i] == c ? 1.0f : 0.0f;
}

// info[b] = first zero U(k, k) + 1 (0: none), as getri reports a singular factor.
__global__ __launch_bounds__(256) void getri_first_zero_kernel(int n, const float* const* __restrict__ Aarr, int lda,
                                                               int* __restrict__ info)
{
    __shared__ int s_first;
    if (threadIdx.x == 0) s_first = INT_MAX;
    __syncthreads();
    const float* A = Aarr[blockIdx.x];
    for (int k = threadIdx.x; k < n; k += 256)
        if (A[static_cast<size_t>(k) * lda + k] == 0.0f) atomicMin(&s_first, k);
    __syncthreads();
    if (threadIdx.x == 0) info[blockIdx.x] = s_first == INT_MAX ? 0 : s_first + 1;
}

// Inverses of the 64 x 64 diagonal blocks of every matrix's unit-lower L (UPPER = false)
// or upper U (UPPER = true): workgroup (block bi, matrix b) writes the inverse to
// inv + (b * nblk + bi) * 4096, column-major, ld 64, identity beyond the block. Column c is
// the substitution of e_c, a quad of lanes per column (16 rows each), as trtri_unit_lower64.
template <bool UPPER>
__global__ __launch_bounds__(256) void trtri_diag64_batched_kernel(int K, const float* const* __restrict__ Tarr,
                                                                   int ldt, float* __restrict__ inv)
{
    __shared__ __attribute__((aligned(16))) float Ts[64][64];  // Ts[s][i] = T(i, s) of the block
    const int nblk = (K + 63) / 64;
    const int bi = blockIdx.x, k0 = bi * 64, kb = min(64, K - k0);
    const float* Tb = Tarr[blockIdx.y] + static_cast<size_t>(k0) * ldt + k0;
    for (int e = threadIdx.x; e < 64 * 64; e += 256)
    {
        const int s2 = e >> 6, i = e & 63;
        Ts[s2][i] = (i < kb && s2 < kb) ? Tb[static_cast<size_t>(s2) * ldt + i] : 0.0f;
    }
    __syncthreads();
    const int col = threadIdx.x >> 2, q = threadIdx.x & 3;
    float x[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) x[t] = (16 * q + t == col) ? 1.0f : 0.0f;
#pragma unroll
    for (int u = 0; u < 64; ++u)
    {
        const int s2 = UPPER ? 63 - u : u;
        if (s2 >= kb) continue;  // uniform
        const int o = s2 >> 4, so = s2 & 15;
        float xs;
        if (o == 0)
            xs = QuadBroadcast<0x00>(x[so]);
        else if (o == 1)
            xs = QuadBroadcast<0x55>(x[so]);
        else if (o == 2)
            xs = QuadBroadcast<0xAA>(x[so]);
        else
            xs = QuadBroadcast<0xFF>(x[so]);
        if (UPPER)
        {
            xs = xs / Ts[s2][s2];
            if (q == o) x[so] = xs;
        }
#pragma unroll
        for (int t = 0; t < 16; ++t)
        {
            const int row = 16 * q + t;
            const bool upd = UPPER ? row < s2 : (row > s2 && row < kb);
            const float nx = fmaf(-Ts[s2][row], xs, x[t]);
            x[t] = upd ? nx : x[t];
        }
    }
    float* out = inv + (static_cast<size_t>(blockIdx.y) * nblk + bi) * 4096 + col * 64 + 16 * q;
#pragma unroll
    for (int t = 0; t < 16; t += 4)
        *reinterpret_cast<float4*>(out + t) = make_float4(x[t], x[t + 1], x[t + 2], x[t + 3]);
}

// B <- T^{-1} B for every matrix of the batch (T unit lower from L, or upper from U; K <=
// 512): workgroup (strip, matrix) holds a 32-column strip of B in LDS; per 64-row block,
// in ascending (lower) or descending (upper) order: X_b = inv(T_bb) B_b, then the other
// unsolved rows -= T(rows, b) X_b, both on FP32 MFMA -- trsm_lu_kernel's scheme.
template <bool UPPER>
__global__ __launch_bounds__(256) void trsm_blocked_batched_kernel(int K, int N, const float* const* __restrict__ Tarr,
                                                                   int ldt, const float* __restrict__ inv,
                                                                   float* const* __restrict__ Barr, int ldb)
{
    __shared__ float Xs[kTrsmLuMaxK][33];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const float* T = Tarr[blockIdx.y];
    float* B = Barr[blockIdx.y];
    const int c0 = blockIdx.x * 32, nc = min(32, N - c0);
    const int Kp = (K + 63) & ~63;
    const int nblk = Kp / 64;
    const float* Inv = inv + static_cast<size_t>(blockIdx.y) * nblk * 4096;
    for (int e = tid; e < 32 * Kp; e += 256)
    {
        const int c = e / Kp, r = e - c * Kp;
        Xs[r][c] = (c < nc && r < K) ? B[static_cast<size_t>(c0 + c) * ldb + r] : 0.0f;
    }
    __syncthreads();
    const int li = lane & 31, lk = lane >> 5;
    float lf[32];
    auto loadTile = [&](int k0, int k1, int r0)
    {
        const int row = r0 + li;
        const float* Lc = T + static_cast<size_t>(k0) * ldt + row;
#pragma unroll
        for (int t = 0; t < 32; ++t)
        {
            const int k = 2 * t + lk;
            lf[t] = (row < K && k0 + k < k1) ? Lc[static_cast<size_t>(k) * ldt] : 0.0f;
        }
    };
    for (int q = 0; q < nblk; ++q)
    {
        const int b = UPPER ? nblk - 1 - q : q;
        const int k0 = b * 64, k1 = min(K, k0 + 64);
        // rows still to update: below the block (lower) or above it (upper)
        const int rLo = UPPER ? 0 : k1, rHi = UPPER ? k0 : K;
        float lb[32];
        if (wave < 2)
        {
            const float* Lb = Inv + static_cast<size_t>(b) * 4096 + wave * 32 + li;
#pragma unroll
            for (int t = 0; t < 32; ++t) lb[t] = Lb[(2 * t + lk) * 64];
        }
        int r0 = rLo + wave * 32;
        if (r0 < rHi) loadTile(k0, k1, r0);
        f32x16 acc;
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
        if (wave < 2)
#pragma unroll
            for (int t = 0; t < 32; ++t)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(Xs[k0 + 2 * t + lk][li], lb[t], acc, 0, 0, 0);
        __syncthreads();
        if (wave < 2)
#pragma unroll
            for (int v = 0; v < 16; ++v) Xs[k0 + wave * 32 + li][(v & 3) + 8 * (v >> 2) + 4 * lk] = acc[v];
        __syncthreads();
        for (; r0 < rHi; r0 += 128)
        {
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
#pragma unroll
            for (int t = 0; t < 32; ++t)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(Xs[k0 + 2 * t + lk][li], lf[t], acc, 0, 0, 0);
            const int row = r0 + li;
            if (r0 + 128 < rHi) loadTile(k0, k1, r0 + 128);
            if (row < rHi)
#pragma unroll
                for (int v = 0; v < 16; ++v) Xs[row][(v & 3) + 8 * (v >> 2) + 4 * lk] -= acc[v];
        }
        __syncthreads();
    }
    for (int e = tid; e < 32 * Kp; e += 256)
    {
        const int c = e / Kp, r = e - c * Kp;
        if (c < nc && r < K) B[static_cast<size_t>(c0 + c) * ldb + r] = Xs[r][c];
    }
}

hipError_t TrsmDiag(int kb, int N, const float* T, int ldt, bool fwd, bool unit, bool transT, float* B, int ldb,
                    hipStream_t stream)
{
    const dim3 grid((N + 63) / 64);
    const int tr = transT ? 1 : 0;
    if (fwd && unit)
        hipLaunchKernelGGL((trsm_diag64_kernel<true, true>), grid, dim3(256), 0, stream, kb, N, T, ldt, tr, B, ldb);
    else if (fwd)
        hipLaunchKernelGGL((trsm_diag64_kernel<true, false>), grid, dim3(256), 0, stream, kb, N, T, ldt, tr, B, ldb);
    else if (unit)
        hipLaunchKernelGGL((trsm_diag64_kernel<false, true>), grid, dim3(256), 0, stream, kb, N, T, ldt, tr, B, ldb);
    else
        hipLaunchKernelGGL((trsm_diag64_kernel<false, false>), grid, dim3(256), 0, stream, kb, N, T, ldt, tr, B, ldb);
    return hipGetLastError();
}

}  // namespace

hipError_t GemmSub(int M, int N, int K, const float* A, int lda, int transA, const float* B, int ldb, float* C,
                   int ldc, Precision prec, hipStream_t stream)
{
    if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
    const int tilesM = (M + TM - 1) / TM, tilesN = (N + TN - 1) / TN;
    const int blocks = tilesM * tilesN;
    // The split forms are (near) FP32-accurate; without packed operands they run as FP32.
    if (prec == Precision::FP32 || prec == Precision::BF16X6 || prec == Precision::FP16X3)
    {
        if (!transA && blocks < 512)
        {
            const int tm64 = (M + 63) / 64, tn64 = (N + 63) / 64;
            hipLaunchKernelGGL(gemm_sub_f32_64_kernel, dim3(tm64 * tn64), dim3(256), 0, stream, M, N, K, A, lda, B, ldb,
                               C, ldc, tm64, tn64);
        }
        else if (transA)
            hipLaunchKernelGGL(gemm_sub_f32_kernel<1>, dim3(blocks), dim3(256), 0, stream, M, N, K, A, lda, B, ldb, C,
                               ldc, tilesM, tilesN);
        else
            hipLaunchKernelGGL(gemm_sub_f32_kernel<0>, dim3(blocks), dim3(256), 0, stream, M, N, K, A, lda, B, ldb, C,
                               ldc, tilesM, tilesN);
    }
    else if (prec == Precision::FP16)
    {
        if (transA)
            hipLaunchKernelGGL((gemm_sub_half_kernel<_Float16, 1, true>), dim3(blocks), dim3(256), 0, stream, M, N, K,
                               A, lda, B, ldb, C, ldc, tilesM, tilesN);
        else
            hipLaunchKernelGGL((gemm_sub_half_kernel<_Float16, 0, true>), dim3(blocks), dim3(256), 0, stream, M, N, K,
                               A, lda, B, ldb, C, ldc, tilesM, tilesN);
    }
    else if (prec == Precision::BF16X3)
    {
        if (transA)
            hipLaunchKernelGGL(gemm_sub_bf16x3_kernel<1>, dim3(blocks), dim3(256), 0, stream, M, N, K, A, lda, B, ldb,
                               C, ldc, tilesM, tilesN);
        else
            hipLaunchKernelGGL(gemm_sub_bf16x3_kernel<0>, dim3(blocks), dim3(256), 0, stream, M, N, K, A, lda, B, ldb,
                               C, ldc, tilesM, tilesN);
    }
    else
    {
        if (transA)
            hipLaunchKernelGGL((gemm_sub_half_kernel<__bf16, 1, false>), dim3(blocks), dim3(256), 0, stream, M, N, K,
                               A, lda, B, ldb, C, ldc, tilesM, tilesN);
        else
            hipLaunchKernelGGL((gemm_sub_half_kernel<__bf16, 0, false>), dim3(blocks), dim3(256), 0, stream, M, N, K,
                               A, lda, B, ldb, C, ldc, tilesM, tilesN);
    }
    return hipGetLastError();
}

size_t TrsmLowerUnitWorkspaceBytes(int K) { return static_cast<size_t>((K + 63) / 64) * 4096 * sizeof(float); }

hipError_t TrsmLowerUnitInverse(int K, const float* T, int ldt, float* linv, hipStream_t stream)
{
    if (K <= 0 || K > kTrsmLuMaxK) return K <= 0 ? hipSuccess : hipErrorInvalidValue;
    hipLaunchKernelGGL(trtri_unit_lower64_kernel, dim3((K + 63) / 64), dim3(256), 0, stream, K, T, ldt, linv);
    return hipGetLastError();
}

hipError_t TrsmLowerUnitApply(int K, int N, const float* T, int ldt, const float* linv, float* B, int ldb,
                              hipStream_t stream)
{
    if (K <= 0 || N <= 0) return hipSuccess;
    if (K > kTrsmLuMaxK) return hipErrorInvalidValue;
    hipLaunchKernelGGL(trsm_lu_kernel, dim3((N + 31) / 32), dim3(256), 0, stream, K, N, T, ldt, linv, B, ldb);
    return hipGetLastError();
}

size_t GetriBlockedBatchedWorkspaceBytes(int n, int batch)
{
    return static_cast<size_t>(2) * ((n + 63) / 64) * 4096 * sizeof(float) * static_cast<size_t>(batch > 0 ? batch : 0);
}

hipError_t GetriBlockedBatched(int n, const float* const* Aarr, int lda, const int* piv, float* const* Carr, int ldc,
                               int* info, int batch, float* work, hipStream_t stream)
{
    if (n <= 0 || batch <= 0) return hipSuccess;
    if (n > kTrsmLuMaxK || batch > 65535) return hipErrorInvalidValue;
    const int nblk = (n + 63) / 64;
    float* linv = work;
    float* uinv = work + static_cast<size_t>(batch) * nblk * 4096;
    hipLaunchKernelGGL(getri_first_zero_kernel, dim3(batch), dim3(256), 0, stream, n, Aarr, lda, info);
    hipLaunchKernelGGL(getri_perm_identity_kernel, dim3(batch), dim3(256), static_cast<size_t>(n) * sizeof(int), stream,
                       n, piv, Carr, ldc);
    hipLaunchKernelGGL(trtri_diag64_batched_kernel<false>, dim3(nblk, batch), dim3(256), 0, stream, n, Aarr, lda, linv);
    hipLaunchKernelGGL(trtri_diag64_batched_kernel<true>, dim3(nblk, batch), dim3(256), 0, stream, n, Aarr, lda, uinv);
    const dim3 grid((n + 31) / 32, batch);
    hipLaunchKernelGGL(trsm_blocked_batched_kernel<false>, grid, dim3(256), 0, stream, n, n, Aarr, lda, linv, Carr, ldc);
    hipLaunchKernelGGL(trsm_blocked_batched_kernel<true>, grid, dim3(256), 0, stream, n, n, Aarr, lda, uinv, Carr, ldc);
    return hipGetLastError();
}

hipError_t TrsmBlock(int K, int N, const float* T, int ldt, bool lower, bool unit, bool transT, float* B, int ldb,
                     hipStream_t stream)
{
    if (K <= 0 || N <= 0) return hipSuccess;
    // op(T) is lower-triangular (forward substitution) when lower != transT.
    const bool fwd = lower != transT;
    const int nblk = (K + 63) / 64;
    for (int q = 0; q < nblk; ++q)
    {
        const int bi = fwd ? q : nblk - 1 - q;
        const int k0 = bi * 64, kb = std::min(64, K - k0), k1 = k0 + kb;
        const float* Tkk = T + static_cast<size_t>(k0) * ldt + k0;
        hipError_t e = TrsmDiag(kb, N, Tkk, ldt, fwd, unit, transT, B + k0, ldb, stream);
        if (e != hipSuccess) return e;
        // Fold the solved rows X = B[k0:k1) into the rows still to be solved.
        if (fwd && k1 < K)
        {
            // rows [k1, K) of op(T), columns [k0, k1)
            const float* Ablk = transT ? T + static_cast<size_t>(k1) * ldt + k0 : T + static_cast<size_t>(k0) * ldt + k1;
            e = GemmSub(K - k1, N, kb, Ablk, ldt, transT ? 1 : 0, B + k0, ldb, B + k1, ldb, Precision::FP32, stream);
        }
        else if (!fwd && k0 > 0)
        {
            // rows [0, k0) of op(T), columns [k0, k1)
            const float* Ablk = transT ? T + k0 : T + static_cast<size_t>(k0) * ldt;
            e = GemmSub(k0, N, kb, Ablk, ldt, transT ? 1 : 0, B + k0, ldb, B, ldb, Precision::FP32, stream);
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace ops
