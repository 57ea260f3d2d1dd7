// Mixed-precision trailing update C -= A * B on packed BF16 operands.
//
// The trailing update of the mixed-precision LU (the reference's FP32 update is
// src/sgetrf/real_lu_blocked_kernel.cpp:1119-1158) reads every element of the L panel
// once per 128-column tile of C and every element of the U block row once per 128-row
// tile. Converting FP32 -> BF16 on each of those reads wastes VALU time and doubles the
// bytes moved, so the two operands are converted once per outer step into "fragment
// order": for a 32-row block rb and a 16-deep k block kb, the 64 lanes' 8-element bf16
// vectors that one v_mfma_f32_32x32x16_bf16 consumes are stored contiguously (1 KiB), the
// high part (bf16(x)) followed by the low part (bf16(x - hi)) of the split form.
//
// The GEMM then streams whole fragments (16-byte loads, conflict-free 16-byte LDS
// accesses, double-buffered, one barrier per 32-deep stage) and computes C^T's tile
// (src0 = U fragment, src1 = L fragment) so that each accumulator register maps to 32
// consecutive rows of one column of C: the read-modify-write of C is coalesced.
//
// Split form: x*y ~ hi(x)hi(y) + hi(x)lo(y) + lo(x)hi(y) with FP32 accumulation (about
// 16 significant bits per product); the plain form uses hi(x)hi(y) only.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lu_kernels.h"

namespace ops
{
namespace
{
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kTile = 128;  // C tile per workgroup (rows and columns)
constexpr int kFragRows = 32;
constexpr int kFragK = 16;
constexpr int kStageK = 2;  // k blocks per pipeline stage

inline int RoundUp(int v, int m) { return (v + m - 1) / m * m; }

// Packs the `rows x K` operand X(r, k) = src[r * rs + k * ks] into fragment order:
// fragment f = rb * KB + kb holds, for lane l, X(rb*32 + l%32, kb*16 + (l/32)*8 + i),
// i < 8, as hi (part 0) and, in the split form, lo (part 1) vectors. Out-of-range
// elements are zero.
template <bool SPLIT>
__global__ __launch_bounds__(256) void pack_frag_kernel(int rows, int K, const float* __restrict__ src, int64_t rs,
                                                        int64_t ks, int KB, int nfrag, bf16x8* __restrict__ out)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int f = t >> 6, l = t & 63;
    if (f >= nfrag) return;
    const int rb = f / KB, kb = f % KB;
    const int r = rb * kFragRows + (l & 31);
    const int k0 = kb * kFragK + (l >> 5) * 8;
    bf16x8 hi, lo;
#pragma unroll
    for (int i = 0; i < 8; ++i)
    {
        const int k = k0 + i;
        const float v = (r < rows && k < K) ? src[r * rs + k * ks] : 0.0f;
        const __bf16 h = static_cast<__bf16>(v);
        hi[i] = h;
        lo[i] = SPLIT ? static_cast<__bf16>(v - static_cast<float>(h)) : static_cast<__bf16>(0.0f);
    }
    constexpr int PARTS = SPLIT ? 2 : 1;
    out[(static_cast<size_t>(f) * PARTS + 0) * 64 + l] = hi;
    if (SPLIT) out[(static_cast<size_t>(f) * PARTS + 1) * 64 + l] = lo;
}

__device__ inline void TileOfXcd(int tilesM, int tilesN, int* tm, int* tn)
{
    // Consecutive workgroups go round-robin to the 8 XCDs; give each XCD a contiguous run
    // of tiles (column-major tile order) so that neighbouring tiles share L2.
    const int total = tilesM * tilesN;
    int id = blockIdx.x;
    if (total % 8 == 0)
    {
        const int per = total / 8;
        id = (id % 8) * per + id / 8;
    }
    *tm = id % tilesM;
    *tn = id / tilesM;
}

// One workgroup (4 waves) per 128 x 128 tile of C; wave (wm, wn) owns the 64 x 64
// quadrant rows [wm*64, +64), columns [wn*64, +64) as 2 x 2 accumulators of 32 x 32.
template <bool SPLIT>
__global__ __launch_bounds__(256) void gemm_packed_kernel(int M, int N, int KB, const bf16x8* __restrict__ Ap,
                                                          const bf16x8* __restrict__ Bp, float* __restrict__ C, int ldc,
                                                          int tilesM, int tilesN)
{
    constexpr int PARTS = SPLIT ? 2 : 1;
    // [stage][block 0..3][kk][part][lane]
    __shared__ bf16x8 sA[2][4][kStageK][PARTS][64];
    __shared__ bf16x8 sB[2][4][kStageK][PARTS][64];
    int tm, tn;
    TileOfXcd(tilesM, tilesN, &tm, &tn);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int rb0 = tm * 4, cb0 = tn * 4;

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

    // Each stage moves 4 blocks x kStageK x PARTS fragments of A and of B (64 vectors
    // each); thread t moves vector t + 256 q of each.
    constexpr int VEC = 4 * kStageK * PARTS * 64;  // vectors per operand per stage
    constexpr int PER = VEC / 256;
    bf16x8 ra[PER], rbv[PER];
    auto gload = [&](int stage)
    {
#pragma unroll
        for (int q = 0; q < PER; ++q)
        {
            const int idx = tid + 256 * q;
            const int l = idx & 63, part = (idx >> 6) % PARTS, kk = (idx / (64 * PARTS)) % kStageK,
                      blk = idx / (64 * PARTS * kStageK);
            const int kb = stage * kStageK + kk;
            ra[q] = Ap[((static_cast<size_t>(rb0 + blk) * KB + kb) * PARTS + part) * 64 + l];
            rbv[q] = Bp[((static_cast<size_t>(cb0 + blk) * KB + kb) * PARTS + part) * 64 + l];
        }
    };
    auto lstore = [&](int buf)
    {
#pragma unroll
        for (int q = 0; q < PER; ++q)
        {
            const int idx = tid + 256 * q;
            const int l = idx & 63, part = (idx >> 6) % PARTS, kk = (idx / (64 * PARTS)) % kStageK,
                      blk = idx / (64 * PARTS * kStageK);
            sA[buf][blk][kk][part][l] = ra[q];
            sB[buf][blk][kk][part][l] = rbv[q];
        }
    };

    const int stages = KB / kStageK;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int s = 0; s < stages; ++s)
    {
        const int buf = s & 1;
        if (s + 1 < stages) gload(s + 1);
#pragma unroll
        for (int kk = 0; kk < kStageK; ++kk)
        {
            bf16x8 ah[2], bh[2], al[2], bl[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
            {
                ah[i] = sA[buf][wm * 2 + i][kk][0][lane];
                bh[i] = sB[buf][wn * 2 + i][kk][0][lane];
                if (SPLIT)
                {
                    al[i] = sA[buf][wm * 2 + i][kk][PARTS - 1][lane];
                    bl[i] = sB[buf][wn * 2 + i][kk][PARTS - 1][lane];
                }
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                {
                    // C^T tile: rows of the product are columns of C (src0 = U fragment).
                    if (SPLIT)
                    {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
                    }
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
                }
        }
        if (s + 1 < stages) lstore(buf ^ 1);
        __syncthreads();
    }

    // Accumulator register v of lane l: product row (v%4) + 8(v/4) + 4(l/32) -> column of
    // C, product column l%32 -> row of C.
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        const int row = (rb0 + wm * 2 + i) * kFragRows + (lane & 31);
        if (row >= M) continue;
#pragma unroll
        for (int j = 0; j < 2; ++j)
        {
            const int cbase = (cb0 + wn * 2 + j) * kFragRows + 4 * (lane >> 5);
#pragma unroll
            for (int v = 0; v < 16; ++v)
            {
                const int col = cbase + (v & 3) + 8 * (v >> 2);
                if (col < N)
                {
                    float* p = C + static_cast<size_t>(col) * ldc + row;
                    *p -= acc[i][j][v];
                }
            }
        }
    }
}

struct PackLayout
{
    int KB;        // k blocks (padded to the stage depth)
    int rowBlocks; // 32-row blocks of A (padded to the tile)
    int colBlocks; // 32-column blocks of B (padded to the tile)
    size_t aVecs, bVecs;
};

PackLayout Layout(int M, int N, int K, bool split)
{
    PackLayout p;
    p.KB = RoundUp((K + kFragK - 1) / kFragK, kStageK);
    p.rowBlocks = RoundUp(M, kTile) / kFragRows;
    p.colBlocks = RoundUp(N, kTile) / kFragRows;
    const int parts = split ? 2 : 1;
    p.aVecs = static_cast<size_t>(p.rowBlocks) * p.KB * parts * 64;
    p.bVecs = static_cast<size_t>(p.colBlocks) * p.KB * parts * 64;
    return p;
}

}  // namespace

size_t GemmPackedWorkspaceBytes(int M, int N, int K)
{
    if (M <= 0 || N <= 0 || K <= 0) return 0;
    const PackLayout p = Layout(M, N, K, true);
    return (p.aVecs + p.bVecs) * sizeof(bf16x8) + 256;
}

hipError_t GemmSubPacked(int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
                         Precision prec, void* work, hipStream_t stream)
{
    if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
    if (prec != Precision::BF16 && prec != Precision::BF16X3)
        return GemmSub(M, N, K, A, lda, 0, B, ldb, C, ldc, prec, stream);
    const bool split = prec == Precision::BF16X3;
    const PackLayout p = Layout(M, N, K, split);
    bf16x8* ap = static_cast<bf16x8*>(work);
    bf16x8* bp = ap + p.aVecs;
    const int fa = p.rowBlocks * p.KB, fb = p.colBlocks * p.KB;
    // A(r, k) = A[k * lda + r]; B^T(c, k) = B[c * ldb + k]
    if (split)
    {
        hipLaunchKernelGGL(pack_frag_kernel<true>, dim3((fa * 64 + 255) / 256), dim3(256), 0, stream, M, K, A,
                           int64_t{1}, static_cast<int64_t>(lda), p.KB, fa, ap);
        hipLaunchKernelGGL(pack_frag_kernel<true>, dim3((fb * 64 + 255) / 256), dim3(256), 0, stream, N, K, B,
                           static_cast<int64_t>(ldb), int64_t{1}, p.KB, fb, bp);
    }
    else
    {
        hipLaunchKernelGGL(pack_frag_kernel<false>, dim3((fa * 64 + 255) / 256), dim3(256), 0, stream, M, K, A,
                           int64_t{1}, static_cast<int64_t>(lda), p.KB, fa, ap);
        hipLaunchKernelGGL(pack_frag_kernel<false>, dim3((fb * 64 + 255) / 256), dim3(256), 0, stream, N, K, B,
                           static_cast<int64_t>(ldb), int64_t{1}, p.KB, fb, bp);
    }
    const int tilesM = p.rowBlocks / 4, tilesN = p.colBlocks / 4;
    const dim3 grid(tilesM * tilesN);
    if (split)
        hipLaunchKernelGGL(gemm_packed_kernel<true>, grid, dim3(256), 0, stream, M, N, p.KB, ap, bp, C, ldc, tilesM,
                           tilesN);
    else
        hipLaunchKernelGGL(gemm_packed_kernel<false>, grid, dim3(256), 0, stream, M, N, p.KB, ap, bp, C, ldc, tilesM,
                           tilesN);
    return hipGetLastError();
}

}  // namespace ops
