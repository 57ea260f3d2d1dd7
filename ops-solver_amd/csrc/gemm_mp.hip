// Mixed-precision trailing update C -= A * B on packed BF16 operands.
//
// The trailing update of the mixed-precision LU (the reference's FP32 update is
// src/sgetrf/real_lu_blocked_kernel.cpp:1119-1158) reads every element of the L panel
// once per 128-column tile of C and every element of the U block row once per 128-row
// tile. Converting FP32 -> BF16 on each of those reads wastes VALU time and doubles the
// bytes moved, so the two operands are converted once per outer step into "fragment
// order": for a 32-row block rb and a 16-deep k block kb, the 64 lanes' 8-element bf16
// vectors that one v_mfma_f32_32x32x16_bf16 consumes are stored contiguously (1 KiB), one
// such block per part of the split form.
//
// The GEMM then streams whole fragments (16-byte loads, conflict-free 16-byte LDS
// accesses, double-buffered, one barrier per 32-deep stage) and computes C^T's tile
// (src0 = U fragment, src1 = L fragment) so that each accumulator register maps to 32
// consecutive rows of one column of C: the read-modify-write of C is coalesced.
//
// Split forms (FP32 accumulation): two parts x = h + m (hi = bf16(x), m = bf16(x - h)),
// x*y ~ h h' + h m' + m h' (about 16 significant bits per product); three parts
// x = h + m + l, x*y ~ h h' + h m' + m h' + h l' + m m' + l h' (about 24 bits, the
// accuracy of an FP32 product); the plain form uses h h' only.
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
constexpr int kStageK = 2;  // k blocks per pipeline stage (1 for the three-part form)
template <int PARTS>
constexpr int StageK()
{
    return PARTS == 3 ? 1 : 2;
}

inline int RoundUp(int v, int m) { return (v + m - 1) / m * m; }

// Packs the `rows x K` operand X(r, k) = src[r * rs + k * ks] into fragment order:
// fragment f = rb * KB + kb holds, for lane l, X(rb*32 + l%32, kb*16 + (l/32)*8 + i),
// i < 8, as PARTS split vectors (part 0 = bf16(x), part p = bf16 of what parts < p leave).
// Out-of-range elements are zero.
template <int PARTS>
__global__ __launch_bounds__(256) void pack_frag_kernel(int rows, int K, const float* __restrict__ src, int64_t rs,
                                                        int64_t ks, int KB, int nfrag, bf16x8* __restrict__ out)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int f = t >> 6, l = t & 63;
    if (f >= nfrag) return;
    const int rb = f / KB, kb = f % KB;
    const int r = rb * kFragRows + (l & 31);
    const int k0 = kb * kFragK + (l >> 5) * 8;
    bf16x8 part[PARTS];
#pragma unroll
    for (int i = 0; i < 8; ++i)
    {
        const int k = k0 + i;
        float v = (r < rows && k < K) ? src[r * rs + k * ks] : 0.0f;
#pragma unroll
        for (int p = 0; p < PARTS; ++p)
        {
            const __bf16 h = static_cast<__bf16>(v);
            part[p][i] = h;
            v -= static_cast<float>(h);
        }
    }
#pragma unroll
    for (int p = 0; p < PARTS; ++p) out[(static_cast<size_t>(f) * PARTS + p) * 64 + l] = part[p];
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
template <int PARTS>
__global__ __launch_bounds__(256) void gemm_packed_kernel(int M, int N, int KB, const bf16x8* __restrict__ Ap,
                                                          const bf16x8* __restrict__ Bp, float* __restrict__ C, int ldc,
                                                          int tilesM, int tilesN)
{
    constexpr int SK = StageK<PARTS>();
    // [stage][block 0..3][kk][part][lane]
    __shared__ bf16x8 sA[2][4][SK][PARTS][64];
    __shared__ bf16x8 sB[2][4][SK][PARTS][64];
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
    constexpr int VEC = 4 * SK * PARTS * 64;  // vectors per operand per stage
    constexpr int PER = VEC / 256;
    bf16x8 ra[PER], rbv[PER];
    auto gload = [&](int stage)
    {
#pragma unroll
        for (int q = 0; q < PER; ++q)
        {
            const int idx = tid + 256 * q;
            const int l = idx & 63, part = (idx >> 6) % PARTS, kk = (idx / (64 * PARTS)) % SK,
                      blk = idx / (64 * PARTS * SK);
            const int kb = stage * SK + kk;
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
            const int l = idx & 63, part = (idx >> 6) % PARTS, kk = (idx / (64 * PARTS)) % SK,
                      blk = idx / (64 * PARTS * SK);
            sA[buf][blk][kk][part][l] = ra[q];
            sB[buf][blk][kk][part][l] = rbv[q];
        }
    };

    const int stages = KB / SK;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int s = 0; s < stages; ++s)
    {
        const int buf = s & 1;
        if (s + 1 < stages) gload(s + 1);
#pragma unroll
        for (int kk = 0; kk < SK; ++kk)
        {
            bf16x8 a[PARTS][2], bb[PARTS][2];
#pragma unroll
            for (int p = 0; p < PARTS; ++p)
#pragma unroll
                for (int i = 0; i < 2; ++i)
                {
                    a[p][i] = sA[buf][wm * 2 + i][kk][p][lane];
                    bb[p][i] = sB[buf][wn * 2 + i][kk][p][lane];
                }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                {
                    // C^T tile: rows of the product are columns of C (src0 = U fragment).
                    // Smallest terms first: l h', m m', h l', m h', h m', h h'.
#define MF(X, Y) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bb[X][j], a[Y][i], acc[i][j], 0, 0, 0)
                    if constexpr (PARTS == 3)
                    {
                        MF(0, 2);
                        MF(1, 1);
                        MF(2, 0);
                    }
                    if constexpr (PARTS >= 2)
                    {
                        MF(0, 1);
                        MF(1, 0);
                    }
                    MF(0, 0);
#undef MF
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

PackLayout Layout(int M, int N, int K, int parts)
{
    PackLayout p;
    p.KB = RoundUp((K + kFragK - 1) / kFragK, kStageK);
    p.rowBlocks = RoundUp(M, kTile) / kFragRows;
    p.colBlocks = RoundUp(N, kTile) / kFragRows;
    p.aVecs = static_cast<size_t>(p.rowBlocks) * p.KB * parts * 64;
    p.bVecs = static_cast<size_t>(p.colBlocks) * p.KB * parts * 64;
    return p;
}

}  // namespace

size_t GemmPackedWorkspaceBytes(int M, int N, int K)
{
    if (M <= 0 || N <= 0 || K <= 0) return 0;
    const PackLayout p = Layout(M, N, K, 3);
    return (p.aVecs + p.bVecs) * sizeof(bf16x8) + 256;
}

hipError_t GemmSubPacked(int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
                         Precision prec, void* work, hipStream_t stream)
{
    if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
    if (prec != Precision::BF16 && prec != Precision::BF16X3 && prec != Precision::BF16X6)
        return GemmSub(M, N, K, A, lda, 0, B, ldb, C, ldc, prec, stream);
    const int parts = prec == Precision::BF16X6 ? 3 : (prec == Precision::BF16X3 ? 2 : 1);
    const PackLayout p = Layout(M, N, K, parts);
    bf16x8* ap = static_cast<bf16x8*>(work);
    bf16x8* bp = ap + p.aVecs;
    const int fa = p.rowBlocks * p.KB, fb = p.colBlocks * p.KB;
    const int tilesM = p.rowBlocks / 4, tilesN = p.colBlocks / 4;
    const dim3 grid(tilesM * tilesN);
    // A(r, k) = A[k * lda + r]; B^T(c, k) = B[c * ldb + k]
#define PACK_AND_RUN(P)                                                                                           \
    hipLaunchKernelGGL(pack_frag_kernel<P>, dim3((fa * 64 + 255) / 256), dim3(256), 0, stream, M, K, A, int64_t{1}, \
                       static_cast<int64_t>(lda), p.KB, fa, ap);                                                 \
    hipLaunchKernelGGL(pack_frag_kernel<P>, dim3((fb * 64 + 255) / 256), dim3(256), 0, stream, N, K, B,           \
                       static_cast<int64_t>(ldb), int64_t{1}, p.KB, fb, bp);                                     \
    hipLaunchKernelGGL(gemm_packed_kernel<P>, grid, dim3(256), 0, stream, M, N, p.KB, ap, bp, C, ldc, tilesM, tilesN)
    if (parts == 3)
    {
        PACK_AND_RUN(3);
    }
    else if (parts == 2)
    {
        PACK_AND_RUN(2);
    }
    else
    {
        PACK_AND_RUN(1);
    }
#undef PACK_AND_RUN
    return hipGetLastError();
}

}  // namespace ops
