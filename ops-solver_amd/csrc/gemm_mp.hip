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
// FP16 form: the two-part split in FP16 (about 22 bits per product), every row of the L
// panel and column of the U block row scaled by a power of two first (FP16's range is
// narrow), the factors undone exactly in the epilogue.
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
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// Element type of the packed operands and the matching MFMA.
template <typename ET>
struct Mfma;
template <>
struct Mfma<__bf16>
{
    typedef bf16x8 vec;
    static __device__ inline f32x16 run(vec a, vec b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
};
template <>
struct Mfma<_Float16>
{
    typedef f16x8 vec;
    static __device__ inline f32x16 run(vec a, vec b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }
};

// Power of two that brings a row's largest magnitude into [2^14, 2^15) (FP16 keeps 2^15 <
// 65504); clamped so the factor and its inverse stay finite normal floats.
__device__ inline float RowScale(float maxAbs)
{
    if (!(maxAbs > 0.0f) || !(maxAbs <= 3.0e38f)) return 1.0f;
    int e = 14 - ilogbf(maxAbs);
    e = e > 100 ? 100 : (e < -100 ? -100 : e);
    return ldexpf(1.0f, e);
}

// scale[r] = RowScale(max_k |X(r, k)|) for X(r, k) = src[r * rs + k * ks]; zero-padded
// rows past `rows` get 1.
__global__ __launch_bounds__(256) void row_scale_kernel(int rows, int padded, int K, const float* __restrict__ src,
                                                        int64_t rs, int64_t ks, float* __restrict__ scale)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= padded) return;
    float m = 0.0f;
    if (r < rows)
        for (int k = 0; k < K; ++k) m = fmaxf(m, fabsf(src[r * rs + k * ks]));
    scale[r] = RowScale(m);
}

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

// Packs the `rows x K` operand X(r, k) = src[r * rs + k * ks] (times scale[r] when given)
// into fragment order: fragment f = rb * KB + kb holds, for lane l,
// X(rb*32 + l%32, kb*16 + (l/32)*8 + i), i < 8, as PARTS split vectors (part 0 = ET(x),
// part p = ET of what parts < p leave). Out-of-range elements are zero.
template <typename ET, int PARTS>
__global__ __launch_bounds__(256) void pack_frag_kernel(int rows, int K, const float* __restrict__ src, int64_t rs,
                                                        int64_t ks, int KB, int nfrag, const float* __restrict__ scale,
                                                        typename Mfma<ET>::vec* __restrict__ out)
{
    typedef typename Mfma<ET>::vec VT;
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int f = t >> 6, l = t & 63;
    if (f >= nfrag) return;
    const int rb = f / KB, kb = f % KB;
    const int r = rb * kFragRows + (l & 31);
    const int k0 = kb * kFragK + (l >> 5) * 8;
    const float sc = (scale != nullptr && r < rows) ? scale[r] : 1.0f;
    VT part[PARTS];
#pragma unroll
    for (int i = 0; i < 8; ++i)
    {
        const int k = k0 + i;
        float v = (r < rows && k < K) ? src[r * rs + k * ks] * sc : 0.0f;
#pragma unroll
        for (int p = 0; p < PARTS; ++p)
        {
            const ET h = static_cast<ET>(v);
            part[p][i] = h;
            v -= static_cast<float>(h);
        }
    }
#pragma unroll
    for (int p = 0; p < PARTS; ++p) out[(static_cast<size_t>(f) * PARTS + p) * 64 + l] = part[p];
}

constexpr int kXcds = 8;     // workgroups are dealt round-robin to the 8 XCDs (each its own L2)
constexpr int kGroupRows = 8;  // tile rows per L2 group: ~64 concurrent tiles per XCD form 8 x 8

// Tile of this workgroup, or false for a spare workgroup of the padded grid (grid:
// 8 * TilesPerXcd). With 8 or more tile columns, XCD x = blockIdx % 8 owns the tile columns
// [x * tilesN / 8, (x + 1) * tilesN / 8) and runs through them in groups of 8 tile rows,
// column by column, so that the ~64 tiles in flight on one XCD share 8 A strips and about
// 8 B strips in its L2 (instead of one B strip and every A strip). With fewer columns (a
// tall update) XCD x takes the x-th eighth of the tiles in column-major order.
__device__ inline bool TileOfXcd(int tilesM, int tilesN, int* tm, int* tn)
{
    const int x = blockIdx.x % kXcds, local = blockIdx.x / kXcds;
    if (tilesN < kXcds)
    {
        const int total = tilesM * tilesN;
        const int id = x * total / kXcds + local;
        if (id >= (x + 1) * total / kXcds) return false;
        *tm = id % tilesM;
        *tn = id / tilesM;
        return true;
    }
    const int c0 = x * tilesN / kXcds, c1 = (x + 1) * tilesN / kXcds, cols = c1 - c0;
    if (local >= tilesM * cols) return false;
    const int perGroup = kGroupRows * cols;  // tiles of a full group
    const int grp = local / perGroup, idx = local - grp * perGroup;
    const int rows = min(kGroupRows, tilesM - grp * kGroupRows);
    *tm = grp * kGroupRows + idx % rows;
    *tn = c0 + idx / rows;
    return true;
}

// Workgroups per XCD of the padded grid of TileOfXcd.
inline int TilesPerXcd(int tilesM, int tilesN)
{
    return tilesN < kXcds ? (tilesM * tilesN + kXcds - 1) / kXcds : tilesM * ((tilesN + kXcds - 1) / kXcds);
}

// One workgroup (4 waves) per 128 x 128 tile of C; wave (wm, wn) owns the 64 x 64
// quadrant rows [wm*64, +64), columns [wn*64, +64) as 2 x 2 accumulators of 32 x 32.
template <typename ET, int PARTS>
__global__ __launch_bounds__(256) void gemm_packed_kernel(int M, int N, int KB, const typename Mfma<ET>::vec* __restrict__ Ap,
                                                          const typename Mfma<ET>::vec* __restrict__ Bp,
                                                          float* __restrict__ C, int ldc, int tilesM, int tilesN,
                                                          const float* __restrict__ rowScale,
                                                          const float* __restrict__ colScale)
{
    typedef typename Mfma<ET>::vec VT;
    constexpr int SK = StageK<PARTS>();
    // [stage][block 0..3][kk][part][lane]
    __shared__ VT sA[2][4][SK][PARTS][64];
    __shared__ VT sB[2][4][SK][PARTS][64];
    int tm, tn;
    if (!TileOfXcd(tilesM, tilesN, &tm, &tn)) return;  // uniform: a spare workgroup
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
    VT ra[PER], rbv[PER];
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
            VT a[PARTS][2], bb[PARTS][2];
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
#define MF(X, Y) acc[i][j] = Mfma<ET>::run(bb[X][j], a[Y][i], acc[i][j])
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
    // Scaled operands: undo the power-of-two row (A) and column (B) factors exactly.
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        const int row = (rb0 + wm * 2 + i) * kFragRows + (lane & 31);
        if (row >= M) continue;
        const float ra_inv = rowScale != nullptr ? 1.0f / rowScale[row] : 1.0f;
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
                    const float prod = colScale != nullptr ? (acc[i][j][v] * ra_inv) * (1.0f / colScale[col]) : acc[i][j][v];
                    *p -= prod;
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
    return (p.aVecs + p.bVecs) * sizeof(bf16x8) + (static_cast<size_t>(p.rowBlocks) + p.colBlocks) * kFragRows * 4 +
           512;
}

hipError_t GemmSubPacked(int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
                         Precision prec, void* work, hipStream_t stream)
{
    if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
    if (prec != Precision::BF16 && prec != Precision::BF16X3 && prec != Precision::BF16X6 &&
        prec != Precision::FP16X3)
        return GemmSub(M, N, K, A, lda, 0, B, ldb, C, ldc, prec, stream);
    const int parts = prec == Precision::BF16X6 ? 3 : ((prec == Precision::BF16X3 || prec == Precision::FP16X3) ? 2 : 1);
    const PackLayout p = Layout(M, N, K, parts);
    char* base = static_cast<char*>(work);
    void* ap = base;
    void* bp = base + p.aVecs * sizeof(bf16x8);
    float* sa = reinterpret_cast<float*>(base + (p.aVecs + p.bVecs) * sizeof(bf16x8));
    float* sb = sa + static_cast<size_t>(p.rowBlocks) * kFragRows;
    const int fa = p.rowBlocks * p.KB, fb = p.colBlocks * p.KB;
    const int tilesM = p.rowBlocks / 4, tilesN = p.colBlocks / 4;
    const dim3 grid(kXcds * TilesPerXcd(tilesM, tilesN));
    const int64_t one = 1, la = lda, lb = ldb;
    // A(r, k) = A[k * lda + r]; B^T(c, k) = B[c * ldb + k]
#define PACK_AND_RUN(ET, P, SA, SB)                                                                                 \
    hipLaunchKernelGGL((pack_frag_kernel<ET, P>), dim3((fa * 64 + 255) / 256), dim3(256), 0, stream, M, K, A, one, la, \
                       p.KB, fa, SA, static_cast<Mfma<ET>::vec*>(ap));                                            \
    hipLaunchKernelGGL((pack_frag_kernel<ET, P>), dim3((fb * 64 + 255) / 256), dim3(256), 0, stream, N, K, B, lb, one, \
                       p.KB, fb, SB, static_cast<Mfma<ET>::vec*>(bp));                                            \
    {                                                                                                               \
        ProfileScope prof(ProfileKernel::TrailingGemm, 2.0 * M * N * K, stream);                                    \
        hipLaunchKernelGGL((gemm_packed_kernel<ET, P>), grid, dim3(256), 0, stream, M, N, p.KB,                     \
                           static_cast<const Mfma<ET>::vec*>(ap), static_cast<const Mfma<ET>::vec*>(bp), C, ldc,    \
                           tilesM, tilesN, SA, SB);                                                                 \
    }
    if (prec == Precision::FP16X3)
    {
        // FP16 has 5 exponent bits: scale every row of A and column of B by a power of two.
        const int pa = p.rowBlocks * kFragRows, pb = p.colBlocks * kFragRows;
        hipLaunchKernelGGL(row_scale_kernel, dim3((pa + 255) / 256), dim3(256), 0, stream, M, pa, K, A, one, la, sa);
        hipLaunchKernelGGL(row_scale_kernel, dim3((pb + 255) / 256), dim3(256), 0, stream, N, pb, K, B, lb, one, sb);
        PACK_AND_RUN(_Float16, 2, sa, sb);
    }
    else if (parts == 3)
    {
        PACK_AND_RUN(__bf16, 3, nullptr, nullptr);
    }
    else if (parts == 2)
    {
        PACK_AND_RUN(__bf16, 2, nullptr, nullptr);
    }
    else
    {
        PACK_AND_RUN(__bf16, 1, nullptr, nullptr);
    }
#undef PACK_AND_RUN
    return hipGetLastError();
}

}  // namespace ops
