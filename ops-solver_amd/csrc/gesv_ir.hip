// Mixed-precision dense solve with GMRES iterative refinement (EXTENSION; the reference
// has no refinement, its trailing GEMM is plain FP32: src/sgetrf/real_lu_blocked_kernel.cpp:1127).
//
// Argument convention: the reference's cuSOLVER-IRS style aclsolverSSgels
// (include/cann_ops_solver.h:815-862): dA, dB read-only, dX output, dWorkspace / lwork_bytes,
// host niter, device dinfo.
//
// Algorithm (per right-hand side):
//   factors: U^T A V = L U without pivoting after two levels of random butterflies (FP16 /
//   BF16 forms, n % 4 == 0), or P A = L U with partial pivoting (FP32, other orders, and
//   the fallback); trailing updates on MFMA in FP16 (per-tile power-of-two scaling), BF16
//   or split BF16, FP32 accumulation, FP32 storage of the factors;
//   x = M^-1 b;
//   repeat: r = b - A x in FP64 (A read in FP32); stop when
//     ||r||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= tol;
//     d = GMRES(M^-1 A, M^-1 r) with M^-1 applied by FP64 substitution on the FP32
//     factors and the Krylov basis, the Hessenberg least squares and the update in FP64;
//     x += d.
// Every reduction runs in a fixed order, so results are bitwise repeatable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "lu_kernels.h"
#include "trsv.h"

namespace ops
{
hipError_t RealLuSolveInPlace(int trans, int n, int nrhs, const float* lu, int lda, float* x, int ldx,
                              hipStream_t stream);
hipError_t RbtMakeDiagonals(int n, unsigned long long seed, float* du, float* dv, hipStream_t s);
hipError_t RbtTransform(int n, const float* A, int lda, float* dst, int ldd, const float* du, const float* dv,
                        hipStream_t s);
hipError_t RbtApplyVecLeft(int n, const double* x, double* y, const float* du, hipStream_t s);
hipError_t RbtApplyVecRight(int n, const double* x, double* y, const float* dv, hipStream_t s);
hipError_t GetrfNoPivot(int n, float* a, int lda, int* info, Precision trailing, void* pack, hipStream_t s);
int GenpOuterBlock();

namespace
{
constexpr int RB = 256;  // rows per workgroup in the GEMV
constexpr int CB = 256;  // columns per chunk in the GEMV

// partial[chunk * n + i] = sum_{j in chunk} f(A(i, j)) * x_j   (FP64 accumulation)
template <bool ABS>
__global__ __launch_bounds__(RB) void gemv_partial_kernel(int n, const float* __restrict__ A, int lda,
                                                          const double* __restrict__ x, double* __restrict__ partial)
{
    const int i = blockIdx.x * RB + threadIdx.x;
    const int chunk = blockIdx.y;
    const int j0 = chunk * CB, j1 = min(n, j0 + CB);
    __shared__ double s_x[CB];
    for (int j = j0 + threadIdx.x; j < j1; j += RB) s_x[j - j0] = ABS ? 1.0 : x[j];
    __syncthreads();
    if (i >= n) return;
    double acc = 0.0;
    const float* col = A + static_cast<size_t>(j0) * lda + i;
    for (int j = j0; j < j1; ++j, col += lda)
    {
        const double v = static_cast<double>(*col);
        acc = fma(ABS ? fabs(v) : v, s_x[j - j0], acc);
    }
    partial[static_cast<size_t>(chunk) * n + i] = acc;
}

// out_i = b_i - sum_chunks partial (or the plain sum when b == nullptr)
__global__ void gemv_reduce_kernel(int n, int chunks, const double* __restrict__ partial, const double* __restrict__ b,
                                   double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < chunks; ++c) s += partial[static_cast<size_t>(c) * n + i];
    out[i] = b ? b[i] - s : s;
}

// Maximum that keeps NaN (fmax drops it): a NaN anywhere in the vector makes the norm NaN,
// so the refinement sees a non-finite backward error instead of a spurious zero.
__device__ inline double NanMax(double a, double b) { return (a != a || b != b) ? __builtin_nan("") : fmax(a, b); }

// Single-workgroup reductions into out[slot] (fixed tree order).
template <int OP>  // 0: max |v| (NaN-propagating), 1: sum v^2
__global__ __launch_bounds__(1024) void reduce_kernel(int n, const double* __restrict__ v, double* __restrict__ out,
                                                      int slot)
{
    __shared__ double s[1024];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024)
    {
        const double x = v[i];
        acc = OP == 0 ? NanMax(acc, fabs(x)) : fma(x, x, acc);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1)
    {
        if (threadIdx.x < w)
            s[threadIdx.x] = OP == 0 ? NanMax(s[threadIdx.x], s[threadIdx.x + w]) : s[threadIdx.x] + s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[slot] = OP == 0 ? s[0] : sqrt(s[0]);
}

// dots[q] = V(:, q) . w for q < cols (one workgroup per column)
__global__ __launch_bounds__(256) void dots_kernel(int n, const double* __restrict__ V, int ldv,
                                                   const double* __restrict__ w, double* __restrict__ dots)
{
    __shared__ double s[256];
    const double* v = V + static_cast<size_t>(blockIdx.x) * ldv;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc = fma(v[i], w[i], acc);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1)
    {
        if (threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) dots[blockIdx.x] = s[0];
}

// w -= V(:, 0:cols) * h ; optionally h_acc += h
__global__ void orth_update_kernel(int n, int cols, const double* __restrict__ V, int ldv,
                                   const double* __restrict__ h, double* __restrict__ w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = w[i];
    for (int q = 0; q < cols; ++q) s = fma(-V[static_cast<size_t>(q) * ldv + i], h[q], s);
    w[i] = s;
}

__global__ void scale_copy_kernel(int n, const double* __restrict__ src, const double* __restrict__ scal, int slot,
                                  bool invert, double* __restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = scal[slot];
    dst[i] = invert ? (s != 0.0 ? src[i] / s : 0.0) : src[i] * s;
}

// out[i] = v[perm[i]] (row interchanges of the pivoted factorization)
__global__ void gather_f64_kernel(int n, const double* __restrict__ v, const int* __restrict__ perm,
                                  double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v[perm[i]];
}

// x += V(:, 0:k) * y
__global__ void correct_kernel(int n, int k, const double* __restrict__ V, int ldv, const double* __restrict__ y,
                               double* __restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = x[i];
    for (int q = 0; q < k; ++q) s = fma(V[static_cast<size_t>(q) * ldv + i], y[q], s);
    x[i] = s;
}

__global__ void identity_ipiv_kernel(int n, int* __restrict__ ipiv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ipiv[i] = i + 1;
}

__global__ void f32_to_f64_kernel(int n, const float* __restrict__ a, int64_t stride, double* __restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = static_cast<double>(a[i * stride]);
}

__global__ void f64_to_f32_kernel(int n, const double* __restrict__ a, float* __restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = static_cast<float>(a[i]);
}

__global__ void copy_cm_kernel(int rows, int cols, const float* __restrict__ src, int lds, float* __restrict__ dst,
                               int ldd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i < rows) dst[static_cast<size_t>(c) * ldd + i] = src[static_cast<size_t>(c) * lds + i];
}

inline dim3 G1(int n) { return dim3((n + 255) / 256); }

struct IrWork
{
    float* lu;
    int* perm;
    void* getrfWork;
    double *x, *b, *r, *w, *V, *partial, *scal, *small, *t64;
    float *du, *dv;  // random butterfly diagonals (2n each)
    void* trsv;      // chained triangular-solve state
    void* pack;      // packed BF16 operands of the trailing update
    size_t bytes;
};

size_t Align256(size_t v) { return (v + 255) / 256 * 256; }

IrWork IrLayout(void* base, int n, int restart)
{
    IrWork w{};
    char* p = static_cast<char*>(base);
    size_t off = 0;
    auto take = [&](size_t bytes) -> char*
    {
        char* q = p ? p + off : nullptr;
        off += Align256(bytes);
        return q;
    };
    const int chunks = (n + CB - 1) / CB;
    w.lu = reinterpret_cast<float*>(take(static_cast<size_t>(n) * n * sizeof(float)));
    w.perm = reinterpret_cast<int*>(take(static_cast<size_t>(n) * sizeof(int)));
    w.getrfWork = take(RealGetrfWorkspaceBytes(n, n));
    w.x = reinterpret_cast<double*>(take(static_cast<size_t>(n) * sizeof(double)));
    w.b = reinterpret_cast<double*>(take(static_cast<size_t>(n) * sizeof(double)));
    w.r = reinterpret_cast<double*>(take(static_cast<size_t>(n) * sizeof(double)));
    w.w = reinterpret_cast<double*>(take(static_cast<size_t>(n) * sizeof(double)));
    w.V = reinterpret_cast<double*>(take(static_cast<size_t>(n) * (restart + 1) * sizeof(double)));
    w.partial = reinterpret_cast<double*>(take(static_cast<size_t>(chunks) * n * sizeof(double)));
    w.scal = reinterpret_cast<double*>(take(64 * sizeof(double)));
    w.small = reinterpret_cast<double*>(take(static_cast<size_t>(restart + 2) * 4 * sizeof(double)));
    w.t64 = reinterpret_cast<double*>(take(static_cast<size_t>(n) * sizeof(double)));
    w.du = reinterpret_cast<float*>(take(static_cast<size_t>(2 * n) * sizeof(float)));
    w.dv = reinterpret_cast<float*>(take(static_cast<size_t>(2 * n) * sizeof(float)));
    w.trsv = take(TrsvChainStateBytes(n));
    w.pack = take(GemmPackedWorkspaceBytes(n, n, std::max(512, GenpOuterBlock())));
    w.bytes = off;
    return w;
}

// scal slots
enum
{
    S_ANORM = 0,
    S_RNORM = 1,
    S_XNORM = 2,
    S_BNORM = 3,
    S_BETA = 4,
    S_HNEXT = 5
};

}  // namespace

size_t GesvIrWorkspaceBytes(int n, int restart) { return IrLayout(nullptr, n, restart).bytes + 4096; }

// Solves A X = B. prec: FP32 (plain, no refinement), FP16 or BF16 (GMRES-IR).
// Returns hipSuccess; *niterOut = total GMRES iterations (negative when refinement did not
// reach tol within maxOuter outer steps).
hipError_t GesvIR(int n, int nrhs, const float* A, int lda, int* ipiv, const float* B, int ldb, float* X, int ldx,
                  void* work, Precision prec, double tol, int maxOuter, int restart, bool butterfly, bool tournament,
                  int* info,
                  int* niterOut, hipStream_t stream, const LookAhead* la)
{
    IrWork w = IrLayout(work, n, restart);
    // OPS_SOLVER_TRACE=1 prints one line per refinement step to stderr (the loop already
    // synchronizes there, so tracing adds no extra waits).
    static const bool trace = std::getenv("OPS_SOLVER_TRACE") != nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };
    const int chunks = (n + CB - 1) / CB;
    // Partial pivoting by default. With `butterfly` the mixed precisions factor U^T A V
    // without pivoting (random butterflies, n % 4 == 0) and fall back to pivoting after a
    // zero pivot or a refinement failure.
    bool useRbt = butterfly && prec != Precision::FP32 && n % 4 == 0 && n >= 8;
    hipError_t e = hipSuccess;
    auto factor = [&]() -> hipError_t
    {
        if (useRbt)
        {
            hipError_t err = RbtMakeDiagonals(n, 0x5EED0000ull + static_cast<unsigned long long>(n), w.du, w.dv, stream);
            if (err == hipSuccess) err = RbtTransform(n, A, lda, w.lu, n, w.du, w.dv, stream);
            if (err == hipSuccess) err = GetrfNoPivot(n, w.lu, n, info, prec, w.pack, stream);
            hipLaunchKernelGGL(identity_ipiv_kernel, G1(n), dim3(256), 0, stream, n, ipiv);
            return err == hipSuccess ? hipGetLastError() : err;
        }
        hipLaunchKernelGGL(copy_cm_kernel, dim3((n + 255) / 256, n), dim3(256), 0, stream, n, n, A, lda, w.lu, n);
        return RealGetrfPerm(n, n, w.lu, n, ipiv, info, w.getrfWork, prec, w.perm, stream, w.pack, la, tournament);
    };
    TrsvChainState ts = TrsvChainLayout(w.trsv, n);
    if ((e = TrsvChainReset(ts, stream)) != hipSuccess) return e;
    if ((e = factor()) != hipSuccess) return e;
    if (trace)
    {
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
        std::fprintf(stderr, "[gesv-ir] n=%d factor (%s, precision %d) t=%.2f ms\n", n, useRbt ? "rbt" : "pivoted",
                     static_cast<int>(prec), elapsed());
        std::fflush(stderr);
    }
    const dim3 gemvGrid((n + RB - 1) / RB, chunks);
    auto solveM = [&](const double* v, double* out) -> hipError_t
    {
        // out = M^-1 v, M = P^T L U (pivoted) or U^-T L U V^-1 (butterflies). The two
        // substitutions run in FP64 on the FP32 factors: GMRES-IR only converges at large
        // condition numbers when the preconditioned operator is applied at a precision
        // beyond that of the factors (Carson & Higham, SIAM J. Sci. Comput. 40(2), 2018).
        double* t = useRbt ? w.t64 : out;
        hipError_t err;
        if (useRbt)
            err = RbtApplyVecLeft(n, v, t, w.du, stream);
        else
        {
            hipLaunchKernelGGL(gather_f64_kernel, G1(n), dim3(256), 0, stream, n, v, w.perm, t);
            err = hipGetLastError();
        }
        if (err == hipSuccess) err = TrsvChain(ts, false, w.lu, n, t, stream);
        if (err == hipSuccess) err = TrsvChain(ts, true, w.lu, n, t, stream);
        if (err == hipSuccess && useRbt) err = RbtApplyVecRight(n, t, out, w.dv, stream);
        return err;
    };
    auto residual = [&](const double* x, double* r)
    {
        hipLaunchKernelGGL(gemv_partial_kernel<false>, gemvGrid, dim3(RB), 0, stream, n, A, lda, x, w.partial);
        hipLaunchKernelGGL(gemv_reduce_kernel, G1(n), dim3(256), 0, stream, n, chunks, w.partial, w.b, r);
    };
    auto matvec = [&](const double* x, double* y)
    {
        hipLaunchKernelGGL(gemv_partial_kernel<false>, gemvGrid, dim3(RB), 0, stream, n, A, lda, x, w.partial);
        hipLaunchKernelGGL(gemv_reduce_kernel, G1(n), dim3(256), 0, stream, n, chunks, w.partial, nullptr, y);
    };
    // ||A||_inf: row sums of |A|
    double* rowsum = w.r;
    hipLaunchKernelGGL(gemv_partial_kernel<true>, gemvGrid, dim3(RB), 0, stream, n, A, lda, w.x, w.partial);
    hipLaunchKernelGGL(gemv_reduce_kernel, G1(n), dim3(256), 0, stream, n, chunks, w.partial, nullptr, rowsum);
    hipLaunchKernelGGL(reduce_kernel<0>, dim3(1), dim3(1024), 0, stream, n, rowsum, w.scal, S_ANORM);

    int totalIters = 0;
    bool converged = true;
    std::vector<double> hscal(8);
    for (int c = 0; c < nrhs; ++c)
    {
        hipLaunchKernelGGL(f32_to_f64_kernel, G1(n), dim3(256), 0, stream, n, B + static_cast<size_t>(c) * ldb, 1,
                           w.b);
        hipLaunchKernelGGL(reduce_kernel<0>, dim3(1), dim3(1024), 0, stream, n, w.b, w.scal, S_BNORM);
    restart_rhs:
        if ((e = solveM(w.b, w.x)) != hipSuccess) return e;
        bool done = prec == Precision::FP32;
        int outer = 0;
        if (useRbt)
        {
            // A zero pivot without pivoting means the butterflies did not help: pivot instead.
            int hinfo = 0;
            if ((e = hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, stream)) != hipSuccess ||
                (e = hipStreamSynchronize(stream)) != hipSuccess)
                return e;
            if (hinfo != 0)
            {
                useRbt = false;
                if ((e = factor()) != hipSuccess) return e;
                goto restart_rhs;
            }
        }
        while (!done)
        {
            residual(w.x, w.r);
            hipLaunchKernelGGL(reduce_kernel<0>, dim3(1), dim3(1024), 0, stream, n, w.r, w.scal, S_RNORM);
            hipLaunchKernelGGL(reduce_kernel<0>, dim3(1), dim3(1024), 0, stream, n, w.x, w.scal, S_XNORM);
            int chainErr = 0, leafErr = 0;
            int* leafErrWord = useRbt ? nullptr : RealGetrfErrorWord(w.getrfWork, n);
            if ((e = hipMemcpyAsync(hscal.data(), w.scal, 8 * sizeof(double), hipMemcpyDeviceToHost, stream)) !=
                    hipSuccess ||
                (e = hipMemcpyAsync(&chainErr, ts.err, sizeof(int), hipMemcpyDeviceToHost, stream)) != hipSuccess ||
                (leafErrWord != nullptr &&
                 (e = hipMemcpyAsync(&leafErr, leafErrWord, sizeof(int), hipMemcpyDeviceToHost, stream)) !=
                     hipSuccess) ||
                (e = hipStreamSynchronize(stream)) != hipSuccess)
                return e;
            if (chainErr != 0 || leafErr != 0) return hipErrorLaunchTimeOut;
            const double berr = hscal[S_RNORM] / (hscal[S_ANORM] * hscal[S_XNORM] + hscal[S_BNORM]);
            if (trace)
            {
                std::fprintf(stderr, "[gesv-ir] n=%d rhs=%d %s outer=%d berr=%.3e gmres=%d t=%.2f ms\n", n, c,
                             useRbt ? "rbt" : "pivoted", outer, berr, totalIters, elapsed());
                std::fflush(stderr);
            }
            if (!std::isfinite(berr))
            {
                // Overflow in the factors (e.g. a tiny unpivoted pivot) or non-finite data:
                // retry with partial pivoting once, otherwise report non-convergence.
                if (useRbt)
                {
                    useRbt = false;
                    if ((e = factor()) != hipSuccess) return e;
                    goto restart_rhs;
                }
                converged = false;
                break;
            }
            if (berr <= tol) break;
            if (outer++ >= maxOuter)
            {
                if (useRbt)  // refinement stalled on the unpivoted factors: refactor with pivoting
                {
                    useRbt = false;
                    if ((e = factor()) != hipSuccess) return e;
                    goto restart_rhs;
                }
                converged = false;
                break;
            }
            // GMRES(restart) on M^-1 A d = M^-1 r, d0 = 0.
            if ((e = solveM(w.r, w.w)) != hipSuccess) return e;
            hipLaunchKernelGGL(reduce_kernel<1>, dim3(1), dim3(1024), 0, stream, n, w.w, w.scal, S_BETA);
            hipLaunchKernelGGL(scale_copy_kernel, G1(n), dim3(256), 0, stream, n, w.w, w.scal, S_BETA, true, w.V);
            if ((e = hipMemcpyAsync(hscal.data(), w.scal, 8 * sizeof(double), hipMemcpyDeviceToHost, stream)) !=
                    hipSuccess ||
                (e = hipStreamSynchronize(stream)) != hipSuccess)
                return e;
            const double beta = hscal[S_BETA];
            if (!(beta > 0.0)) break;
            std::vector<double> H(static_cast<size_t>(restart + 1) * restart, 0.0), cs(restart), sn(restart),
                g(restart + 1, 0.0), h(restart + 1);
            g[0] = beta;
            int k = 0;
            for (int j = 0; j < restart; ++j)
            {
                double* vj = w.V + static_cast<size_t>(j) * n;
                matvec(vj, w.r);
                if ((e = solveM(w.r, w.w)) != hipSuccess) return e;
                // classical Gram-Schmidt, twice
                double* hdev = w.small;
                for (int pass = 0; pass < 2; ++pass)
                {
                    hipLaunchKernelGGL(dots_kernel, dim3(j + 1), dim3(256), 0, stream, n, w.V, n, w.w,
                                       hdev + pass * (restart + 2));
                    hipLaunchKernelGGL(orth_update_kernel, G1(n), dim3(256), 0, stream, n, j + 1, w.V, n,
                                       hdev + pass * (restart + 2), w.w);
                }
                hipLaunchKernelGGL(reduce_kernel<1>, dim3(1), dim3(1024), 0, stream, n, w.w, w.scal, S_HNEXT);
                std::vector<double> hh(2 * (restart + 2));
                if ((e = hipMemcpyAsync(hh.data(), hdev, hh.size() * sizeof(double), hipMemcpyDeviceToHost,
                                        stream)) != hipSuccess ||
                    (e = hipMemcpyAsync(hscal.data(), w.scal, 8 * sizeof(double), hipMemcpyDeviceToHost, stream)) !=
                        hipSuccess ||
                    (e = hipStreamSynchronize(stream)) != hipSuccess)
                    return e;
                for (int q = 0; q <= j; ++q) h[q] = hh[q] + hh[(restart + 2) + q];
                h[j + 1] = hscal[S_HNEXT];
                if (h[j + 1] > 0.0)
                    hipLaunchKernelGGL(scale_copy_kernel, G1(n), dim3(256), 0, stream, n, w.w, w.scal, S_HNEXT, true,
                                       w.V + static_cast<size_t>(j + 1) * n);
                for (int q = 0; q < j; ++q)  // apply previous Givens rotations
                {
                    const double t = cs[q] * h[q] + sn[q] * h[q + 1];
                    h[q + 1] = -sn[q] * h[q] + cs[q] * h[q + 1];
                    h[q] = t;
                }
                const double den = std::hypot(h[j], h[j + 1]);
                cs[j] = den > 0 ? h[j] / den : 1.0;
                sn[j] = den > 0 ? h[j + 1] / den : 0.0;
                h[j] = den;
                h[j + 1] = 0.0;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                for (int q = 0; q <= j; ++q) H[static_cast<size_t>(q) * restart + j] = h[q];
                k = j + 1;
                ++totalIters;
                if (std::fabs(g[j + 1]) <= 1e-4 * beta || !(hscal[S_HNEXT] > 0.0)) break;
            }
            if (trace)
            {
                std::fprintf(stderr, "[gesv-ir]   gmres cycle: %d iterations, beta %.3e, |g|/beta %.3e, h(1,0) %.3e\n", k,
                             beta, k > 0 ? std::fabs(g[k]) / beta : 1.0, restart > 0 ? H[0] : 0.0);
                std::fflush(stderr);
            }
            // back substitution H y = g
            std::vector<double> y(k, 0.0);
            for (int q = k - 1; q >= 0; --q)
            {
                double s = g[q];
                for (int t = q + 1; t < k; ++t) s -= H[static_cast<size_t>(q) * restart + t] * y[t];
                y[q] = s / H[static_cast<size_t>(q) * restart + q];
            }
            if ((e = hipMemcpyAsync(w.small, y.data(), k * sizeof(double), hipMemcpyHostToDevice, stream)) !=
                hipSuccess)
                return e;
            hipLaunchKernelGGL(correct_kernel, G1(n), dim3(256), 0, stream, n, k, w.V, n, w.small, w.x);
            if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;  // y is host memory
        }
        hipLaunchKernelGGL(f64_to_f32_kernel, G1(n), dim3(256), 0, stream, n, w.x, X + static_cast<size_t>(c) * ldx);
    }
    // A bounded wait that gave up (factor leaves or chained solves) reaches dinfo too; the
    // refinement loop has already turned it into an error for the refined precisions.
    if (!useRbt) FoldWaitError(RealGetrfErrorWord(w.getrfWork, n), info, stream);
    FoldWaitError(ts.err, info, stream);
    if (niterOut) *niterOut = converged ? totalIters : -totalIters - 1;
    return hipGetLastError();
}

}  // namespace ops
