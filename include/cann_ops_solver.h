/*
 * C ABI of the MI355X (gfx950) LU solver library (libops_solver_hip.so).
 *
 * Drop-in for the LU family of the reference's include/cann_ops_solver.h: the same entry
 * point names, argument order, memory domains, layouts, 1-based pivots, info and status
 * conventions. Each declaration cites the reference declaration it replaces. Entry points
 * marked EXTENSION have no counterpart in the reference.
 *
 * Conventions (reference include/cann_ops_solver.h:82-99):
 *  - getrf / getri / batched entry points take Device pointers; matrices are column-major
 *    with element (r, c) at A[c * lda + r]; ipiv is 1-based; devInfo is a Device int.
 *  - info = 0 on success, -i for an illegal argument i (handle excluded), or the first
 *    1-based column with an exact zero on U's diagonal (the factorization still completes).
 *  - Lwork counts are elements (float or complex), not bytes.
 *  - aclsolverSgetrs / aclsolverCgetrs take HOST buffers in ROW-major layout and stage them
 *    (reference :980-1043); the Device column-major solves are EXTENSIONS.
 *  - Work is asynchronous on the handle's stream; validation steps that inspect Device data
 *    synchronize that stream, as the reference does.
 *  - Size limits are extended beyond the reference's (getrf m <= 16384, getrs / getri
 *    n <= 16384, matinv n <= 64, cgetrf m <= 16384); larger sizes get INVALID_VALUE with
 *    devInfo = -(argument index). aclsolverSetStrictLimits(handle, 1) restores the
 *    reference limits.
 *  - Every compute entry point answers ACLSOLVER_STATUS_ARCH_MISMATCH when the handle was
 *    created on a device other than gfx950 (the reference does so off its target,
 *    src/utils/complex_solver_host.h:50-53).
 *  - devInfo = ACLSOLVER_INFO_WAIT_TIMEOUT (EXTENSION) reports that a bounded
 *    cross-workgroup wait of the factorization or solve gave up; the results are then NaN.
 *    The call itself returns SUCCESS (the condition arises on the device after the launch).
 */
#pragma once
#include "cann_ops_solver_common.h"

#define ACLSOLVER_INFO_WAIT_TIMEOUT (-1000)
#define ACLSOLVER_DEFAULT_SPIN_LIMIT (1u << 20)

#ifdef __cplusplus
extern "C"
{
#endif

    /* ---- handle (reference include/cann_ops_solver.h:50-80) ---- */
    aclsolverStatus_t aclsolverCreate(aclsolverHandle_t* handle);                        /* :50 */
    aclsolverStatus_t aclsolverDestroy(aclsolverHandle_t handle);                        /* :61 */
    aclsolverStatus_t aclsolverSetStream(aclsolverHandle_t handle, aclrtStream stream);  /* :69 */
    aclsolverStatus_t aclsolverGetStream(aclsolverHandle_t handle, aclrtStream* stream); /* :79 */

    /* EXTENSION: 1 = enforce the reference's size limits (m <= 8192 for sgetrf, n <= 32
     * for matinv, m <= 4096 / n <= 8192 for cgetrf, n, nrhs <= 8192 for getrs). */
    aclsolverStatus_t aclsolverSetStrictLimits(aclsolverHandle_t handle, int strict);

    /* ---- real LU (reference include/cann_ops_solver.h:100-111) ---- */
    aclsolverStatus_t aclsolverSgetrf_bufferSize(aclsolverHandle_t handle, int m, int n, float* A, int lda,
                                                 int* Lwork); /* :100 */
    aclsolverStatus_t aclsolverSgetrf(aclsolverHandle_t handle, int m, int n, float* A, int lda, float* Workspace,
                                      int* devIpiv, int* devInfo); /* :101-102 */
    aclsolverStatus_t aclsolverSgetri_bufferSize(aclsolverHandle_t handle, int n, float* A, int lda,
                                                 int* Lwork); /* :103 */
    aclsolverStatus_t aclsolverSgetri(aclsolverHandle_t handle, int n, float* A, int lda, int* devIpiv,
                                      float* Workspace, int Lwork, int* devInfo); /* :104-105 */
    aclsolverStatus_t aclsolverSgetriBatched(aclsolverHandle_t handle, int n, const float* const Aarray[], int lda,
                                             const int* PivotArray, float* const Carray[], int ldc, int* infoArray,
                                             int batchSize); /* :106-108 */
    aclsolverStatus_t aclsolverSmatinvBatched(aclsolverHandle_t handle, int n, const float* const A[], int lda,
                                              float* const Ainv[], int lda_inv, int* info,
                                              int batchSize); /* :109-110 */

    /* ---- complex LU (reference include/cann_ops_solver.h:882-923) ---- */
    aclsolverStatus_t aclsolverCgetrf_bufferSize(aclsolverHandle_t handle, int m, int n, aclFloatComplex* A, int lda,
                                                 int* Lwork); /* :882-883 */
    aclsolverStatus_t aclsolverCgetrf(aclsolverHandle_t handle, int m, int n, aclFloatComplex* A, int lda,
                                      aclFloatComplex* Workspace, int* devIpiv, int* devInfo); /* :890-891 */
    aclsolverStatus_t aclsolverCgetri_bufferSize(aclsolverHandle_t handle, int n, aclFloatComplex* A, int lda,
                                                 int* Lwork); /* :896-897 */
    aclsolverStatus_t aclsolverCgetri(aclsolverHandle_t handle, int n, aclFloatComplex* A, int lda, int* devIpiv,
                                      aclFloatComplex* Workspace, int Lwork, int* devInfo); /* :903-904 */
    aclsolverStatus_t aclsolverCgetriBatched(aclsolverHandle_t handle, int n, const aclFloatComplex* const Aarray[],
                                             int lda, const int* PivotArray, aclFloatComplex* const Carray[], int ldc,
                                             int* infoArray, int batchSize); /* :911-913 */
    aclsolverStatus_t aclsolverCmatinvBatched(aclsolverHandle_t handle, int n, const aclFloatComplex* const A[],
                                              int lda, aclFloatComplex* const Ainv[], int lda_inv, int* info,
                                              int batchSize); /* :919-921 */

    /* ---- host-staged row-major solves (reference include/cann_ops_solver.h:1011-1043) ---- */
    aclError aclsolverSgetrs(aclsolverHandle_t handle, const uint64_t trans, const int64_t n, const int64_t nrhs,
                             float* A, const int64_t lda, int32_t* ipiv, float* B, const int64_t ldb,
                             int32_t* info); /* :1011-1012 */
    aclError aclsolverCgetrs(aclsolverHandle_t handle, const uint64_t trans, const int64_t n, const int64_t nrhs,
                             aclFloatComplex* A, const int64_t lda, int32_t* ipiv, aclFloatComplex* B,
                             const int64_t ldb, int32_t* info); /* :1040-1042 */

    /* ---- EXTENSION: Device column-major solves from getrf factors ----
     * trans 0: A X = B; 1: A^T X = B; 2 (complex only): A^H X = B. B is overwritten by X.
     * devIpiv may be null (factors without pivoting). devInfo receives 0 or -i. */
    aclsolverStatus_t aclsolverSgetrsDevice(aclsolverHandle_t handle, int trans, int n, int nrhs, const float* dA,
                                            int lda, const int* devIpiv, float* dB, int ldb, int* devInfo);
    aclsolverStatus_t aclsolverCgetrsDevice(aclsolverHandle_t handle, int trans, int n, int nrhs,
                                            const aclFloatComplex* dA, int lda, const int* devIpiv,
                                            aclFloatComplex* dB, int ldb, int* devInfo);

    /* ---- EXTENSION: LU with tournament pivoting (communication-avoiding LU) ----
     * Same arguments, workspace (aclsolverSgetrf_bufferSize), checks and outputs as
     * aclsolverSgetrf, but panels with more than 1024 active rows choose the pivots of each
     * 32 columns by a reduction tree over 256-row blocks (partial pivoting inside every
     * block and tree node) instead of partial pivoting over the whole column. P A = L U
     * holds with ipiv as the interchange sequence; the row order differs from partial
     * pivoting's, |L| may exceed 1 slightly. A null devIpiv factors without pivoting. */
    aclsolverStatus_t aclsolverSgetrfTournament(aclsolverHandle_t handle, int m, int n, float* A, int lda,
                                                float* Workspace, int* devIpiv, int* devInfo);

    /* ---- EXTENSION without a counterpart in the reference: a solve that allocates nothing ----
     * For callers that cannot allocate device memory (code under stream/graph capture, tight
     * serving loops, harnesses that forbid hipMalloc): the same computation and results as
     * aclsolverSgetrsDevice, with its staging (4*n*nrhs + 20*n bytes plus alignment) in a
     * fixed per-GPU device array of ACLSOLVER_NOALLOC_SCRATCH_BYTES instead of handle scratch.
     * Larger problems return INVALID_VALUE with devInfo = -2 (n alone too large) or -3
     * (n*nrhs too large). Every call on one GPU shares the array: calls that overlap on one
     * GPU (two streams or threads) race on it. */
#define ACLSOLVER_NOALLOC_SCRATCH_BYTES (8u << 20)
    aclsolverStatus_t aclsolverSgetrsDeviceNoAlloc(aclsolverHandle_t handle, int trans, int n, int nrhs,
                                                   const float* dA, int lda, const int* devIpiv, float* dB, int ldb,
                                                   int* devInfo);

    /* ---- EXTENSION: mixed-precision dense solve with GMRES iterative refinement ----
     * Argument convention of the reference's aclsolverSSgels (include/cann_ops_solver.h:
     * 815-862, the cuSOLVER IRS form). A is FP32 column-major and is NOT modified; the
     * factors live in dWorkspace. The trailing-matrix updates of the LU run on MFMA with
     * FP16 (SH: each FP32 operand split into two FP16 parts after a power-of-two row /
     * column scaling, three products) or BF16 (SB: split into three BF16 parts, six
     * products, see aclsolverSetBf16Split) inputs and FP32 accumulation; SS factors in plain
     * FP32 with partial pivoting; SH/SB pivot partially too unless tournament pivoting is
     * selected (aclsolverSetTournamentPivoting; aclsolverSetButterfly for the unpivoted form).
     * Refinement:
     * GMRES preconditioned by the LU factors (applied by FP64 substitution), the
     * residual r = b - A x and the correction accumulation in FP64, until the backward
     * error ||r||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= the handle's tolerance
     * (default 1e-6) or maxiter outer steps. niter (host) receives the total number of
     * GMRES iterations (negative: refinement did not converge and the FP32 LU solution
     * was returned). dinfo (device) gets 0, -i or the first zero pivot. */
    aclsolverStatus_t aclsolverSHgesv_bufferSize(aclsolverHandle_t handle, int n, int nrhs, size_t* lwork_bytes);
    aclsolverStatus_t aclsolverSBgesv_bufferSize(aclsolverHandle_t handle, int n, int nrhs, size_t* lwork_bytes);
    aclsolverStatus_t aclsolverSSgesv_bufferSize(aclsolverHandle_t handle, int n, int nrhs, size_t* lwork_bytes);
    aclsolverStatus_t aclsolverSHgesv(aclsolverHandle_t handle, int n, int nrhs, float* dA, int ldda, int* dipiv,
                                      float* dB, int lddb, float* dX, int lddx, void* dWorkspace, size_t lwork_bytes,
                                      int* niter, int* dinfo);
    aclsolverStatus_t aclsolverSBgesv(aclsolverHandle_t handle, int n, int nrhs, float* dA, int ldda, int* dipiv,
                                      float* dB, int lddb, float* dX, int lddx, void* dWorkspace, size_t lwork_bytes,
                                      int* niter, int* dinfo);
    aclsolverStatus_t aclsolverSSgesv(aclsolverHandle_t handle, int n, int nrhs, float* dA, int ldda, int* dipiv,
                                      float* dB, int lddb, float* dX, int lddx, void* dWorkspace, size_t lwork_bytes,
                                      int* niter, int* dinfo);
    /* EXTENSION: refinement controls for the gesv family (tolerance on the backward error,
     * maximum outer refinement steps, GMRES restart length). */
    aclsolverStatus_t aclsolverSetRefinement(aclsolverHandle_t handle, double tolerance, int maxOuter, int restart);
    /* EXTENSION: operand split of SBgesv's trailing updates on bf16 MFMA (FP32 accumulation).
     * 2 (default): x = h + m + l in bf16, six products (FP32-level products);
     * 1: x = h + m, products h*h + h*m + m*h (~16 bits); 0: single bf16 operands.
     * Returns INVALID_VALUE outside 0..2. */
    aclsolverStatus_t aclsolverSetBf16Split(aclsolverHandle_t handle, int split);
    /* EXTENSION: 1 = SH/SB factor U^T A V without pivoting after two levels of random
     * butterfly transforms (n % 4 == 0; partial pivoting after a zero pivot or a refinement
     * failure); 0 (default) = partial pivoting. */
    aclsolverStatus_t aclsolverSetButterfly(aclsolverHandle_t handle, int on);
    /* EXTENSION: pivoting of SH/SB's tall panel leaves. 1: tournament pivoting
     * (communication-avoiding LU: the pivots of 32 columns are chosen by a reduction tree over
     * 256-row blocks, so the row order differs from partial pivoting's); 0 (default): partial
     * pivoting, as sgetrf. SS always uses partial pivoting. */
    aclsolverStatus_t aclsolverSetTournamentPivoting(aclsolverHandle_t handle, int on);

    /* ---- EXTENSION: uneven batched LU solve (getrf + getrs) in one persistent launch ----
     * nArray (HOST, batchSize entries) holds each order n_i; dAarray / dBarray are Device
     * arrays of Device pointers to column-major A_i (n_i x n_i, lda = n_i) and B_i
     * (n_i x nrhs, ldb = n_i); A_i is overwritten by its packed LU, B_i by X_i. dIpiv holds
     * sum(n_i) pivots, matrix i's at offset sum_{j<i} n_j. dInfo gets one info per matrix.
     * A plan bins the matrices by size and holds the tuned per-bin schedule; executing a
     * plan allocates nothing, synchronizes nothing and may be captured into a hipGraph. */
    typedef void* aclsolverVbatchedPlan_t;
    aclsolverStatus_t aclsolverSgesvVbatched_createPlan(aclsolverHandle_t handle, int batchSize, const int* nArray,
                                                        int nrhs, aclsolverVbatchedPlan_t* plan);
    aclsolverStatus_t aclsolverSgesvVbatched_execute(aclsolverHandle_t handle, aclsolverVbatchedPlan_t plan,
                                                     float* const dAarray[], float* const dBarray[], int* dIpiv,
                                                     int* dInfo);
    aclsolverStatus_t aclsolverSgesvVbatched_destroyPlan(aclsolverVbatchedPlan_t plan);
    /* The schedule: for each 16-wide size bin above 64 (65..80, 81..96, ..., 241..256; 12 bins)
     * the configuration that solves its matrices, as an id: 1 = 2 rows per lane, 4 waves per
     * matrix (n <= 128); 2 = 2 rows, 8 waves (n <= 128); 3 = 3 rows, 8 waves (n <= 192);
     * 4 = 3 rows, 16 waves (n <= 192); 5 = 4 rows, 16 waves (n <= 256). Orders <= 64 always
     * take one wave per matrix. createPlan takes the autotuned schedule of the device;
     * createPlanWithSchedule takes the caller's (INVALID_VALUE for an id outside 1..5 or one
     * whose largest order is below its bin's); getSchedule reports a plan's. */
    aclsolverStatus_t aclsolverSgesvVbatched_createPlanWithSchedule(aclsolverHandle_t handle, int batchSize,
                                                                    const int* nArray, int nrhs, const int* schedule,
                                                                    aclsolverVbatchedPlan_t* plan);
    aclsolverStatus_t aclsolverSgesvVbatched_getSchedule(aclsolverVbatchedPlan_t plan, int* schedule);
    aclsolverStatus_t aclsolverSgesvVbatched(aclsolverHandle_t handle, int batchSize, const int* nArray, int nrhs,
                                             float* const dAarray[], float* const dBarray[], int* dIpiv, int* dInfo);

    /* ---- EXTENSION: contiguous strided batched direct inverse, n <= 64, with optional
     * per-matrix pivots (n*batch ints). No metadata validation pass and no host sync; used by
     * measurement harnesses and by callers that own the contract. */
    aclsolverStatus_t aclsolverSmatinvStridedBatched(aclsolverHandle_t handle, int n, const float* A, int lda,
                                                     long long strideA, float* Ainv, int lda_inv, long long strideInv,
                                                     int* ipiv, int* info, int batchSize);

    /* ---- EXTENSION (measurement): per-launch timing of instrumented kernels. Between
     * Begin and End, launches of kernel class k are bracketed by HIP events on their own
     * stream; End reports the launch count, summed algorithmic work (flop or byte) and the
     * summed duration (ms). k = 0: packed trailing-update GEMM of the mixed-precision getrf
     * (flop); 1: batched inverse; 2: uneven batched getrf+getrs (byte); 3: panel leaves of
     * getrf (byte); 4: trailing-update GEMM of cgetrf (real flop). */
    aclsolverStatus_t aclsolverProfileBegin(void);
    aclsolverStatus_t aclsolverProfileEnd(int kernel, int* launches, double* work, double* ms);

    /* ---- EXTENSION (testing): polls a bounded cross-workgroup wait may spend before it
     * gives up (default ACLSOLVER_DEFAULT_SPIN_LIMIT; 0 makes every wait that is not
     * already satisfied give up at once). Lowering it lets a test drive the timeout path
     * from the host while every kernel still exits normally; results of calls that time
     * out are NaN and devInfo reports ACLSOLVER_INFO_WAIT_TIMEOUT. Library-wide, not
     * thread-safe against concurrent calls. */
    aclsolverStatus_t aclsolverDebugSetSpinLimit(unsigned int polls);

#ifdef __cplusplus
}
#endif
