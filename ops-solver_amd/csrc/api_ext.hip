// EXTENSION entry points: the mixed-precision gesv family with GMRES refinement.
// Argument convention follows the reference's aclsolverSSgels (include/cann_ops_solver.h:815-862).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"

namespace ops
{
size_t GesvIrWorkspaceBytes(int n, int restart);
hipError_t GesvIR(int n, int nrhs, const float* A, int lda, int* ipiv, const float* B, int ldb, float* X, int ldx,
                  void* work, Precision prec, double tol, int maxOuter, int restart, bool butterfly, bool tournament,
                  int* info,
                  int* niterOut, hipStream_t stream, const LookAhead* la);
}  // namespace ops

using namespace ops;

namespace
{
constexpr int kGesvMax = 16384;

aclsolverStatus_t GesvBufferSize(aclsolverHandle_t handle, int n, int nrhs, size_t* bytes)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    if (n < 0 || n > kGesvMax || nrhs < 0 || bytes == nullptr) return ACLSOLVER_STATUS_INVALID_VALUE;
    *bytes = n == 0 ? 0 : GesvIrWorkspaceBytes(n, H(handle)->irRestart);
    return ACLSOLVER_STATUS_SUCCESS;
}

// Argument indices (handle excluded): n=1 nrhs=2 dA=3 ldda=4 dipiv=5 dB=6 lddb=7 dX=8
// lddx=9 dWorkspace=10 lwork_bytes=11 niter=12 dinfo=13.
aclsolverStatus_t Gesv(aclsolverHandle_t handle, Precision prec, int n, int nrhs, float* dA, int ldda, int* dipiv,
                       float* dB, int lddb, float* dX, int lddx, void* dWork, size_t lwork, int* niter, int* dinfo)
{
    if (const aclsolverStatus_t g = CheckHandle(handle)) return g;
    if (dinfo == nullptr || !IsDevice(dinfo)) return ACLSOLVER_STATUS_INVALID_VALUE;
    Handle* h = H(handle);
    int arg = 0;
    const size_t need = n > 0 ? GesvIrWorkspaceBytes(n, h->irRestart) : 0;
    if (n < 0 || n > kGesvMax)
        arg = 1;
    else if (nrhs < 0)
        arg = 2;
    else if (n > 0 && !IsDevice(dA))
        arg = 3;
    else if (ldda < std::max(1, n))
        arg = 4;
    else if (dipiv != nullptr && !IsDevice(dipiv))
        arg = 5;
    else if (n > 0 && nrhs > 0 && !IsDevice(dB))
        arg = 6;
    else if (lddb < std::max(1, n))
        arg = 7;
    else if (n > 0 && nrhs > 0 && (!IsDevice(dX) || dX == dB || dX == dA))
        arg = 8;
    else if (lddx < std::max(1, n))
        arg = 9;
    else if (n > 0 && !IsDevice(dWork))
        arg = 10;
    else if (lwork < need)
        arg = 11;
    else if (niter == nullptr)
        arg = 12;
    if (arg) return Invalid(handle, dinfo, 1, arg);
    hipStream_t s = h->stream;
    *niter = 0;
    FillInfo(dinfo, 1, 0, s);
    if (n == 0 || nrhs == 0) return Launched();
    // dipiv is optional; keep pivots inside the workspace tail when absent.
    int* piv = dipiv;
    if (piv == nullptr)
    {
        void* extra = HandleScratch(handle, static_cast<size_t>(n) * sizeof(int));
        if (extra == nullptr) return ACLSOLVER_STATUS_ALLOC_FAILED;
        piv = static_cast<int*>(extra);
    }
    if (GesvIR(n, nrhs, dA, ldda, piv, dB, lddb, dX, lddx, dWork, prec, h->irTolerance, h->irMaxOuter, h->irRestart,
               h->butterfly != 0, prec != Precision::FP32 && h->tournament != 0, dinfo, niter, s,
               HandleLookAhead(handle)) != hipSuccess)
        return ACLSOLVER_STATUS_EXECUTION_FAILED;
    return Launched();
}
}  // namespace

OPS_API aclsolverStatus_t aclsolverSHgesv_bufferSize(aclsolverHandle_t handle, int n, int nrhs, size_t* lwork_bytes)
{
    return GesvBufferSize(handle, n, nrhs, lwork_bytes);
}
OPS_API aclsolverStatus_t aclsolverSBgesv_bufferSize(aclsolverHandle_t handle, int n, int nrhs, size_t* lwork_bytes)
{
    return GesvBufferSize(handle, n, nrhs, lwork_bytes);
}
OPS_API aclsolverStatus_t aclsolverSSgesv_bufferSize(aclsolverHandle_t handle, int n, int nrhs, size_t* lwork_bytes)
{
    return GesvBufferSize(handle, n, nrhs, lwork_bytes);
}
OPS_API aclsolverStatus_t aclsolverSHgesv(aclsolverHandle_t handle, int n, int nrhs, float* dA, int ldda, int* dipiv,
                                          float* dB, int lddb, float* dX, int lddx, void* dWorkspace,
                                          size_t lwork_bytes, int* niter, int* dinfo)
{
    // FP16 matrix cores on the two-part split with power-of-two row/column scaling.
    return Gesv(handle, Precision::FP16X3, n, nrhs, dA, ldda, dipiv, dB, lddb, dX, lddx, dWorkspace, lwork_bytes, niter,
                dinfo);
}
OPS_API aclsolverStatus_t aclsolverSBgesv(aclsolverHandle_t handle, int n, int nrhs, float* dA, int ldda, int* dipiv,
                                          float* dB, int lddb, float* dX, int lddx, void* dWorkspace,
                                          size_t lwork_bytes, int* niter, int* dinfo)
{
    const int split = handle != nullptr ? H(handle)->bf16Split : 2;
    const Precision p = split >= 2 ? Precision::BF16X6 : (split == 1 ? Precision::BF16X3 : Precision::BF16);
    return Gesv(handle, p, n, nrhs, dA, ldda, dipiv, dB, lddb, dX, lddx, dWorkspace, lwork_bytes, niter, dinfo);
}

// EXTENSION: operand split of SBgesv's trailing updates on bf16 MFMA. 2 (default): three
// bf16 parts per FP32 value, six products (~24-bit, FP32-level products); 1: two parts,
// three products (~16 bits); 0: plain bf16 (8 bits). Fewer bits leave more work to the
// refinement: at condition 1e6 only the three-part form keeps GMRES-IR short.
OPS_API aclsolverStatus_t aclsolverSetBf16Split(aclsolverHandle_t handle, int split)
{
    if (handle == nullptr) return ACLSOLVER_STATUS_HANDLE_IS_NULLPTR;
    if (split < 0 || split > 2) return ACLSOLVER_STATUS_INVALID_VALUE;
    H(handle)->bf16Split = split;
    return ACLSOLVER_STATUS_SUCCESS;
}
// EXTENSION: 1 = SH/SB factor U^T A V without pivoting after two levels of random
// butterfly transforms (n % 4 == 0), falling back to partial pivoting after a zero pivot or
// a refinement failure; 0 (default) = partial pivoting.
OPS_API aclsolverStatus_t aclsolverSetButterfly(aclsolverHandle_t handle, int on)
{
    if (handle == nullptr) return ACLSOLVER_STATUS_HANDLE_IS_NULLPTR;
    H(handle)->butterfly = on != 0;
    return ACLSOLVER_STATUS_SUCCESS;
}
// EXTENSION: pivoting of SH/SB's tall panel leaves. 1: tournament pivoting (the pivots of 32
// columns chosen by a reduction tree over 256-row blocks, no exchange between CUs per
// column); 0 (default): partial pivoting, as sgetrf. SS always pivots partially.
OPS_API aclsolverStatus_t aclsolverSetTournamentPivoting(aclsolverHandle_t handle, int on)
{
    if (handle == nullptr) return ACLSOLVER_STATUS_HANDLE_IS_NULLPTR;
    H(handle)->tournament = on != 0;
    return ACLSOLVER_STATUS_SUCCESS;
}
OPS_API aclsolverStatus_t aclsolverSSgesv(aclsolverHandle_t handle, int n, int nrhs, float* dA, int ldda, int* dipiv,
                                          float* dB, int lddb, float* dX, int lddx, void* dWorkspace,
                                          size_t lwork_bytes, int* niter, int* dinfo)
{
    return Gesv(handle, Precision::FP32, n, nrhs, dA, ldda, dipiv, dB, lddb, dX, lddx, dWorkspace, lwork_bytes, niter,
                dinfo);
}
