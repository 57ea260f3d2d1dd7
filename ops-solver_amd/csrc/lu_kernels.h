// Device building blocks of the real LU path, shared by getrf, getrs and the gesv family.
#pragma once
#include <hip/hip_runtime.h>

#include "internal.h"

namespace ops
{

// C(MxN) -= op(A)(MxK) * B(KxN), column-major, on MFMA with FP32 accumulation. op(A) = A
// (transA = 0) or A^T (transA = 1). prec selects the MFMA input type (FP32, or FP16 with
// per-tile power-of-two scaling, or BF16). The product is accumulated from zero and
// subtracted from C once (reference src/sgetrf/real_lu_blocked_kernel.cpp:1152-1153).
hipError_t GemmSub(int M, int N, int K, const float* A, int lda, int transA, const float* B, int ldb, float* C,
                   int ldc, Precision prec, hipStream_t stream);

// C(MxN) -= A(MxK) * B(KxN) (column-major, no transposes) on BF16 MFMA from operands packed
// once into fragment order in `work` (GemmPackedWorkspaceBytes): prec BF16 (hi*hi) or
// BF16X3 (split form); other precisions go to GemmSub.
size_t GemmPackedWorkspaceBytes(int M, int N, int K);
hipError_t GemmSubPacked(int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
                         Precision prec, void* work, hipStream_t stream);

// Triangular solve with a K x K diagonal block T (column-major, ldt) on a K x N block B:
// B <- op(T)^{-1} B. lower: T's lower triangle is used (else upper); unit: unit diagonal;
// transT: op(T) = T^T. Runs as 64-row diagonal solves (a quad of lanes per column, in
// registers) interleaved with MFMA GEMM updates of the remaining rows.
hipError_t TrsmBlock(int K, int N, const float* T, int ldt, bool lower, bool unit, bool transT, float* B, int ldb,
                     hipStream_t stream);

// B <- T^{-1} B for a unit-lower K x K block T (K <= 512), in two steps: Inverse writes
// the inverses of T's 64 x 64 diagonal blocks to linv (TrsmLowerUnitWorkspaceBytes), Apply
// then solves any number of column blocks in one launch (one workgroup per 32 columns,
// products on FP32 MFMA). One Inverse serves every Apply with the same T.
size_t TrsmLowerUnitWorkspaceBytes(int K);
hipError_t TrsmLowerUnitInverse(int K, const float* T, int ldt, float* linv, hipStream_t stream);
hipError_t TrsmLowerUnitApply(int K, int N, const float* T, int ldt, const float* linv, float* B, int ldb,
                              hipStream_t stream);

// Batched inverse from LU factors for orders up to 512: C_b = U_b^{-1} L_b^{-1} P_b for
// every matrix (pointer arrays), info_b = first zero U(k, k) + 1. piv (batch * n, 1-based)
// may be null (no pivoting). work: GetriBlockedBatchedWorkspaceBytes(n, batch); at most
// 65535 matrices per call.
size_t GetriBlockedBatchedWorkspaceBytes(int n, int batch);
hipError_t GetriBlockedBatched(int n, const float* const* Aarr, int lda, const int* piv, float* const* Carr, int ldc,
                               int* info, int batch, float* work, hipStream_t stream);

}  // namespace ops
