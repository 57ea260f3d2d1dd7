// Internal declarations shared by the host and kernel translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "cann_ops_solver.h"

#define OPS_API extern "C" __attribute__((visibility("default")))

namespace ops
{

// Side stream and events for look-ahead factorizations: the trailing update of block
// step k runs on `side` while the caller's stream factors panel k + 1.
struct LookAhead
{
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr;
    hipEvent_t panel[2] = {nullptr, nullptr};
    hipEvent_t bulk[2] = {nullptr, nullptr};
    // Optional CU partition: the panel path runs on `panelStream`, restricted to a reserved
    // set of CUs, and `side` on the others, so the latency-bound panel kernels do not share
    // CUs with the trailing update; `join` hands the result back to the caller's stream.
    hipStream_t panelStream = nullptr;
    hipEvent_t join = nullptr;
};

// Solver handle: the stream plus per-handle options and a reusable scratch allocation.
// Reference: src/utils/aclsolver_handle_internal.h:33-100 (stream + per-context LU scratch).
struct Handle
{
    uint32_t magic = 0x4C55AC01u;
    hipStream_t stream = nullptr;
    int device = 0;
    int strict = 0;           // enforce the reference's size limits
    double irTolerance = 1e-6;
    int irMaxOuter = 30;
    int irRestart = 30;
    int bf16Split = 2;        // SBgesv trailing updates: 0 plain bf16, 1 two-part, 2 three-part split
    int butterfly = 0;        // SH/SB: random butterflies + no pivoting instead of partial pivoting
    int tournament = 0;       // SH/SB: tall panel leaves pivot by tournament (0: partial pivoting)
    void* scratch = nullptr;  // handle-owned Device scratch (grown on demand)
    int* dflag = nullptr;     // Device flag word for validation scans
    int* hflag = nullptr;     // pinned host mirror of dflag
    size_t scratchBytes = 0;
    int numCUs = 256;
    LookAhead la;             // created on first use
    bool laReady = false;
    bool archOk = true;       // the current device at creation is a gfx950
};

// The handle's look-ahead resources (created on first call, default stream priority);
// nullptr when they cannot be created, in which case factorizations run on one stream.
const LookAhead* HandleLookAhead(aclsolverHandle_t h);

inline Handle* H(aclsolverHandle_t h) { return static_cast<Handle*>(h); }
inline hipStream_t Stream(aclsolverHandle_t h) { return H(h)->stream; }

// Entry-point guard: HANDLE_IS_NULLPTR for a null handle, ARCH_MISMATCH when the handle
// was created on a device other than gfx950, the only target the kernels are built for
// (reference src/utils/complex_solver_host.h:50-53 answers ARCH_MISMATCH off its target).
inline aclsolverStatus_t CheckHandle(aclsolverHandle_t h)
{
    if (h == nullptr) return ACLSOLVER_STATUS_HANDLE_IS_NULLPTR;
    return H(h)->archOk ? ACLSOLVER_STATUS_SUCCESS : ACLSOLVER_STATUS_ARCH_MISMATCH;
}

// Largest order the register-resident factorization engines support (rows of getrf,
// order of getrs / getri / gesv).
constexpr int kEngineMaxOrder = 16384;

// Polls a bounded wait may spend before it gives up (multi-workgroup panel leaves,
// chained triangular solves). Library-wide; lowered only to test the failure path.
unsigned SpinLimit();
// devInfo value a factorization or solve reports when one of its bounded waits gave up
// (its results are then NaN); ACLSOLVER_INFO_WAIT_TIMEOUT in the public header.
constexpr int kInfoWaitTimeout = -1000;
// info <- kInfoWaitTimeout when *err != 0 (one thread, stream-ordered after the waits).
void FoldWaitError(const int* err, int* info, hipStream_t stream);

// True when pointer is Device memory of the current process (hipMalloc / managed).
bool IsDevice(const void* p);
// True when pointer is NOT Device memory (plain host or pinned host).
bool IsHostOnly(const void* p);

// Grows the handle scratch to at least bytes; returns nullptr on failure. The scratch is
// only touched by work on the handle's stream, so stream order protects reuse.
void* HandleScratch(aclsolverHandle_t h, size_t bytes);

// Device address of this device's fixed no-allocation scratch array (noalloc.hip), looked
// up once per device with the device current; nullptr when the lookup fails.
void* NoAllocScratch(int device);
// Once per device at handle creation: may the tall panel leaves exchange through one XCD's
// L2 (getrf.hip)?
void LeafSpreadProbe(int device);
// May the full-width tall leaves of the current device run spread over 8x their grid and
// exchange through one XCD's L2 (probe result, OPS_SOLVER_LEAF_XCD)?
bool LeafSpread();

// Writes value into count ints of a Device info array on stream.
void FillInfo(int* info, int count, int value, hipStream_t stream);
// Invalid argument: writes -argument into info (when it is a Device pointer) and returns
// INVALID_VALUE (reference src/utils/real_lu_runtime.h:61-78).
aclsolverStatus_t Invalid(aclsolverHandle_t h, int* info, int count, int argument);
// Returns EXECUTION_FAILED when the last launch on this thread failed.
aclsolverStatus_t Launched();
// Scans a Device matrix (or batch of them) for NaN/Inf; returns true when all finite.
// Synchronizes the stream (reference src/utils/real_lu_common.h:90-132).
bool AllFinite(aclsolverHandle_t h, const float* a, int rows, int cols, int lda, int64_t batchStride, int batch,
               bool* ok);
bool AllFiniteArrays(aclsolverHandle_t h, const float* const* arrays, int rows, int cols, int lda, int batch,
                     bool* ok);

// ---- optional launch timing (profile.hip) ----
enum class ProfileKernel
{
    TrailingGemm = 0,  // the packed BF16/FP16 MFMA trailing-update kernel alone (2 M N K flops)
    BatchedInverse = 1,
    Vbatched = 2,
    PanelLeaf = 3,   // one panel leaf of the LU (all its launches); work = 8 * rows * columns bytes
    ComplexGemm = 4  // the complex LU's trailing update (cgemm_sub_kernel); work = 8 M N K real flops
};
bool ProfilingOn();
// Brackets the launches issued during its lifetime with HIP events on `stream`.
class ProfileScope
{
   public:
    ProfileScope(ProfileKernel k, double work, hipStream_t stream);
    ~ProfileScope();
    ProfileScope(const ProfileScope&) = delete;
    ProfileScope& operator=(const ProfileScope&) = delete;

   private:
    ProfileKernel kernel_;
    double work_;
    hipStream_t stream_;
    hipEvent_t start_ = nullptr;
};

// ---- real LU engine (getrf.hip) ----
enum class Precision
{
    FP32 = 0,
    FP16 = 1,
    BF16 = 2,
    BF16X3 = 3,  // two-part split bf16 (h h' + h m' + m h'), ~16-bit products
    BF16X6 = 4,  // three-part split bf16 (six products), ~24-bit (FP32-level) products
    FP16X3 = 5   // two-part split fp16 with power-of-two row/column scaling, ~22-bit products
};
// Workspace bytes needed by RealGetrf for an m x n matrix.
size_t RealGetrfWorkspaceBytes(int m, int n);
// Device address of the factorization's timeout flag inside its workspace: nonzero after a
// bounded wait of a panel leaf gave up (a multi-workgroup leaf waiting for its peers, or a
// tournament node for its children's candidates); the factors are then invalid.
int* RealGetrfErrorWord(void* work, int m);
// Blocked right-looking LU with partial pivoting (or none when ipiv == nullptr).
// Writes ipiv (1-based) and info (first zero pivot). Returns hipSuccess on launch success.
hipError_t RealGetrf(int m, int n, float* a, int lda, int* ipiv, int* info, void* work, Precision trailing,
                     hipStream_t stream);
// The blocked LU with look-ahead behind RealGetrf. permOut (optional, m ints) receives the
// row permutation P with (P A)[i] = A[perm[i]]; pack (optional): the packed-operand buffer
// of the mixed-precision trailing updates; la (optional): the side stream and events of the
// look-ahead. tournament: tall panel leaves choose their pivots by tournament (a reduction
// tree over row blocks) instead of partial pivoting -- for callers that only need stable
// factors, not the partial-pivoting row order.
hipError_t RealGetrfPerm(int m, int n, float* a, int lda, int* ipiv, int* info, void* work, Precision trailing,
                         int* permOut, hipStream_t stream, void* pack, const LookAhead* la, bool tournament = false);
// Device column-major getrs from RealGetrf factors (trans 0 / 1). work: RealGetrsWorkspaceBytes.
size_t RealGetrsWorkspaceBytes(int n, int nrhs);
// info (optional, Device): set to kInfoWaitTimeout when a bounded wait of the solve gave up.
hipError_t RealGetrs(int trans, int n, int nrhs, const float* lu, int lda, const int* ipiv, float* b, int ldb,
                     void* work, hipStream_t stream, int* info = nullptr);

// ---- complex LU engine (complex_engine.hip) ----
// ipiv (1-based interchange sequence) -> perm (perm[i] = source row of row i of P*B), one
// launch; clampBelow also takes entries pointing above their own row as no interchange.
hipError_t IpivToPerm(int n, const int* ipiv, int* perm, hipStream_t stream, bool clampBelow = false);
size_t ComplexGetrfWorkspaceBytes(int m, int n);
// la (optional): look-ahead streams and events of the handle (the trailing updates then run
// on la->side beside the next panel).
hipError_t ComplexGetrf(int m, int n, float2* a, int lda, int* ipiv, int* info, void* work, hipStream_t stream,
                        const LookAhead* la = nullptr);
size_t ComplexGetrsWorkspaceBytes(int n, int nrhs);
hipError_t ComplexGetrs(int trans, int n, int nrhs, const float2* lu, int lda, const int* ipiv, float2* b, int ldb,
                        void* work, hipStream_t stream);

}  // namespace ops
