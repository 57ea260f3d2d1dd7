// Single-launch triangular solve of one right-hand side (trsv.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace ops
{
// Device state of the chained solve: one publication word per row, a ticket counter and
// an error word. Reset it (TrsvChainReset) before the first solve on a new buffer; after
// that, consecutive solves on the same stream reuse it without clearing.
struct TrsvChainState
{
    unsigned long long* pub;
    unsigned* ticket;
    int* err;  // set to 1 when a bounded wait ran out (the solve's result is then invalid)
    int n;
    long long launches;
};
size_t TrsvChainStateBytes(int n);
TrsvChainState TrsvChainLayout(void* base, int n);
hipError_t TrsvChainReset(TrsvChainState& s, hipStream_t stream);
// x <- T^{-1} x for the unit lower (upper = false) or non-unit upper (upper = true)
// triangle of the packed LU factors T (n x n, column-major, ldt). The upper solve divides
// through a rounded reciprocal of each pivot. The double overload carries the whole
// substitution in FP64 on the FP32 factors.
hipError_t TrsvChain(TrsvChainState& s, bool upper, const float* T, int ldt, float* x, hipStream_t stream);
hipError_t TrsvChain(TrsvChainState& s, bool upper, const float* T, int ldt, double* x, hipStream_t stream);
}  // namespace ops
