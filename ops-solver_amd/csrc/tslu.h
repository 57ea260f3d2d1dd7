// Tournament-pivoting panel leaf (tslu.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace ops
{
// Columns per tournament leaf.
constexpr int kTsluWidth = 32;
// Largest number of active rows a tournament leaf takes (256-row blocks, at most 64).
constexpr int kTsluMaxRows = 16384;

// Device scratch of the tournament leaves of an m-row factorization.
size_t TsluScratchBytes(int m);
// Zeroes the arrival counters and row flags (once per factorization, before its first leaf;
// every leaf leaves them zero again).
hipError_t TsluClear(void* scratch, int m, hipStream_t stream);
// Factors columns [c0, c0 + wact) (wact <= kTsluWidth) of rows [c0, m) with tournament
// pivoting, with the same outputs as the partial-pivoting leaves: ipiv[c0 .. c0+wact) as a
// LAPACK interchange sequence (1-based, global rows), each row of the leaf columns at its
// final position (L below the diagonal, U on and above it), rowsrc for the panel starting
// at j0 (newPanel: c0 == j0), the interchanges applied to the other panel columns
// [j0, c0) and [c0 + wact, panelEnd), and info = first zero pivot (1-based) if none yet.
// tag: nonzero and unique among the leaves since the last TsluClear. The tree's nodes hand
// their candidates on through tagged words; a node that does not see its children's words
// within the spin limit raises *err and the leaf (and every later one) leaves its outputs
// unwritten.
hipError_t TsluLeaf(float* A, int lda, int m, int j0, int c0, int wact, int panelEnd, int newPanel, int* ipiv,
                    int* info, int* rowsrc, void* scratch, unsigned tag, int* err, hipStream_t stream);
}  // namespace ops
