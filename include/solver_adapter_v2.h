/*
 * The operator-adapter protocol (version 2) through which an external test package drives
 * one LU operator: create / reset / execute / readback / destroy over compact row-major
 * HOST transport buffers, the adapter owning the device copies in the operator's own layout.
 *
 * The structure layouts and function signatures are an ABI shared with the reference's
 * harness (reference test/sgetrf/solver_adapter.h:25-69) and must match it field for field;
 * the adapters behind them (ops-solver_amd/adapter/lu_adapter.hip) are this library's own.
 *
 * Adapter return codes: 0 success, 1 invalid case, 2 allocation failed, 3 runtime failure,
 * 4 workspace query failed. The operator's own status is returned separately (dut_status).
 */
#ifndef OPS_SOLVER_ADAPTER_V2_H
#define OPS_SOLVER_ADAPTER_V2_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C"
{
#endif

#define SOLVER_ADAPTER_VERSION 1u
#define SOLVER_ADAPTER_VERSION_V2 2u

    /* What the adapter drives: operator name, element type, memory domain and layout of the
     * operator's buffers, kind of info output, and the address of the operator function. */
    typedef struct
    {
        uint32_t version, struct_size;
        const char *op, *dtype, *dut_abi, *dut_layout, *info_kind;
        const void *target_symbol;
    } SolverAdapterDesc;

    /* One test case (version-1 part). Transport buffers are compact row-major host arrays
     * of batch matrices; complex64 is interleaved (real, imag) float32. */
    typedef struct
    {
        uint32_t version, struct_size;
        int64_t n, nrhs, batch, lda, ldb;
        char uplo;
        const char *info_probe, *input_kind;
        size_t a_bytes, b_bytes, out_bytes, info_count;
    } SolverAdapterCase;

    /* Version-2 case: the version-1 case plus rows, a second order, leading dimensions of
     * the output and solution, side / trans flags and the pivot (1-based int32) and tau
     * counts of the inputs and outputs. */
    typedef struct
    {
        SolverAdapterCase base;
        int64_t m, k, ldc, ldx;
        char side, trans;
        size_t in_ipiv_count, in_tau_count, out_ipiv_count, out_tau_count;
    } SolverAdapterCaseV2;

    const SolverAdapterDesc *solver_adapter_descriptor(void);
    /* The harness owns handle and stream; a partially built state is published through
     * *state before anything can fail, so destroy can release it. */
    int solver_adapter_create_v2(const SolverAdapterCaseV2 *c, void *handle, void *stream, void **state);
    /* Restores every input, output, pivot and info buffer; never calls the operator. */
    int solver_adapter_reset_v2(void *state, const void *a, size_t a_bytes, const void *b, size_t b_bytes,
                                const int32_t *ipiv, size_t ipiv_count, const void *tau, size_t tau_bytes);
    /* Calls the operator once; *dut_status receives its unmodified return code. */
    int solver_adapter_execute(void *state, int *dut_status);
    /* After the harness synchronized the stream: outputs back to row-major host transport. */
    int solver_adapter_readback_v2(void *state, void *out, size_t out_bytes, int32_t *info, size_t info_count,
                                   int32_t *ipiv, size_t ipiv_count, void *tau, size_t tau_bytes);
    /* Releases everything; a null or partially built state is valid. */
    int solver_adapter_destroy(void *state);

#ifdef __cplusplus
}
#endif
#endif
