/*
 * Common ABI types of the MI355X (gfx950) LU solver library.
 *
 * Mirrors include/cann_ops_solver_common.h of the reference (status codes :36-50, handle
 * :52, complex64 :65-69, static_asserts :125-139) so that code written against the
 * reference compiles unchanged. The ACL runtime types the reference takes from
 * <acl/acl.h> are mapped onto HIP: aclrtStream is a hipStream_t (an opaque pointer, null
 * = the default stream) and aclError is an int with ACL_SUCCESS = 0 and
 * ACL_ERROR_INVALID_PARAM = 100000 (the reference's docs/QUICKSTART.md:140 values).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C"
{
#endif

    typedef void* aclrtStream;
    typedef int aclError;
#ifndef ACL_SUCCESS
#define ACL_SUCCESS 0
#endif
#ifndef ACL_ERROR_INVALID_PARAM
#define ACL_ERROR_INVALID_PARAM 100000
#endif
#ifndef ACL_ERROR_BAD_ALLOC
#define ACL_ERROR_BAD_ALLOC 200000
#endif
#ifndef ACL_ERROR_RT_FAILURE
#define ACL_ERROR_RT_FAILURE 500000
#endif

    /* include/cann_ops_solver_common.h:36-50 */
    typedef enum aclsolverStatus
    {
        ACLSOLVER_STATUS_SUCCESS = 0,
        ACLSOLVER_STATUS_NOT_INITIALIZED = 1,
        ACLSOLVER_STATUS_ALLOC_FAILED = 2,
        ACLSOLVER_STATUS_INVALID_VALUE = 3,
        ACLSOLVER_STATUS_MAPPING_ERROR = 4,
        ACLSOLVER_STATUS_EXECUTION_FAILED = 5,
        ACLSOLVER_STATUS_INTERNAL_ERROR = 6,
        ACLSOLVER_STATUS_NOT_SUPPORTED = 7,
        ACLSOLVER_STATUS_ARCH_MISMATCH = 8,
        ACLSOLVER_STATUS_HANDLE_IS_NULLPTR = 9,
        ACLSOLVER_STATUS_INVALID_ENUM = 10,
        ACLSOLVER_STATUS_UNKNOWN = 11
    } aclsolverStatus_t;

    /* include/cann_ops_solver_common.h:52 */
    typedef void* aclsolverHandle_t;

    /* include/cann_ops_solver_common.h:58-69: two consecutive IEEE float32 values. */
    typedef struct
    {
        float real;
        float imag;
    } aclsolverFloatComplex;

    typedef struct
    {
        float real;
        float imag;
    } aclFloatComplex;

    /* include/cann_ops_solver_common.h:85-90 */
    typedef enum
    {
        ACLSOLVER_OP_N = 0,
        ACLSOLVER_OP_T = 1,
        ACLSOLVER_OP_C = 2
    } aclsolverOperation_t;

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
static_assert(sizeof(aclFloatComplex) == 2 * sizeof(float), "complex size must be two FP32 components");
static_assert(offsetof(aclFloatComplex, real) == 0, "real must be the first component");
static_assert(offsetof(aclFloatComplex, imag) == sizeof(float), "imag must follow real");
static_assert(sizeof(aclsolverStatus_t) == 4, "status is a 4-byte enum");
#endif
