/*
 * CPU oracle for the LU family (getrf / getrs / getri / matinv), real and complex.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or calls this
 * file; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, and
 * only as the checker of results produced by the HIP path.
 *
 * Every routine restates the algorithm of the reference operator it checks, citing the
 * reference file:line it follows (paths relative to the reference repository root):
 *
 *  - pivot rule: first row holding the maximum |x| in the active column, lower rows
 *    only replace the candidate on a strictly larger magnitude
 *      src/utils/kernel/real_lu.hpp:59-77 (Pivot), :100-139 (PivotRF tie rule),
 *      src/sgetrf/real_lu_blocked_kernel.cpp:342-389 (PanelPivot, first-row ties)
 *  - zero pivot: info = first k+1 with U(k,k) == 0, elimination of that column skipped,
 *    the factorization keeps going         src/utils/kernel/real_lu.hpp:419-437
 *  - column scaling: the small (unblocked) path divides by the pivot
 *      src/utils/kernel/real_lu.hpp:226-247 (ScaleColumn); the blocked panel multiplies
 *    by the rounded reciprocal when |d| >= FLT_MIN and divides otherwise
 *      src/sgetrf/real_lu_blocked_kernel.cpp:390-410 (PreparePanelReciprocal), :467-482
 *  - rank-1 update a(i,j) <- a(i,j) - l(i) * u(j) as one fused multiply-add
 *      src/utils/kernel/real_lu.hpp:312-327 (UpdateColumn: Axpy with -coefficient)
 *  - getrs: row interchanges in ipiv order, unit-lower then upper solve (trans 0),
 *    U^T then L^T then interchanges in reverse (trans 1/2)
 *      src/sgetrs/sgetrs_kernel.cpp:33-62, src/sgetrs/kernel/laswp.hpp:107-136,
 *      test/utils/lu_verify_utils.py:67-191
 *  - getri: LAPACK order: inv(U), then X*L = inv(U) solved column by column from the
 *    right, then the column interchanges applied in reverse pivot order
 *      src/utils/kernel/real_lu_small_inverse.hpp:183-216, :220-242
 *  - complex pivot score |re|+|im| with NaN inactive, both zeros equal and an
 *    overflow-safe half sum               src/utils/complex_solver_kernel.hpp:47-63
 *  - complex rank-1 update x - a*b with the fused order of SubMul
 *                                          src/utils/complex_solver_kernel.hpp:65-68
 *
 * All matrices are column-major: element (r, c) sits at a[c * lda + r]. Complex matrices
 * are interleaved (re, im) float or double pairs. ipiv is 1-based.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------------------ */
/* Real single precision                                                                 */
/* ------------------------------------------------------------------------------------ */

/* First maximum of |a(r, k)| over r in [k, m); also reports the relative gap between the
 * winner and the best other candidate (the tie margin a parity test filters on). */
static int pivot_row_f(const float *col, int k, int m, float *margin)
{
    int row = k;
    float best = fabsf(col[k]);
    for (int i = k + 1; i < m; ++i)
    {
        float v = fabsf(col[i]);
        if (v > best)
        {
            best = v;
            row = i;
        }
    }
    if (margin)
    {
        float second = 0.0f;
        for (int i = k; i < m; ++i)
            if (i != row && fabsf(col[i]) > second) second = fabsf(col[i]);
        *margin = best > 0.0f ? (best - second) / best : 0.0f;
    }
    return row;
}

/* sgetrf. divide != 0 selects the small-path division scaling, divide == 0 the blocked
 * panel's reciprocal scaling. ipiv == NULL factors without pivoting. margin (nullable)
 * receives min(m, n) tie margins. Returns info (0 or first zero pivot, 1-based). */
ORC_EXPORT int orc_sgetrf(int m, int n, float *a, int lda, int *ipiv, int divide, float *margin)
{
    int info = 0;
    int steps = m < n ? m : n;
    for (int k = 0; k < steps; ++k)
    {
        float *col = a + (size_t)k * lda;
        int p = k;
        if (ipiv)
        {
            p = pivot_row_f(col, k, m, margin ? margin + k : NULL);
            ipiv[k] = p + 1;
            if (p != k)
                for (int j = 0; j < n; ++j)
                {
                    float t = a[(size_t)j * lda + k];
                    a[(size_t)j * lda + k] = a[(size_t)j * lda + p];
                    a[(size_t)j * lda + p] = t;
                }
        }
        else if (margin)
            margin[k] = 1.0f;
        float d = col[k];
        if (d == 0.0f)
        {
            if (info == 0) info = k + 1;
            continue;
        }
        if (divide || fabsf(d) < FLT_MIN)
        {
            for (int i = k + 1; i < m; ++i) col[i] = col[i] / d;
        }
        else
        {
            float r = 1.0f / d;
            for (int i = k + 1; i < m; ++i) col[i] = col[i] * r;
        }
        for (int j = k + 1; j < n; ++j)
        {
            float *cj = a + (size_t)j * lda;
            float u = cj[k];
            for (int i = k + 1; i < m; ++i) cj[i] = fmaf(-col[i], u, cj[i]);
        }
    }
    return info;
}

/* Tournament pivot selection of one panel leaf (no reference counterpart: the library's
 * communication-avoiding factorization, aclsolverSgetrfTournament, and the SH/SB solves).
 * Rows [0, m) x columns [0, w) of a. Level 0 splits the rows into blocks of node_rows; every
 * node factors its rows (in stacking order) with orc_sgetrf's blocked rule and keeps the
 * rows of its first min(w, rows) pivots, in pivot order, with their ORIGINAL values; a node
 * of the next level stacks the kept rows of `group` consecutive children. The root's
 * pivot order, turned into a LAPACK interchange sequence, is written to ipiv[0, w)
 * (1-based, relative to row 0). Returns the root's info. A matrix of at most node_rows
 * rows is a single node: the result is orc_sgetrf's. */
ORC_EXPORT int orc_stslu_pivots(int m, int w, const float *a, int lda, int node_rows, int group, int *ipiv)
{
    int nodes = (m + node_rows - 1) / node_rows;
    int *ids = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    int *cnt = (int *)malloc(sizeof(int) * (size_t)(nodes > 0 ? nodes : 1));
    int *next = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    int *ncnt = (int *)malloc(sizeof(int) * (size_t)(nodes > 0 ? nodes : 1));
    float *blk = (float *)malloc(sizeof(float) * (size_t)node_rows * (size_t)w);
    int *piv = (int *)malloc(sizeof(int) * (size_t)w);
    int *order = (int *)malloc(sizeof(int) * (size_t)node_rows);
    int info = 0;
    /* level 0: node b holds rows [b*node_rows, ...) */
    for (int b = 0; b < nodes; ++b)
    {
        int r0 = b * node_rows, nr = m - r0 < node_rows ? m - r0 : node_rows;
        for (int i = 0; i < nr; ++i) ids[(size_t)b * node_rows + i] = r0 + i;
        cnt[b] = nr;
    }
    for (;;)
    {
        /* factor every node; keep its first min(w, rows) pivot rows */
        for (int b = 0; b < nodes; ++b)
        {
            int nr = cnt[b];
            const int *id = ids + (size_t)b * node_rows;
            for (int j = 0; j < w; ++j)
                for (int i = 0; i < nr; ++i) blk[(size_t)j * nr + i] = a[(size_t)j * lda + id[i]];
            int steps = nr < w ? nr : w;
            int nodeinfo = orc_sgetrf(nr, w, blk, nr, piv, 0, NULL);
            for (int i = 0; i < nr; ++i) order[i] = id[i];
            for (int k = 0; k < steps; ++k)
            {
                int p = piv[k] - 1, t = order[k];
                order[k] = order[p];
                order[p] = t;
            }
            if (nodes == 1)
            {
                info = nodeinfo;
                /* root: interchange sequence of its pivot rows (row ids are panel rows) */
                int *cpos = (int *)malloc(sizeof(int) * (size_t)w);
                for (int l = 0; l < steps; ++l) cpos[l] = order[l];
                for (int j = 0; j < steps; ++j)
                {
                    int x = cpos[j];
                    ipiv[j] = x + 1;
                    for (int l = j + 1; l < steps; ++l)
                        if (cpos[l] == j) cpos[l] = x;
                }
                free(cpos);
                goto done;
            }
            for (int k = 0; k < steps; ++k) next[(size_t)b * w + k] = order[k];
            ncnt[b] = steps;
        }
        /* stack the kept rows of `group` children per parent */
        int parents = (nodes + group - 1) / group;
        for (int q = 0; q < parents; ++q)
        {
            int nr = 0;
            for (int c = q * group; c < nodes && c < (q + 1) * group; ++c)
                for (int k = 0; k < ncnt[c]; ++k) ids[(size_t)q * node_rows + nr++] = next[(size_t)c * w + k];
            cnt[q] = nr;
        }
        nodes = parents;
    }
done:
    free(ids);
    free(cnt);
    free(next);
    free(ncnt);
    free(blk);
    free(piv);
    free(order);
    return info;
}

/* Row interchanges of a column-major block b (n rows, nrhs columns), forward order. */
static void laswp_f(float *b, int ldb, int nrhs, const int *ipiv, int k0, int k1, int forward)
{
    if (!ipiv) return;
    if (forward)
    {
        for (int k = k0; k < k1; ++k)
        {
            int p = ipiv[k] - 1;
            if (p != k)
                for (int j = 0; j < nrhs; ++j)
                {
                    float t = b[(size_t)j * ldb + k];
                    b[(size_t)j * ldb + k] = b[(size_t)j * ldb + p];
                    b[(size_t)j * ldb + p] = t;
                }
        }
    }
    else
    {
        for (int k = k1 - 1; k >= k0; --k)
        {
            int p = ipiv[k] - 1;
            if (p != k)
                for (int j = 0; j < nrhs; ++j)
                {
                    float t = b[(size_t)j * ldb + k];
                    b[(size_t)j * ldb + k] = b[(size_t)j * ldb + p];
                    b[(size_t)j * ldb + p] = t;
                }
        }
    }
}

/* sgetrs on column-major data: trans 0 solves A X = B, trans 1 solves A^T X = B. */
ORC_EXPORT void orc_sgetrs(int trans, int n, int nrhs, const float *lu, int lda, const int *ipiv, float *b, int ldb)
{
    for (int j = 0; j < nrhs; ++j)
    {
        float *x = b + (size_t)j * ldb;
        if (trans == 0)
        {
            laswp_f(x, ldb, 1, ipiv, 0, n, 1);
            for (int k = 0; k < n; ++k) /* L y = P b, unit diagonal */
            {
                float v = x[k];
                const float *c = lu + (size_t)k * lda;
                for (int i = k + 1; i < n; ++i) x[i] = fmaf(-c[i], v, x[i]);
            }
            for (int k = n - 1; k >= 0; --k) /* U x = y */
            {
                const float *c = lu + (size_t)k * lda;
                x[k] = x[k] / c[k];
                float v = x[k];
                for (int i = 0; i < k; ++i) x[i] = fmaf(-c[i], v, x[i]);
            }
        }
        else
        {
            for (int k = 0; k < n; ++k) /* U^T w = b */
            {
                const float *c = lu + (size_t)k * lda;
                float s = x[k];
                for (int i = 0; i < k; ++i) s = fmaf(-c[i], x[i], s);
                x[k] = s / c[k];
            }
            for (int k = n - 1; k >= 0; --k) /* L^T z = w */
            {
                const float *c = lu + (size_t)k * lda;
                float s = x[k];
                for (int i = k + 1; i < n; ++i) s = fmaf(-c[i], x[i], s);
                x[k] = s;
            }
            laswp_f(x, ldb, 1, ipiv, 0, n, 0);
        }
    }
}

/* sgetri: in-place inverse from packed LU + pivots (LAPACK order). work holds n floats.
 * Returns info: first zero U diagonal (1-based), in which case a is left untouched. */
ORC_EXPORT int orc_sgetri(int n, float *a, int lda, const int *ipiv)
{
    for (int k = 0; k < n; ++k)
        if (a[(size_t)k * lda + k] == 0.0f) return k + 1;
    /* inv(U) in place (column-oriented trtri, upper, non-unit) */
    for (int j = 0; j < n; ++j)
    {
        float *cj = a + (size_t)j * lda;
        cj[j] = 1.0f / cj[j];
        float ajj = -cj[j];
        /* x = triu(inv(U))(0:j,0:j) * a(0:j, j) */
        for (int k = 0; k < j; ++k)
        {
            float t = cj[k];
            if (t != 0.0f)
            {
                const float *ck = a + (size_t)k * lda;
                for (int i = 0; i < k; ++i) cj[i] = fmaf(ck[i], t, cj[i]);
                cj[k] = ck[k] * t;
            }
        }
        for (int i = 0; i < j; ++i) cj[i] *= ajj;
    }
    /* solve inv(A) * L = inv(U) for inv(A), from the last column to the first */
    float *work = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    for (int j = n - 1; j >= 0; --j)
    {
        float *cj = a + (size_t)j * lda;
        for (int i = j + 1; i < n; ++i)
        {
            work[i] = cj[i];
            cj[i] = 0.0f;
        }
        for (int k = j + 1; k < n; ++k)
        {
            const float *ck = a + (size_t)k * lda;
            float w = work[k];
            for (int i = 0; i < n; ++i) cj[i] = fmaf(-ck[i], w, cj[i]);
        }
    }
    free(work);
    /* column interchanges in reverse pivot order */
    if (ipiv)
        for (int j = n - 2; j >= 0; --j)
        {
            int p = ipiv[j] - 1;
            if (p != j)
                for (int i = 0; i < n; ++i)
                {
                    float t = a[(size_t)j * lda + i];
                    a[(size_t)j * lda + i] = a[(size_t)p * lda + i];
                    a[(size_t)p * lda + i] = t;
                }
        }
    return 0;
}

/* smatinv: direct inverse of an original matrix; factor with division scaling (the small
 * path, src/utils/kernel/real_lu_small_inverse.hpp:426-468) then LAPACK-order inverse.
 * c receives the inverse (ldc), ipiv_out (nullable) the pivots. Returns info. */
ORC_EXPORT int orc_smatinv(int n, const float *a, int lda, float *c, int ldc, int *ipiv_out)
{
    int *piv = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; ++j) memcpy(c + (size_t)j * ldc, a + (size_t)j * lda, sizeof(float) * (size_t)n);
    int info = orc_sgetrf(n, n, c, ldc, piv, 1, NULL);
    if (info == 0) orc_sgetri(n, c, ldc, piv);
    if (ipiv_out) memcpy(ipiv_out, piv, sizeof(int) * (size_t)n);
    free(piv);
    return info;
}

/* ------------------------------------------------------------------------------------ */
/* Real double precision (the FP64 oracle used for forward/backward errors)              */
/* ------------------------------------------------------------------------------------ */

ORC_EXPORT int orc_dgetrf(int m, int n, double *a, int lda, int *ipiv)
{
    int info = 0;
    int steps = m < n ? m : n;
    for (int k = 0; k < steps; ++k)
    {
        double *col = a + (size_t)k * lda;
        int p = k;
        if (ipiv)
        {
            double best = fabs(col[k]);
            for (int i = k + 1; i < m; ++i)
                if (fabs(col[i]) > best)
                {
                    best = fabs(col[i]);
                    p = i;
                }
            ipiv[k] = p + 1;
            if (p != k)
                for (int j = 0; j < n; ++j)
                {
                    double t = a[(size_t)j * lda + k];
                    a[(size_t)j * lda + k] = a[(size_t)j * lda + p];
                    a[(size_t)j * lda + p] = t;
                }
        }
        double d = col[k];
        if (d == 0.0)
        {
            if (info == 0) info = k + 1;
            continue;
        }
        for (int i = k + 1; i < m; ++i) col[i] /= d;
        for (int j = k + 1; j < n; ++j)
        {
            double *cj = a + (size_t)j * lda;
            double u = cj[k];
            if (u != 0.0)
                for (int i = k + 1; i < m; ++i) cj[i] -= col[i] * u;
        }
    }
    return info;
}

ORC_EXPORT void orc_dgetrs(int trans, int n, int nrhs, const double *lu, int lda, const int *ipiv, double *b,
                           int ldb)
{
    for (int j = 0; j < nrhs; ++j)
    {
        double *x = b + (size_t)j * ldb;
        if (trans == 0)
        {
            if (ipiv)
                for (int k = 0; k < n; ++k)
                {
                    int p = ipiv[k] - 1;
                    double t = x[k];
                    x[k] = x[p];
                    x[p] = t;
                }
            for (int k = 0; k < n; ++k)
            {
                const double *c = lu + (size_t)k * lda;
                for (int i = k + 1; i < n; ++i) x[i] -= c[i] * x[k];
            }
            for (int k = n - 1; k >= 0; --k)
            {
                const double *c = lu + (size_t)k * lda;
                x[k] /= c[k];
                for (int i = 0; i < k; ++i) x[i] -= c[i] * x[k];
            }
        }
        else
        {
            for (int k = 0; k < n; ++k)
            {
                const double *c = lu + (size_t)k * lda;
                double s = x[k];
                for (int i = 0; i < k; ++i) s -= c[i] * x[i];
                x[k] = s / c[k];
            }
            for (int k = n - 1; k >= 0; --k)
            {
                const double *c = lu + (size_t)k * lda;
                double s = x[k];
                for (int i = k + 1; i < n; ++i) s -= c[i] * x[i];
                x[k] = s;
            }
            if (ipiv)
                for (int k = n - 1; k >= 0; --k)
                {
                    int p = ipiv[k] - 1;
                    double t = x[k];
                    x[k] = x[p];
                    x[p] = t;
                }
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* Complex single precision                                                              */
/* ------------------------------------------------------------------------------------ */

typedef struct
{
    float re, im;
} cf;

/* Pivot score of src/utils/complex_solver_kernel.hpp:47-63: an unsigned key that orders
 * by |re|+|im|, extends the exponent by one bit when that sum overflows, gives NaN the
 * inactive score 0 and both signed zeros the score 1. */
ORC_EXPORT uint32_t orc_cpivot_score(float re, float im)
{
    union
    {
        float f;
        uint32_t u;
    } s;
    float r = fabsf(re), i = fabsf(im);
    s.f = r + i;
    if (!(s.f >= 0.0f)) return 0u;
    if (s.f == 0.0f) return 1u;
    if (s.f > FLT_MAX)
    {
        s.f = r * 0.5f + i * 0.5f;
        return s.u + 0x00800001u;
    }
    return s.u + 1u;
}

/* Complex quotient a / b computed in double and rounded once: within one rounding of the
 * exact quotient for every finite float input (the reference's Divide is specified to 4
 * ulp, test/complex_solver/complex_solver_test.cpp:520-556). */
static cf cdiv_f(cf a, cf b)
{
    double ar = a.re, ai = a.im, br = b.re, bi = b.im;
    double q_re, q_im;
    if (fabs(br) >= fabs(bi))
    {
        double r = bi / br, den = br + bi * r;
        q_re = (ar + ai * r) / den;
        q_im = (ai - ar * r) / den;
    }
    else
    {
        double r = br / bi, den = br * r + bi;
        q_re = (ar * r + ai) / den;
        q_im = (ai * r - ar) / den;
    }
    cf q = {(float)q_re, (float)q_im};
    return q;
}

/* x - a * b with the fused order of SubMul (complex_solver_kernel.hpp:65-68). */
static cf csubmul_f(cf x, cf a, cf b)
{
    cf r;
    r.re = fmaf(a.im, b.im, fmaf(-a.re, b.re, x.re));
    r.im = fmaf(-a.im, b.re, fmaf(-a.re, b.im, x.im));
    return r;
}

/* margin (nullable): per step, relative gap of |re|+|im| between the winner and the best
 * other candidate (the tie margin a parity test filters on). */
ORC_EXPORT int orc_cgetrf(int m, int n, float *af, int lda, int *ipiv, float *margin)
{
    cf *a = (cf *)af;
    int info = 0;
    int steps = m < n ? m : n;
    for (int k = 0; k < steps; ++k)
    {
        cf *col = a + (size_t)k * lda;
        int p = k;
        if (ipiv)
        {
            uint32_t best = orc_cpivot_score(col[k].re, col[k].im);
            for (int i = k + 1; i < m; ++i)
            {
                uint32_t s = orc_cpivot_score(col[i].re, col[i].im);
                if (s > best)
                {
                    best = s;
                    p = i;
                }
            }
            if (margin)
            {
                double w = fabs((double)col[p].re) + fabs((double)col[p].im), second = 0.0;
                for (int i = k; i < m; ++i)
                {
                    double v = fabs((double)col[i].re) + fabs((double)col[i].im);
                    if (i != p && v > second) second = v;
                }
                margin[k] = w > 0.0 ? (float)((w - second) / w) : 0.0f;
            }
            ipiv[k] = p + 1;
            if (p != k)
                for (int j = 0; j < n; ++j)
                {
                    cf t = a[(size_t)j * lda + k];
                    a[(size_t)j * lda + k] = a[(size_t)j * lda + p];
                    a[(size_t)j * lda + p] = t;
                }
        }
        cf d = col[k];
        if (d.re == 0.0f && d.im == 0.0f)
        {
            if (info == 0) info = k + 1;
            continue;
        }
        for (int i = k + 1; i < m; ++i) col[i] = cdiv_f(col[i], d);
        for (int j = k + 1; j < n; ++j)
        {
            cf *cj = a + (size_t)j * lda;
            cf u = cj[k];
            for (int i = k + 1; i < m; ++i) cj[i] = csubmul_f(cj[i], col[i], u);
        }
    }
    return info;
}

static cf conj_if(cf v, int c)
{
    if (c) v.im = -v.im;
    return v;
}

/* cgetrs, column-major: trans 0 A X = B, 1 A^T X = B, 2 A^H X = B. */
ORC_EXPORT void orc_cgetrs(int trans, int n, int nrhs, const float *luf, int lda, const int *ipiv, float *bf,
                           int ldb)
{
    const cf *lu = (const cf *)luf;
    cf *b = (cf *)bf;
    int cj = trans == 2;
    for (int j = 0; j < nrhs; ++j)
    {
        cf *x = b + (size_t)j * ldb;
        if (trans == 0)
        {
            if (ipiv)
                for (int k = 0; k < n; ++k)
                {
                    int p = ipiv[k] - 1;
                    cf t = x[k];
                    x[k] = x[p];
                    x[p] = t;
                }
            for (int k = 0; k < n; ++k)
            {
                const cf *c = lu + (size_t)k * lda;
                for (int i = k + 1; i < n; ++i) x[i] = csubmul_f(x[i], c[i], x[k]);
            }
            for (int k = n - 1; k >= 0; --k)
            {
                const cf *c = lu + (size_t)k * lda;
                x[k] = cdiv_f(x[k], c[k]);
                for (int i = 0; i < k; ++i) x[i] = csubmul_f(x[i], c[i], x[k]);
            }
        }
        else
        {
            for (int k = 0; k < n; ++k)
            {
                const cf *c = lu + (size_t)k * lda;
                cf s = x[k];
                for (int i = 0; i < k; ++i) s = csubmul_f(s, conj_if(c[i], cj), x[i]);
                x[k] = cdiv_f(s, conj_if(c[k], cj));
            }
            for (int k = n - 1; k >= 0; --k)
            {
                const cf *c = lu + (size_t)k * lda;
                cf s = x[k];
                for (int i = k + 1; i < n; ++i) s = csubmul_f(s, conj_if(c[i], cj), x[i]);
                x[k] = s;
            }
            if (ipiv)
                for (int k = n - 1; k >= 0; --k)
                {
                    int p = ipiv[k] - 1;
                    cf t = x[k];
                    x[k] = x[p];
                    x[p] = t;
                }
        }
    }
}

ORC_EXPORT int orc_cgetri(int n, float *af, int lda, const int *ipiv)
{
    cf *a = (cf *)af;
    for (int k = 0; k < n; ++k)
        if (a[(size_t)k * lda + k].re == 0.0f && a[(size_t)k * lda + k].im == 0.0f) return k + 1;
    cf one = {1.0f, 0.0f}, zero = {0.0f, 0.0f};
    for (int j = 0; j < n; ++j)
    {
        cf *c = a + (size_t)j * lda;
        c[j] = cdiv_f(one, c[j]);
        cf ajj = {-c[j].re, -c[j].im};
        for (int k = 0; k < j; ++k)
        {
            cf t = c[k];
            const cf *ck = a + (size_t)k * lda;
            cf nt = {-t.re, -t.im};
            for (int i = 0; i < k; ++i) c[i] = csubmul_f(c[i], ck[i], nt);
            c[k] = csubmul_f(zero, ck[k], nt);
        }
        for (int i = 0; i < j; ++i)
        {
            cf v = c[i];
            c[i].re = v.re * ajj.re - v.im * ajj.im;
            c[i].im = v.re * ajj.im + v.im * ajj.re;
        }
    }
    cf *work = (cf *)malloc(sizeof(cf) * (size_t)(n > 0 ? n : 1));
    for (int j = n - 1; j >= 0; --j)
    {
        cf *c = a + (size_t)j * lda;
        for (int i = j + 1; i < n; ++i)
        {
            work[i] = c[i];
            c[i] = zero;
        }
        for (int k = j + 1; k < n; ++k)
        {
            const cf *ck = a + (size_t)k * lda;
            for (int i = 0; i < n; ++i) c[i] = csubmul_f(c[i], ck[i], work[k]);
        }
    }
    free(work);
    if (ipiv)
        for (int j = n - 2; j >= 0; --j)
        {
            int p = ipiv[j] - 1;
            if (p != j)
                for (int i = 0; i < n; ++i)
                {
                    cf t = a[(size_t)j * lda + i];
                    a[(size_t)j * lda + i] = a[(size_t)p * lda + i];
                    a[(size_t)p * lda + i] = t;
                }
        }
    return 0;
}

ORC_EXPORT int orc_cmatinv(int n, const float *af, int lda, float *cfp, int ldc, int *ipiv_out)
{
    const cf *a = (const cf *)af;
    cf *c = (cf *)cfp;
    int *piv = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; ++j) memcpy(c + (size_t)j * ldc, a + (size_t)j * lda, sizeof(cf) * (size_t)n);
    int info = orc_cgetrf(n, n, cfp, ldc, piv, NULL);
    if (info == 0) orc_cgetri(n, cfp, ldc, piv);
    if (ipiv_out) memcpy(ipiv_out, piv, sizeof(int) * (size_t)n);
    free(piv);
    return info;
}

/* ------------------------------------------------------------------------------------ */
/* Complex double precision (FP64 oracle)                                                */
/* ------------------------------------------------------------------------------------ */

typedef struct
{
    double re, im;
} cd;

static cd cdiv_d(cd a, cd b)
{
    cd q;
    if (fabs(b.re) >= fabs(b.im))
    {
        double r = b.im / b.re, den = b.re + b.im * r;
        q.re = (a.re + a.im * r) / den;
        q.im = (a.im - a.re * r) / den;
    }
    else
    {
        double r = b.re / b.im, den = b.re * r + b.im;
        q.re = (a.re * r + a.im) / den;
        q.im = (a.im * r - a.re) / den;
    }
    return q;
}

ORC_EXPORT int orc_zgetrf(int m, int n, double *af, int lda, int *ipiv)
{
    cd *a = (cd *)af;
    int info = 0;
    int steps = m < n ? m : n;
    for (int k = 0; k < steps; ++k)
    {
        cd *col = a + (size_t)k * lda;
        int p = k;
        if (ipiv)
        {
            double best = fabs(col[k].re) + fabs(col[k].im);
            for (int i = k + 1; i < m; ++i)
            {
                double s = fabs(col[i].re) + fabs(col[i].im);
                if (s > best)
                {
                    best = s;
                    p = i;
                }
            }
            ipiv[k] = p + 1;
            if (p != k)
                for (int j = 0; j < n; ++j)
                {
                    cd t = a[(size_t)j * lda + k];
                    a[(size_t)j * lda + k] = a[(size_t)j * lda + p];
                    a[(size_t)j * lda + p] = t;
                }
        }
        cd d = col[k];
        if (d.re == 0.0 && d.im == 0.0)
        {
            if (info == 0) info = k + 1;
            continue;
        }
        for (int i = k + 1; i < m; ++i) col[i] = cdiv_d(col[i], d);
        for (int j = k + 1; j < n; ++j)
        {
            cd *cj = a + (size_t)j * lda;
            cd u = cj[k];
            for (int i = k + 1; i < m; ++i)
            {
                cj[i].re -= col[i].re * u.re - col[i].im * u.im;
                cj[i].im -= col[i].re * u.im + col[i].im * u.re;
            }
        }
    }
    return info;
}
