/*
 * hprqp_oracle.c -- CPU FP64 restatement of the HPR-QP reference CPU path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that parity tests,
 * the smoke check and the benchmark's CPU baseline compare against.  The
 * product library (libhprqp) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  The reference is
 * Julia (HPRQP.jl); its CPU path runs single-threaded SparseArrays mul! and
 * @simd loops, which this file restates as plain sequential C loops.
 *
 * Parity status of each part (what pins it) is listed in the tests under
 * tests/test_oracle_*.py.  scaling and compute_M_norm, which only shape the
 * trajectory, are pinned on their own through oracle_scale / oracle_mnorm
 * (tests/test_oracle_pinning_cpu.py): the Ruiz entry bound and
 * equilibration, the Pock-Chambolle spectral bound, the scaling bookkeeping,
 * and the M-norm as a sum of squares, each check shown to fail on a copy of
 * this file with a planted slip.  One part is "parity unpinned": the power-method
 * start vector.  The reference draws it from Julia's MersenneTwister(1)
 * randn stream (src/utils.jl:1339-1343), which cannot be regenerated without
 * Julia; this oracle uses its own counter-based N(0,1) stream instead
 * (oracle_randn).  The power method's answer is pinned to its 1e-4
 * tolerance against LAPACK eigenvalues, not to Julia's bit pattern.
 */
#include "hprqp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* small helpers                                                       */
/* ------------------------------------------------------------------ */

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static double *dalloc(int64_t n) {
    double *p = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    return p;
}

static double dot(int64_t n, const double *a, const double *b) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static double nrm2(int64_t n, const double *a) {
    /* BLAS dnrm2 semantics: scaled sum of squares (overflow-safe). */
    double scale = 0.0, ssq = 1.0;
    for (int64_t i = 0; i < n; ++i) {
        if (a[i] != 0.0) {
            double ax = fabs(a[i]);
            if (scale < ax) {
                ssq = 1.0 + ssq * (scale / ax) * (scale / ax);
                scale = ax;
            } else {
                ssq += (ax / scale) * (ax / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}

static double nrminf(int64_t n, const double *a) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double v = fabs(a[i]);
        if (v > s) s = v;
    }
    return s;
}

static double fmax_j(double a, double b) { return a > b ? a : b; }
static double fmin_j(double a, double b) { return a < b ? a : b; }
/* Julia clamp(x, lo, hi) = ifelse(x > hi, hi, ifelse(x < lo, lo, x)) */
static double clamp_j(double x, double lo, double hi) {
    return x > hi ? hi : (x < lo ? lo : x);
}
/* Julia min(max(z, l), u) for box projection */
static double proj(double z, double lo, double hi) {
    return fmin_j(fmax_j(z, lo), hi);
}

/* ------------------------------------------------------------------ */
/* random stream (power-method start vector)                           */
/* ------------------------------------------------------------------ */

static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double u01(uint64_t seed, uint64_t k) {
    uint64_t h = mix64(mix64(seed + 0x9E3779B97F4A7C15ULL) ^
                       (k * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

void oracle_randn(uint64_t seed, int64_t count, double *out) {
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < count; ++i) {
        double a = u01(seed, 2 * (uint64_t)i);
        double b = u01(seed, 2 * (uint64_t)i + 1);
        out[i] = sqrt(-2.0 * log(1.0 - a)) * cos(two_pi * b);
    }
}

/* ------------------------------------------------------------------ */
/* CSC storage                                                         */
/* ------------------------------------------------------------------ */

typedef struct {
    int32_t nrow, ncol;
    int64_t nnz;
    int64_t *cp;
    int32_t *ri;
    double *v;
} csc_t;

static csc_t csc_copy(int32_t nrow, int32_t ncol, const int64_t *cp,
                      const int32_t *ri, const double *v) {
    csc_t M;
    M.nrow = nrow;
    M.ncol = ncol;
    M.nnz = cp[ncol];
    M.cp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncol + 1));
    M.ri = (int32_t *)malloc(sizeof(int32_t) * (size_t)(M.nnz > 0 ? M.nnz : 1));
    M.v = dalloc(M.nnz);
    memcpy(M.cp, cp, sizeof(int64_t) * (size_t)(ncol + 1));
    if (M.nnz > 0) {
        memcpy(M.ri, ri, sizeof(int32_t) * (size_t)M.nnz);
        memcpy(M.v, v, sizeof(double) * (size_t)M.nnz);
    }
    return M;
}

static void csc_free(csc_t *M) {
    free(M->cp);
    free(M->ri);
    free(M->v);
    memset(M, 0, sizeof(*M));
}

/* transpose (materialises A' as its own CSC, like the reference's qp.AT = A') */
static csc_t csc_transpose(const csc_t *A) {
    csc_t T;
    T.nrow = A->ncol;
    T.ncol = A->nrow;
    T.nnz = A->nnz;
    T.cp = (int64_t *)calloc((size_t)T.ncol + 1, sizeof(int64_t));
    T.ri = (int32_t *)malloc(sizeof(int32_t) * (size_t)(T.nnz > 0 ? T.nnz : 1));
    T.v = dalloc(T.nnz);
    for (int64_t k = 0; k < A->nnz; ++k) T.cp[A->ri[k] + 1]++;
    for (int32_t i = 0; i < T.ncol; ++i) T.cp[i + 1] += T.cp[i];
    int64_t *next = (int64_t *)malloc(sizeof(int64_t) * (size_t)(T.ncol + 1));
    memcpy(next, T.cp, sizeof(int64_t) * (size_t)(T.ncol + 1));
    for (int32_t j = 0; j < A->ncol; ++j)
        for (int64_t k = A->cp[j]; k < A->cp[j + 1]; ++k) {
            int64_t p = next[A->ri[k]]++;
            T.ri[p] = j;
            T.v[p] = A->v[k];
        }
    free(next);
    return T;
}

/* mul!(y, M, x) for SparseMatrixCSC: zero y, then scatter column by column
 * (Julia SparseArrays, called from src/kernels.jl:2122,2504-2509 etc.). */
void oracle_csc_mul(int32_t nrow, int32_t ncol, const int64_t *cp,
                    const int32_t *ri, const double *v, const double *x,
                    double *y) {
    for (int32_t i = 0; i < nrow; ++i) y[i] = 0.0;
    for (int32_t j = 0; j < ncol; ++j) {
        double xj = x[j];
        for (int64_t k = cp[j]; k < cp[j + 1]; ++k) y[ri[k]] += v[k] * xj;
    }
}

static void csc_mul(const csc_t *M, const double *x, double *y) {
    oracle_csc_mul(M->nrow, M->ncol, M->cp, M->ri, M->v, x, y);
}

/* ------------------------------------------------------------------ */
/* Q operators (CPU Qmap! of src/Q_operators/)                         */
/* ------------------------------------------------------------------ */

typedef struct {
    int32_t kind;
    int32_t n;
    csc_t Q;           /* sparse */
    csc_t LA, LAT;     /* LASSO data matrix and its transpose */
    double *ltemp;     /* LASSO temp (rows) */
    int32_t nq;        /* QAP */
    const double *qa, *qb, *qs, *qt;
    double *qtemp;     /* QAP temp (nq*nq) */
    int32_t lr_k;      /* diagonal plus low rank */
    const double *lr_U, *lr_d;
    double *lr_t;
} qop_t;

/* C = alpha * X * Y + beta * C, dense column-major nq x nq (BLAS dgemm). */
static void gemm_nn(int32_t nq, double alpha, const double *X, const double *Y,
                    double beta, double *C) {
    for (int32_t j = 0; j < nq; ++j) {
        for (int32_t i = 0; i < nq; ++i) {
            double s = 0.0;
            for (int32_t k = 0; k < nq; ++k)
                s += X[i + (int64_t)k * nq] * Y[k + (int64_t)j * nq];
            double prev = beta == 0.0 ? 0.0 : beta * C[i + (int64_t)j * nq];
            C[i + (int64_t)j * nq] = alpha * s + prev;
        }
    }
}

/* src/Q_operators/sparse_matrix_operator.jl:45-51 (CSC mul!),
 * src/Q_operators/LASSO_operator.jl:112-117 (A*x then A'*temp),
 * src/Q_operators/QAP_operator.jl:65-76 (four dgemm calls). */
static void qmap(qop_t *q, const double *x, double *Qx) {
    if (q->kind == ORACLE_Q_SPARSE) {
        csc_mul(&q->Q, x, Qx);
    } else if (q->kind == ORACLE_Q_LASSO) {
        csc_mul(&q->LA, x, q->ltemp);
        csc_mul(&q->LAT, q->ltemp, Qx);
    } else if (q->kind == ORACLE_Q_QAP) {
        int32_t nq = q->nq;
        gemm_nn(nq, 1.0, q->qa, x, 0.0, q->qtemp);   /* TMP = A*X */
        gemm_nn(nq, 2.0, q->qtemp, q->qb, 0.0, Qx);  /* QX = 2*TMP*B */
        gemm_nn(nq, -2.0, q->qs, x, 1.0, Qx);        /* QX -= 2*S*X */
        gemm_nn(nq, -2.0, x, q->qt, 1.0, Qx);        /* QX -= 2*X*T */
    } else {
        /* Qx = diag .* x + U * (U' * x) (docs/src/guide/q_operators_overview.md:135-137) */
        int32_t n = q->n, k = q->lr_k;
        for (int32_t j = 0; j < k; ++j) q->lr_t[j] = dot(n, q->lr_U + (int64_t)j * n, x);
        for (int32_t i = 0; i < n; ++i) Qx[i] = q->lr_d[i] * x[i];
        for (int32_t j = 0; j < k; ++j) {
            const double tj = q->lr_t[j];
            const double *col = q->lr_U + (int64_t)j * n;
            for (int32_t i = 0; i < n; ++i) Qx[i] += col[i] * tj;
        }
    }
}

static int qop_build(int32_t n, const oracle_qop *in, qop_t *q) {
    memset(q, 0, sizeof(*q));
    q->kind = in->kind;
    q->n = n;
    if (in->kind == ORACLE_Q_SPARSE) {
        q->Q = csc_copy(n, n, in->colptr, in->rowidx, in->val);
    } else if (in->kind == ORACLE_Q_LASSO) {
        q->LA = csc_copy(in->rows, n, in->l_colptr, in->l_rowidx, in->l_val);
        q->LAT = csc_transpose(&q->LA);
        q->ltemp = dalloc(in->rows);
    } else if (in->kind == ORACLE_Q_QAP) {
        if ((int64_t)in->nq * in->nq != n) return -1;
        q->nq = in->nq;
        q->qa = in->qa;
        q->qb = in->qb;
        q->qs = in->qs;
        q->qt = in->qt;
        q->qtemp = dalloc((int64_t)n);
    } else if (in->kind == ORACLE_Q_LOWRANK) {
        q->lr_k = in->lr_k;
        q->lr_U = in->lr_U;
        q->lr_d = in->lr_d;
        q->lr_t = dalloc(in->lr_k);
    } else {
        return -1;
    }
    return 0;
}

static void qop_free(qop_t *q) {
    if (q->kind == ORACLE_Q_SPARSE) csc_free(&q->Q);
    if (q->kind == ORACLE_Q_LASSO) {
        csc_free(&q->LA);
        csc_free(&q->LAT);
        free(q->ltemp);
    }
    if (q->kind == ORACLE_Q_QAP) free(q->qtemp);
    if (q->kind == ORACLE_Q_LOWRANK) free(q->lr_t);
}

int oracle_qmap(int32_t n, const oracle_qop *in, const double *x, double *Qx) {
    qop_t q;
    if (qop_build(n, in, &q) != 0) return -1;
    qmap(&q, x, Qx);
    qop_free(&q);
    return 0;
}

static int qop_is_operator(const qop_t *q) { return q->kind != ORACLE_Q_SPARSE; }
static int q_nonempty(const qop_t *q) {
    return q->kind != ORACLE_Q_SPARSE || q->Q.nnz > 0;
}

/* ------------------------------------------------------------------ */
/* scaling! CPU branch (src/utils.jl:350-804)                           */
/* ------------------------------------------------------------------ */

typedef struct {
    double *row_norm, *col_norm;
    double b_scale, c_scale;
    double norm_b, norm_c, norm_b_org, norm_c_org;
} scaling_t;

static double bound_inf_norm(int32_t m, const double *AL, const double *AU,
                             int two) {
    /* max.(abs.(AL_nInf), abs.(AU_nInf)) with infinite entries set to 0,
     * then an inf-norm or a 2-norm (src/utils.jl:404-408, 784-789) */
    double *t = dalloc(m);
    for (int32_t i = 0; i < m; ++i) {
        double a = AL[i] == -INFINITY ? 0.0 : AL[i];
        double b = AU[i] == INFINITY ? 0.0 : AU[i];
        t[i] = fmax_j(fabs(a), fabs(b));
    }
    double r = two ? nrm2(m, t) : nrminf(m, t);
    free(t);
    return r;
}

/* (Q + Q')/2 over the union pattern; row indices sorted within columns.
 * Entries present on one side only become half their value. */
static csc_t csc_sym_avg(const csc_t *Q) {
    csc_t T = csc_transpose(Q); /* transpose emits sorted row indices */
    csc_t R;
    R.nrow = Q->nrow;
    R.ncol = Q->ncol;
    R.cp = (int64_t *)calloc((size_t)Q->ncol + 1, sizeof(int64_t));
    R.ri = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * Q->nnz + 1));
    R.v = dalloc(2 * Q->nnz + 1);
    int64_t p = 0;
    for (int32_t j = 0; j < Q->ncol; ++j) {
        int64_t a = Q->cp[j], ae = Q->cp[j + 1], b = T.cp[j], be = T.cp[j + 1];
        while (a < ae || b < be) {
            int32_t ra = a < ae ? Q->ri[a] : 0x7fffffff;
            int32_t rb = b < be ? T.ri[b] : 0x7fffffff;
            if (ra == rb) {
                R.ri[p] = ra;
                R.v[p++] = (Q->v[a++] + T.v[b++]) / 2.0;
            } else if (ra < rb) {
                R.ri[p] = ra;
                R.v[p++] = (Q->v[a++] + 0.0) / 2.0;
            } else {
                R.ri[p] = rb;
                R.v[p++] = (0.0 + T.v[b++]) / 2.0;
            }
        }
        R.cp[j + 1] = p;
    }
    R.nnz = p;
    csc_free(&T);
    return R;
}

/* DR * M * DC for CSC M: entry (i,j) becomes (dr[i] * v) * dc[j] with
 * dr = 1 ./ row scale, dc = 1 ./ col scale (sparse diagonal products). */
static void csc_scale(csc_t *M, const double *inv_r, const double *inv_c) {
    for (int32_t j = 0; j < M->ncol; ++j)
        for (int64_t k = M->cp[j]; k < M->cp[j + 1]; ++k)
            M->v[k] = (inv_r[M->ri[k]] * M->v[k]) * inv_c[j];
}

static void scaling(int32_t n, int32_t m, qop_t *q, csc_t *A, double *c,
                    double *AL, double *AU, double *l, double *u,
                    const oracle_params *p, scaling_t *sc) {
    sc->row_norm = dalloc(m);
    sc->col_norm = dalloc(n);
    for (int32_t i = 0; i < m; ++i) sc->row_norm[i] = 1.0;
    for (int32_t j = 0; j < n; ++j) sc->col_norm[j] = 1.0;
    sc->b_scale = 1.0;
    sc->c_scale = 1.0;
    sc->norm_b_org = m > 0 ? bound_inf_norm(m, AL, AU, 0) : 0.0;
    sc->norm_c_org = nrminf(n, c);
    if (qop_is_operator(q)) {
        /* operator Q skips all scaling (src/utils.jl:364-406) */
        sc->norm_b = m > 0 ? bound_inf_norm(m, AL, AU, 1) : 0.0;
        sc->norm_c = nrm2(n, c);
        return;
    }
    double *tr = dalloc(m), *tc = dalloc(n), *ir = dalloc(m), *ic = dalloc(n);
    int passes_ruiz = p->use_Ruiz_scaling ? 10 : 0;
    int total = passes_ruiz + (p->use_Pock_Chambolle_scaling ? 1 : 0);
    for (int pass = 0; pass < total; ++pass) {
        int ruiz = pass < passes_ruiz;
        /* column statistic of |A| and |Q| (src/utils.jl:468-475, 595-600) */
        for (int32_t j = 0; j < n; ++j) {
            double sa = 0.0, sq = 0.0;
            for (int64_t k = A->cp[j]; k < A->cp[j + 1]; ++k)
                sa = ruiz ? fmax_j(sa, fabs(A->v[k])) : sa + fabs(A->v[k]);
            for (int64_t k = q->Q.cp[j]; k < q->Q.cp[j + 1]; ++k)
                sq = ruiz ? fmax_j(sq, fabs(q->Q.v[k])) : sq + fabs(q->Q.v[k]);
            double t = ruiz ? sqrt(fmax_j(sa, sq)) : sqrt(sa + sq);
            tc[j] = t == 0.0 ? 1.0 : t;
        }
        /* row statistic of |A| (src/utils.jl:486-490, 610-614) */
        if (m > 0) {
            for (int32_t i = 0; i < m; ++i) tr[i] = 0.0;
            for (int32_t j = 0; j < n; ++j)
                for (int64_t k = A->cp[j]; k < A->cp[j + 1]; ++k) {
                    double av = fabs(A->v[k]);
                    int32_t i = A->ri[k];
                    tr[i] = ruiz ? fmax_j(tr[i], av) : tr[i] + av;
                }
            for (int32_t i = 0; i < m; ++i) {
                double t = sqrt(tr[i]);
                tr[i] = t == 0.0 ? 1.0 : t;
            }
        }
        for (int32_t i = 0; i < m; ++i) sc->row_norm[i] *= tr[i];
        for (int32_t j = 0; j < n; ++j) sc->col_norm[j] *= tc[j];
        for (int32_t i = 0; i < m; ++i) ir[i] = 1.0 / tr[i];
        for (int32_t j = 0; j < n; ++j) ic[j] = 1.0 / tc[j];
        /* Q = DC*Q*DC, A = DR*A*DC (src/utils.jl:510-512, 538-541) */
        csc_scale(&q->Q, ic, ic);
        if (m > 0) csc_scale(A, ir, ic);
        /* c ./= d, AL/AU ./= r, l/u .*= d (src/utils.jl:568-575) */
        for (int32_t j = 0; j < n; ++j) c[j] /= tc[j];
        for (int32_t i = 0; i < m; ++i) {
            AL[i] /= tr[i];
            AU[i] /= tr[i];
        }
        for (int32_t j = 0; j < n; ++j) {
            l[j] *= tc[j];
            u[j] *= tc[j];
        }
    }
    free(tr);
    free(tc);
    free(ir);
    free(ic);
    /* Q = (Q + Q')/2 (src/utils.jl:778-780): the pattern is symmetric after
     * model formulation, so each entry becomes (Q_ij + Q_ji) / 2. */
    if (q->Q.nnz > 0) {
        csc_t S = csc_sym_avg(&q->Q);
        csc_free(&q->Q);
        q->Q = S;
    }
    sc->norm_b = m > 0 ? bound_inf_norm(m, AL, AU, 1) : 0.0;
    sc->norm_c = nrm2(n, c);
}

/* ------------------------------------------------------------------ */
/* power method (src/utils.jl:1335-1385, 1419-1461)                     */
/* ------------------------------------------------------------------ */

static double power_A(const csc_t *A, const csc_t *AT, int64_t max_it,
                      double tol, int64_t *its) {
    int32_t m = A->nrow, n = A->ncol;
    double *z = dalloc(m), *q = dalloc(m), *ATq = dalloc(n);
    oracle_randn(1, m, z);
    for (int32_t i = 0; i < m; ++i) z[i] += 1e-8;
    double lambda_max = 1.0, err = 1.0;
    int64_t i;
    for (i = 1; i <= max_it; ++i) {
        double nq = nrm2(m, z);
        for (int32_t k = 0; k < m; ++k) q[k] = z[k] / nq;
        csc_mul(AT, q, ATq);
        csc_mul(A, ATq, z);
        lambda_max = dot(m, q, z);
        for (int32_t k = 0; k < m; ++k) q[k] = z[k] - lambda_max * q[k];
        err = nrm2(m, q) / (nrm2(m, z) + lambda_max);
        if (err < tol) break;
    }
    if (its) *its = i > max_it ? max_it : i;
    free(z);
    free(q);
    free(ATq);
    return lambda_max;
}

static double power_Q(qop_t *Q, int32_t n, int64_t max_it, double tol,
                      int64_t *its) {
    double *z = dalloc(n), *q = dalloc(n);
    oracle_randn(1, n, z);
    for (int32_t i = 0; i < n; ++i) z[i] += 1e-8;
    double lambda_max = 1.0, err = 1.0;
    int64_t i;
    for (i = 1; i <= max_it; ++i) {
        double nq = nrm2(n, z);
        for (int32_t k = 0; k < n; ++k) q[k] = z[k] / nq;
        qmap(Q, q, z);
        lambda_max = dot(n, q, z);
        for (int32_t k = 0; k < n; ++k) q[k] = z[k] - lambda_max * q[k];
        err = nrm2(n, q) / (nrm2(n, z) + lambda_max);
        if (err < tol) break;
    }
    if (its) *its = i > max_it ? max_it : i;
    free(z);
    free(q);
    return lambda_max;
}

double oracle_power_iteration(int32_t n, int32_t m, const oracle_qop *qin,
                              int64_t a_nnz, const int64_t *a_colptr,
                              const int32_t *a_rowidx, const double *a_val,
                              int which, int64_t max_it, double tol,
                              int64_t *iters_out) {
    (void)a_nnz;
    if (which == 0) {
        csc_t A = csc_copy(m, n, a_colptr, a_rowidx, a_val);
        csc_t AT = csc_transpose(&A);
        double r = power_A(&A, &AT, max_it, tol, iters_out);
        csc_free(&A);
        csc_free(&AT);
        return r;
    }
    qop_t q;
    if (qop_build(n, qin, &q) != 0) return NAN;
    double r = power_Q(&q, n, max_it, tol, iters_out);
    qop_free(&q);
    return r;
}

/* ------------------------------------------------------------------ */
/* golden-section searches                                             */
/* ------------------------------------------------------------------ */

/* src/kernels.jl:1449-1486 */
static double f_dev(double x, double a, double b, double c, double d) {
    return a * x + b / x + c * x * x / (1 + d * x);
}

double oracle_golden(double ap, double bp, double cp, double dp, double lo,
                     double hi, double tol, int maxiter) {
    const double phi = (sqrt(5.0) - 1.0) / 2.0;
    double a = lo, b = hi;
    double c = b - phi * (b - a);
    double d = a + phi * (b - a);
    double fc = f_dev(c, ap, bp, cp, dp);
    double fd = f_dev(d, ap, bp, cp, dp);
    for (int i = 1; i <= maxiter; ++i) {
        if (fd < fc) {
            a = c;
            c = d;
            fc = fd;
            d = a + phi * (b - a);
            fd = f_dev(d, ap, bp, cp, dp);
        } else {
            b = d;
            d = c;
            fd = fc;
            c = b - phi * (b - a);
            fc = f_dev(c, ap, bp, cp, dp);
        }
        if ((b - a) < tol) break;
    }
    return 0.5 * (a + b);
}

/* src/unified_operations.jl:348-386 (CPU method) */
double oracle_golden_Q_diag(double a, double b, int32_t n, const double *Q,
                            const double *c, const double *d, double lo,
                            double hi, double tol, int maxiter) {
    const double phi = (sqrt(5.0) - 1.0) / 2.0;
    double *t = dalloc(n);
#define FQ(xx)                                                        \
    ({                                                                \
        double _x = (xx);                                             \
        for (int32_t _i = 0; _i < n; ++_i) t[_i] = d[_i] / (1.0 + _x * Q[_i]); \
        a * _x + b / _x + _x * _x * dot(n, c, t);                     \
    })
    double x1 = hi - phi * (hi - lo);
    double x2 = lo + phi * (hi - lo);
    double f1 = FQ(x1), f2 = FQ(x2);
    int iter = 0;
    while (fabs(hi - lo) > tol * fmax_j(1.0, fabs(lo)) && iter < maxiter) {
        if (f1 > f2) {
            lo = x1;
            x1 = x2;
            f1 = f2;
            x2 = lo + phi * (hi - lo);
            f2 = FQ(x2);
        } else {
            hi = x2;
            x2 = x1;
            f2 = f1;
            x1 = hi - phi * (hi - lo);
            f1 = FQ(x1);
        }
        iter++;
    }
#undef FQ
    free(t);
    return (f1 < f2) ? x1 : x2;
}

/* ------------------------------------------------------------------ */
/* workspace (src/structs.jl:462-523, src/algorithm.jl:760-885)          */
/* ------------------------------------------------------------------ */

typedef struct {
    int32_t n, m;
    double *w, *w_bar, *dw, *x, *x_hat, *x_bar, *dx, *y, *y_bar, *dy, *z_bar;
    double *Rp, *Rd, *Ax, *ATy, *ATy_bar, *ATdy, *QATdy, *s, *Qw, *Qw_bar,
        *Qx, *dQw, *last_x, *last_y, *last_Qw, *last_w, *last_ATy, *tempv;
    double *diag_Q, *fact1, *fact2, *fact, *fact_M, *lambda;
    const double *c, *l, *u, *AL, *AU;
    csc_t *A, *AT;
    qop_t *Q;
    double sigma, lambda_max_A, lambda_max_Q;
    int Q_is_diag, to_check, noC;
} ws_t;

typedef struct {
    int restart_flag, first_restart;
    double last_gap, current_gap, save_gap;
    int64_t inner;
    int64_t sufficient, necessary, long_, times;
    double weighted_norm, best_gap, best_sigma;
} restart_t;

typedef struct {
    double err_Rp, err_Rd, kkt, primal_obj, rel_gap, dual_obj;
} resid_t;

static void ws_alloc(ws_t *ws, int32_t n, int32_t m) {
    memset(ws, 0, sizeof(*ws));
    ws->n = n;
    ws->m = m;
    double **nv[] = {&ws->w, &ws->w_bar, &ws->dw, &ws->x, &ws->x_hat,
                     &ws->x_bar, &ws->dx, &ws->z_bar, &ws->Rd, &ws->ATy,
                     &ws->ATy_bar, &ws->ATdy, &ws->QATdy, &ws->Qw,
                     &ws->Qw_bar, &ws->Qx, &ws->dQw, &ws->last_x,
                     &ws->last_Qw, &ws->last_w, &ws->last_ATy, &ws->tempv,
                     &ws->fact1, &ws->fact2, &ws->fact, &ws->fact_M,
                     &ws->lambda};
    for (size_t k = 0; k < sizeof(nv) / sizeof(nv[0]); ++k) *nv[k] = dalloc(n);
    double **mv[] = {&ws->y, &ws->y_bar, &ws->dy, &ws->Rp, &ws->Ax, &ws->s,
                     &ws->last_y};
    for (size_t k = 0; k < sizeof(mv) / sizeof(mv[0]); ++k) *mv[k] = dalloc(m);
}

static void ws_free(ws_t *ws) {
    double *all[] = {ws->w, ws->w_bar, ws->dw, ws->x, ws->x_hat, ws->x_bar,
                     ws->dx, ws->z_bar, ws->Rd, ws->ATy, ws->ATy_bar,
                     ws->ATdy, ws->QATdy, ws->Qw, ws->Qw_bar, ws->Qx,
                     ws->dQw, ws->last_x, ws->last_Qw, ws->last_w,
                     ws->last_ATy, ws->tempv, ws->fact1, ws->fact2,
                     ws->fact, ws->fact_M, ws->lambda, ws->y, ws->y_bar,
                     ws->dy, ws->Rp, ws->Ax, ws->s, ws->last_y, ws->diag_Q};
    for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); ++k) free(all[k]);
}

/* unified_update_Q_factors! (src/unified_operations.jl:445-475), called as
 * (ws.fact2, ws.fact, ws.fact1, ws.fact_M, ...) from src/algorithm.jl:1949 */
static void update_Q_factors(ws_t *ws) {
    double s = ws->sigma, s2 = s * s;
    for (int32_t i = 0; i < ws->n; ++i) {
        double f2 = 1.0 / (1.0 + s * ws->diag_Q[i]);
        ws->fact2[i] = f2;
        ws->fact[i] = s * f2;
        ws->fact1[i] = s * ws->diag_Q[i] * f2;
        ws->fact_M[i] = s2 * f2;
    }
}

/* ------------------------------------------------------------------ */
/* iteration steps (src/kernels.jl CPU loops)                          */
/* ------------------------------------------------------------------ */

/* update_zxw1_cpu! (src/kernels.jl:2301-2397) */
static void step_zxw1(ws_t *ws, double h1, double h2) {
    qmap(ws->Q, ws->w, ws->Qw);
    double f2s = 1.0 / (1.0 + ws->sigma * ws->lambda_max_Q);
    double f1s = 1.0 - f2s;
    double sigma = ws->sigma;
    for (int32_t i = 0; i < ws->n; ++i) {
        double tmp = -ws->Qw[i] + ws->ATy[i] - ws->c[i];
        double z_raw = ws->x[i] + sigma * tmp;
        double xb = proj(z_raw, ws->l[i], ws->u[i]);
        double xh = 2.0 * xb - ws->x[i];
        double xn = fma(h2, xh, h1 * ws->last_x[i]);
        double wb = ws->Q_is_diag ? fma(ws->fact1[i], ws->w[i], ws->fact2[i] * xh)
                                  : fma(f1s, ws->w[i], f2s * xh);
        if (ws->to_check) {
            ws->dx[i] = xb - ws->x[i];
            ws->x_bar[i] = xb;
            ws->z_bar[i] = (xb - z_raw) / sigma;
        }
        ws->x[i] = xn;
        ws->x_hat[i] = xh;
        ws->w_bar[i] = wb;
    }
}

/* update_y_cpu! (src/kernels.jl:2682-2743) */
static void step_y(ws_t *ws, double h1, double h2) {
    if (ws->m == 0) return;
    qmap(ws->Q, ws->w_bar, ws->Qw_bar);
    for (int32_t i = 0; i < ws->n; ++i)
        ws->tempv[i] = ws->x_hat[i] + ws->sigma * (ws->Qw[i] - ws->Qw_bar[i]);
    csc_mul(ws->A, ws->tempv, ws->Ax);
    double fact1 = ws->lambda_max_A * ws->sigma;
    double fact2 = 1.0 / fact1;
    for (int32_t i = 0; i < ws->m; ++i) {
        double yi = ws->y[i];
        double s_raw = ws->Ax[i] - fact1 * yi;
        double s_proj = proj(s_raw, ws->AL[i], ws->AU[i]);
        double yb = fact2 * (s_proj - s_raw);
        double yh = 2.0 * yb - yi;
        double yn = fma(h2, yh, h1 * ws->last_y[i]);
        if (ws->to_check) {
            ws->s[i] = s_proj;
            ws->dy[i] = yb - yi;
        }
        ws->y_bar[i] = yb;
        ws->y[i] = yn;
    }
}

/* update_w2_cpu! (src/kernels.jl:2747-2832) */
static void step_w2(ws_t *ws, double h1, double h2) {
    csc_mul(ws->AT, ws->y_bar, ws->ATy_bar);
    double fs = ws->sigma / (1.0 + ws->sigma * ws->lambda_max_Q);
    for (int32_t i = 0; i < ws->n; ++i) {
        double wi = ws->w[i], wbi = ws->w_bar[i], atyi = ws->ATy[i],
               atybi = ws->ATy_bar[i];
        double f = ws->Q_is_diag ? ws->fact[i] : fs;
        double wbn = wbi + f * (atybi - atyi);
        double wn = h1 * ws->last_w[i] + h2 * (2.0 * wbn - wi);
        double atyn = h1 * ws->last_ATy[i] + h2 * (2.0 * atybi - atyi);
        ws->w[i] = wn;
        ws->ATy[i] = atyn;
        if (ws->to_check) {
            ws->w_bar[i] = wbn;
            ws->dw[i] = wbn - wi;
            ws->ATdy[i] = atybi - atyi;
        }
    }
}

/* unified_update_zxw_cpu! noC path (src/kernels.jl:2117-2238) */
static void step_noc_zxw(ws_t *ws, double h1, double h2) {
    csc_mul(ws->AT, ws->y, ws->ATy);
    double f2s = 1.0 / (1.0 + ws->sigma * ws->lambda_max_Q);
    double f1s = 1.0 - f2s;
    double sigma = ws->sigma;
    for (int32_t i = 0; i < ws->n; ++i) {
        double wi = ws->w[i];
        double tmp = -ws->Qw[i] + ws->ATy[i] - ws->c[i];
        double z_raw = ws->x[i] + sigma * tmp;
        double xb = proj(z_raw, ws->l[i], ws->u[i]);
        double xh = 2.0 * xb - ws->x[i];
        double xn = fma(h2, xh, h1 * ws->last_x[i]);
        double wb = ws->Q_is_diag ? fma(ws->fact1[i], wi, ws->fact2[i] * xh)
                                  : fma(f1s, wi, f2s * xh);
        double whh = 2.0 * wb - wi;
        double wn = fma(h2, whh, h1 * ws->last_w[i]);
        if (ws->to_check) {
            ws->dx[i] = xb - ws->x[i];
            ws->x_bar[i] = xb;
            ws->z_bar[i] = (xb - z_raw) / sigma;
            ws->dw[i] = wb - wi;
        }
        ws->x[i] = xn;
        ws->x_hat[i] = xh;
        ws->w_bar[i] = wb;
        ws->w[i] = wn;
    }
    qmap(ws->Q, ws->w_bar, ws->Qw_bar);
    for (int32_t i = 0; i < ws->n; ++i) {
        double qw = ws->Qw[i], qwb = ws->Qw_bar[i];
        double qwh = 2.0 * qwb - qw;
        ws->Qw[i] = fma(h2, qwh, h1 * ws->last_Qw[i]);
        ws->tempv[i] = ws->x_hat[i] + sigma * (qw - qwb);
    }
}

/* unified_update_y_cpu! noC path (src/kernels.jl:2242-2298) and
 * unified_update_y_noQ_cpu! LP path (src/kernels.jl:2582-2658): identical
 * except for the input vector and the partial write set. */
static void step_y_from(ws_t *ws, const double *t, double h1, double h2,
                        int write_ybar_partial) {
    if (ws->m == 0) return;
    csc_mul(ws->A, t, ws->Ax);
    double fact1 = ws->lambda_max_A * ws->sigma;
    double fact2 = 1.0 / fact1;
    for (int32_t i = 0; i < ws->m; ++i) {
        double yi = ws->y[i];
        double s_raw = ws->Ax[i] - fact1 * yi;
        double s_proj = proj(s_raw, ws->AL[i], ws->AU[i]);
        double yb = fact2 * (s_proj - s_raw);
        double yh = 2.0 * yb - yi;
        double yn = fma(h2, yh, h1 * ws->last_y[i]);
        if (ws->to_check) {
            ws->s[i] = s_proj;
            ws->dy[i] = yb - yi;
            ws->y_bar[i] = yb;
        } else if (write_ybar_partial) {
            ws->y_bar[i] = yb;
        }
        ws->y[i] = yn;
    }
}

/* update_zxw_LASSO_cpu! (src/kernels.jl:2400-2496) */
static void step_lasso(ws_t *ws, double h1, double h2) {
    double f2s = 1.0 / (1.0 + ws->sigma * ws->lambda_max_Q);
    double f1s = 1.0 - f2s;
    qmap(ws->Q, ws->w, ws->Qw);
    double sigma = ws->sigma;
    for (int32_t i = 0; i < ws->n; ++i) {
        double wi = ws->w[i];
        double tmp = -ws->Qw[i] + ws->ATy[i] - ws->c[i];
        double z_raw = ws->x[i] + sigma * tmp;
        double ls = ws->lambda[i] * sigma;
        double shrink = fmax_j(fabs(z_raw) - ls, 0.0);
        double xb = copysign(shrink, z_raw);
        double xh = 2.0 * xb - ws->x[i];
        double xn = fma(h2, xh, h1 * ws->last_x[i]);
        double wb = fma(f1s, wi, f2s * xh);
        double wn = fma(h2, 2.0 * wb - wi, h1 * ws->last_w[i]);
        if (ws->to_check) {
            ws->dx[i] = xb - ws->x[i];
            ws->x_bar[i] = xb;
            ws->z_bar[i] = (xb - z_raw) / sigma;
            ws->dw[i] = wb - wi;
        }
        ws->x[i] = xn;
        ws->x_hat[i] = xh;
        ws->w_bar[i] = wb;
        ws->w[i] = wn;
    }
}

/* unified_update_zx_cpu! LP path (src/kernels.jl:2504-2578) */
static void step_lp_zx(ws_t *ws, double h1, double h2) {
    csc_mul(ws->AT, ws->y, ws->ATy);
    double sigma = ws->sigma;
    for (int32_t i = 0; i < ws->n; ++i) {
        double xi = ws->x[i];
        double z_raw = xi + sigma * (ws->ATy[i] - ws->c[i]);
        double xb = proj(z_raw, ws->l[i], ws->u[i]);
        double xh = 2.0 * xb - xi;
        double xn = fma(h2, xh, h1 * ws->last_x[i]);
        if (ws->to_check) {
            ws->dx[i] = xb - xi;
            ws->x_bar[i] = xb;
            ws->z_bar[i] = (xb - z_raw) / sigma;
        }
        ws->x[i] = xn;
        ws->x_hat[i] = xh;
    }
}

/* main_update_cpu! (src/algorithm.jl:945-972) */
static void main_update(ws_t *ws, const restart_t *r) {
    double h1 = 1.0 / ((double)r->inner + 2.0);
    double h2 = 1.0 - h1;
    if (ws->Q->kind == ORACLE_Q_LASSO) {
        step_lasso(ws, h1, h2);
        return;
    }
    if (ws->Q->kind == ORACLE_Q_QAP || q_nonempty(ws->Q)) {
        if (ws->noC) {
            step_noc_zxw(ws, h1, h2);
            step_y_from(ws, ws->tempv, h1, h2, 1);
        } else {
            step_zxw1(ws, h1, h2);
            step_y(ws, h1, h2);
            step_w2(ws, h1, h2);
        }
    } else {
        step_lp_zx(ws, h1, h2);
        step_y_from(ws, ws->x_hat, h1, h2, 0);
    }
}

/* ------------------------------------------------------------------ */
/* residuals (src/algorithm.jl:37-133, src/kernels.jl:114-194)          */
/* ------------------------------------------------------------------ */

static void compute_residuals(ws_t *ws, const scaling_t *sc, resid_t *res,
                              double obj_constant, int lasso, int64_t iter) {
    int32_t n = ws->n, m = ws->m;
    double bc = sc->b_scale * sc->c_scale;
    qmap(ws->Q, ws->x_bar, ws->Qx);
    double cx = dot(n, ws->c, ws->x_bar);
    double xQx = dot(n, ws->x_bar, ws->Qx);
    double zx = dot(n, ws->z_bar, ws->x_bar);
    res->primal_obj = bc * (cx + 0.5 * xQx) + obj_constant;
    res->dual_obj = bc * (-0.5 * xQx + zx) + obj_constant;
    if (m > 0) res->dual_obj += bc * dot(m, ws->y_bar, ws->s);
    if (lasso) {
        double l1 = 0.0;
        for (int32_t i = 0; i < n; ++i) l1 += fabs(ws->lambda[i] * ws->x_bar[i]);
        res->primal_obj += bc * l1;
        res->dual_obj += bc * l1;
    }
    res->rel_gap = fabs(res->primal_obj - res->dual_obj) /
                   (1.0 + fmax_j(fabs(res->primal_obj), fabs(res->dual_obj)));
    /* compute_Rd! (src/kernels.jl:114-149) */
    if (m > 0) csc_mul(ws->AT, ws->y_bar, ws->ATdy);
    for (int32_t i = 0; i < n; ++i) {
        double sf = sc->col_norm[i] * sc->c_scale;
        ws->Rd[i] = (ws->Qx[i] + ws->c[i] - ws->ATdy[i] - ws->z_bar[i]) * sf;
        ws->Qx[i] *= sf;
        ws->ATdy[i] *= sf;
    }
    double den = fmax_j(fmax_j(sc->norm_c_org, nrminf(n, ws->ATdy)), nrminf(n, ws->Qx));
    res->err_Rd = nrminf(n, ws->Rd) / (1.0 + den);
    /* compute_Rp! (src/kernels.jl:156-194) */
    if (m > 0) {
        csc_mul(ws->A, ws->x_bar, ws->Ax);
        for (int32_t i = 0; i < m; ++i) {
            double ax = ws->Ax[i];
            double corr = proj(ax, ws->AL[i], ws->AU[i]) - ax;
            double sf = sc->row_norm[i] * sc->b_scale;
            ws->Rp[i] = corr * sf;
            ws->Ax[i] = ax * sf;
        }
        res->err_Rp = nrminf(m, ws->Rp) / (1.0 + fmax_j(sc->norm_b_org, nrminf(m, ws->Ax)));
    } else {
        res->err_Rp = 0.0;
    }
    /* bound violation at iteration 0 (src/algorithm.jl:118-124) */
    if (iter == 0) {
        double v = 0.0;
        for (int32_t i = 0; i < n; ++i) {
            double lo = fmax_j(ws->l[i] - ws->x_bar[i], 0.0);
            double hi = fmax_j(ws->x_bar[i] - ws->u[i], 0.0);
            ws->dx[i] = (lo + hi) * (sc->b_scale / sc->col_norm[i]);
            double a = fabs(ws->dx[i]);
            if (a > v) v = a;
        }
        res->err_Rp = fmax_j(res->err_Rp, v);
    }
    res->kkt = fmax_j(fmax_j(res->err_Rp, res->err_Rd), res->rel_gap);
}

/* compute_M_norm! (src/algorithm.jl:689-739); terms (may be NULL) receives
 * the unclamped M_1, M_2, M_3 */
static double compute_M_norm(ws_t *ws, double *terms) {
    int32_t n = ws->n, m = ws->m;
    double M1 = 0.0, M3 = 0.0;
    double M2 = (1.0 / ws->sigma) * dot(n, ws->dx, ws->dx);
    qmap(ws->Q, ws->dw, ws->dQw);
    M2 -= 2.0 * dot(n, ws->dQw, ws->dx);
    if (m > 0) {
        M1 = ws->sigma * ws->lambda_max_A * dot(m, ws->dy, ws->dy);
        csc_mul(ws->AT, ws->dy, ws->ATdy);
        qmap(ws->Q, ws->ATdy, ws->QATdy);
        M1 -= 2.0 * ws->sigma * dot(n, ws->dQw, ws->ATdy);
        M2 += 2.0 * dot(n, ws->ATdy, ws->dx);
        if (ws->Q_is_diag) {
            for (int32_t i = 0; i < n; ++i) ws->ATdy[i] *= ws->fact_M[i];
            M3 = dot(n, ws->ATdy, ws->QATdy);
            M1 += ws->sigma * dot(n, ws->dQw, ws->dQw);
        } else {
            M3 = (ws->sigma * ws->sigma) / (1.0 + ws->sigma * ws->lambda_max_Q) *
                 dot(n, ws->ATdy, ws->QATdy);
            M1 += ws->sigma * ws->lambda_max_Q * dot(n, ws->dw, ws->dQw);
        }
    } else if (!ws->Q_is_diag) {
        M1 = ws->sigma * ws->lambda_max_Q * dot(n, ws->dw, ws->dQw);
    }
    if (terms) {
        terms[0] = M1;
        terms[1] = M2;
        terms[2] = M3;
    }
    M2 += fmax_j(M1, 0.0);
    double Mn = fmax_j(M2, 0.0) + fmax_j(M3, 0.0);
    return sqrt(Mn);
}

/* update_sigma! (src/algorithm.jl:323-423) */
static void update_sigma(const oracle_params *p, restart_t *r, ws_t *ws,
                         const resid_t *res) {
    if (p->sigma_fixed || r->restart_flag < 1) return;
    int32_t n = ws->n, m = ws->m;
    double sigma_old = ws->sigma;
    for (int32_t i = 0; i < n; ++i) {
        ws->dx[i] = ws->x_bar[i] - ws->last_x[i];
        ws->dw[i] = ws->w_bar[i] - ws->last_w[i];
    }
    qmap(ws->Q, ws->dw, ws->dQw);
    double a = 0.0, b = dot(n, ws->dx, ws->dx), c = 0.0, d = 0.0;
    if (m > 0) {
        for (int32_t i = 0; i < m; ++i) ws->dy[i] = ws->y_bar[i] - ws->last_y[i];
        for (int32_t i = 0; i < n; ++i) ws->ATdy[i] = ws->ATy_bar[i] - ws->last_ATy[i];
        qmap(ws->Q, ws->ATdy, ws->QATdy);
        a = ws->lambda_max_A * dot(m, ws->dy, ws->dy) - 2 * dot(n, ws->dQw, ws->ATdy);
    }
    if (ws->Q_is_diag) {
        double t = nrm2(n, ws->dQw);
        a += t * t;
    } else {
        a += ws->lambda_max_Q * dot(n, ws->dw, ws->dQw);
        if (m > 0) {
            c = dot(n, ws->ATdy, ws->QATdy);
            d = ws->lambda_max_Q;
        }
    }
    a = fmax_j(a, 1e-12);
    b = fmax_j(b, 1e-12);
    double est;
    if (ws->Q_is_diag) {
        est = m > 0 ? oracle_golden_Q_diag(a, b, n, ws->diag_Q, ws->ATdy, ws->QATdy,
                                           1e-12, 1e12, 1e-13, 200)
                    : sqrt(b / a);
    } else {
        est = m > 0 ? oracle_golden(a, b, c, d, 1e-12, 1e12, 1e-13, 200) : sqrt(b / a);
    }
    double fact = exp(-0.05 * (r->current_gap / r->best_gap));
    double temp_1 = fmax_j(fmin_j(res->err_Rd, res->err_Rp), fmin_j(res->rel_gap, r->current_gap));
    double cand = exp(fact * log(est) + (1 - fact) * log(r->best_sigma));
    double kappa;
    if (temp_1 > 9e-10) {
        kappa = 1.0;
    } else if (temp_1 > 5e-10) {
        kappa = clamp_j(sqrt(res->err_Rd / res->err_Rp), 1e-2, 100.0);
    } else {
        kappa = clamp_j(res->err_Rd / res->err_Rp, 1e-2, 100.0);
    }
    ws->sigma = kappa * cand;
    if (ws->Q_is_diag && fabs(sigma_old - ws->sigma) > 1e-15) update_Q_factors(ws);
}

/* check_restart (src/algorithm.jl:441-492) */
static void check_restart(restart_t *r, int64_t iter, int64_t check_iter,
                          double sigma) {
    r->restart_flag = 0;
    if (r->first_restart) {
        if (iter == check_iter) {
            r->first_restart = 0;
            r->restart_flag = 1;
            r->weighted_norm = r->current_gap;
            r->best_gap = r->current_gap;
            r->best_sigma = sigma;
        }
    } else if (iter % check_iter == 0) {
        if (r->current_gap < 0) r->current_gap = 1e-6;
        if (r->current_gap <= 0.2 * r->last_gap) {
            r->sufficient++;
            r->restart_flag = 1;
        }
        if (r->current_gap <= 0.8 * r->last_gap && r->current_gap > 1.00 * r->save_gap) {
            r->necessary++;
            r->restart_flag = 2;
        }
        double fact = (r->current_gap / r->weighted_norm > 1e-1) ? 0.5 : 0.2;
        if ((double)r->inner >= fact * (double)iter) {
            r->long_++;
            r->restart_flag = 3;
        }
        if (r->best_gap > r->current_gap) {
            r->best_gap = r->current_gap;
            r->best_sigma = sigma;
        }
        r->save_gap = r->current_gap;
    }
}

/* do_restart! (src/algorithm.jl:522-555) */
static void do_restart(restart_t *r, ws_t *ws) {
    if (r->restart_flag <= 0) return;
    int32_t n = ws->n, m = ws->m;
    memcpy(ws->x, ws->x_bar, sizeof(double) * (size_t)n);
    memcpy(ws->w, ws->w_bar, sizeof(double) * (size_t)n);
    memcpy(ws->last_x, ws->x_bar, sizeof(double) * (size_t)n);
    memcpy(ws->last_w, ws->w_bar, sizeof(double) * (size_t)n);
    if (m > 0) {
        memcpy(ws->y, ws->y_bar, sizeof(double) * (size_t)m);
        memcpy(ws->last_y, ws->y_bar, sizeof(double) * (size_t)m);
        csc_mul(ws->AT, ws->y_bar, ws->ATy_bar);
        memcpy(ws->last_ATy, ws->ATy_bar, sizeof(double) * (size_t)n);
        memcpy(ws->ATy, ws->ATy_bar, sizeof(double) * (size_t)n);
    }
    if (ws->noC) {
        memcpy(ws->Qw, ws->Qw_bar, sizeof(double) * (size_t)n);
        memcpy(ws->last_Qw, ws->Qw_bar, sizeof(double) * (size_t)n);
    }
    r->last_gap = r->current_gap;
    r->save_gap = INFINITY;
    r->times++;
    r->inner = 0;
}

/* print_step (src/algorithm.jl:1173-1175): max(10^floor(log10 k)/10, 10) */
static int64_t print_step(int64_t k) {
    if (k <= 0) return 10;
    int64_t p = 1;
    while (p * 10 <= k) p *= 10;
    int64_t s = p / 10;
    return s > 10 ? s : 10;
}

static int should_print(int64_t iter, const oracle_params *p, double t0) {
    int64_t step = p->print_frequency == -1 ? print_step(iter) : p->print_frequency;
    return (iter % step == 0) || iter == p->max_iter || (now_s() - t0 > p->time_limit);
}

/* ------------------------------------------------------------------ */
/* solve (src/algorithm.jl:1904-2039)                                   */
/* ------------------------------------------------------------------ */

int oracle_solve(int32_t n, int32_t m, const oracle_qop *qin, int64_t a_nnz,
                 const int64_t *a_colptr, const int32_t *a_rowidx,
                 const double *a_val, const double *c_in, const double *AL_in,
                 const double *AU_in, const double *l_in, const double *u_in,
                 double obj_constant, double lam, const oracle_params *p,
                 const double *x0, const double *y0, double *x_out,
                 double *y_out, double *z_out, double *w_out, double *hist,
                 int64_t hist_cap, oracle_result *out) {
    (void)a_nnz;
    memset(out, 0, sizeof(*out));
    const double setup_start = now_s();
    const int64_t pmax = p->max_power_iter > 0 ? p->max_power_iter : 5000;
    qop_t Q;
    if (qop_build(n, qin, &Q) != 0) return -1;
    csc_t A = csc_copy(m, n, a_colptr, a_rowidx, a_val);
    double *c = dalloc(n), *l = dalloc(n), *u = dalloc(n), *AL = dalloc(m), *AU = dalloc(m);
    memcpy(c, c_in, sizeof(double) * (size_t)n);
    memcpy(l, l_in, sizeof(double) * (size_t)n);
    memcpy(u, u_in, sizeof(double) * (size_t)n);
    if (m > 0) {
        memcpy(AL, AL_in, sizeof(double) * (size_t)m);
        memcpy(AU, AU_in, sizeof(double) * (size_t)m);
    }
    scaling_t sc;
    scaling(n, m, &Q, &A, c, AL, AU, l, u, p, &sc);
    csc_t AT = csc_transpose(&A);

    ws_t ws;
    ws_alloc(&ws, n, m);
    ws.c = c;
    ws.l = l;
    ws.u = u;
    ws.AL = AL;
    ws.AU = AU;
    ws.A = &A;
    ws.AT = &AT;
    ws.Q = &Q;
    /* check_Q_diagonal after scaling (src/utils.jl:827-855) */
    ws.diag_Q = dalloc(n);
    if (Q.kind == ORACLE_Q_SPARSE) {
        int isdiag = 1;
        for (int32_t j = 0; j < n; ++j) {
            double colsum = 0.0, dj = 0.0;
            for (int64_t k = Q.Q.cp[j]; k < Q.Q.cp[j + 1]; ++k) {
                colsum += fabs(Q.Q.v[k]);
                if (Q.Q.ri[k] == j) dj = Q.Q.v[k];
            }
            ws.diag_Q[j] = dj;
            double ad = fabs(dj);
            /* isapprox(x, y) with rtol = sqrt(eps(Float64)) */
            int ok = colsum == ad ||
                     (isfinite(colsum) && isfinite(ad) &&
                      fabs(colsum - ad) <= 1.4901161193847656e-8 * fmax_j(colsum, ad));
            if (!ok) isdiag = 0;
        }
        ws.Q_is_diag = isdiag;
    } else {
        ws.Q_is_diag = 0;
    }
    /* sigma_0 (src/algorithm.jl:839-851) */
    if (p->sigma == -1) {
        if (sc.norm_c > 1e-16 && sc.norm_b > 1e-16 && sc.norm_b < 1e16 && sc.norm_c < 1e16)
            ws.sigma = sc.norm_b / sc.norm_c;
        else
            ws.sigma = 1.0;
    } else {
        ws.sigma = p->sigma;
    }
    out->sigma0 = ws.sigma;
    for (int32_t i = 0; i < n; ++i) ws.lambda[i] = lam;

    out->setup_time = now_s() - setup_start;
    double t_start = now_s();
    /* estimate_eigenvalues (src/algorithm.jl:1407-1440) */
    ws.lambda_max_A = m > 0 ? power_A(&A, &AT, pmax, 1e-4, &out->power_iter_A) * p->eig_factor : 0.0;
    out->power_time_A = now_s() - t_start;
    if (Q.kind != ORACLE_Q_SPARSE) {
        ws.lambda_max_Q = power_Q(&Q, n, pmax, 1e-4, &out->power_iter_Q) * p->eig_factor;
    } else if (Q.Q.nnz > 0) {
        if (!ws.Q_is_diag) {
            ws.lambda_max_Q = power_Q(&Q, n, pmax, 1e-4, &out->power_iter_Q) * p->eig_factor;
        } else {
            double mx = -INFINITY;
            for (int64_t k = 0; k < Q.Q.nnz; ++k) mx = fmax_j(mx, Q.Q.v[k]);
            ws.lambda_max_Q = mx;
        }
    } else {
        ws.lambda_max_Q = 0.0;
    }
    out->power_time = now_s() - t_start;
    out->lambda_max_A = ws.lambda_max_A;
    out->lambda_max_Q = ws.lambda_max_Q;

    /* process_initial_points! (src/algorithm.jl:1783-1836) */
    if (x0) {
        for (int32_t i = 0; i < n; ++i) {
            double v = x0[i] * sc.col_norm[i] / sc.b_scale;
            ws.x[i] = ws.x_bar[i] = ws.last_x[i] = v;
            ws.w[i] = ws.w_bar[i] = ws.last_w[i] = v;
        }
    }
    if (y0) {
        for (int32_t i = 0; i < m; ++i) {
            double v = y0[i];
            v = v * sc.row_norm[i] / sc.c_scale;
            ws.y[i] = ws.y_bar[i] = ws.last_y[i] = v;
        }
        if (m > 0) {
            csc_mul(&AT, ws.y_bar, ws.ATy_bar);
            memcpy(ws.ATy, ws.ATy_bar, sizeof(double) * (size_t)n);
            memcpy(ws.last_ATy, ws.ATy_bar, sizeof(double) * (size_t)n);
        }
        qmap(&Q, ws.x_bar, ws.Qx);
        for (int32_t i = 0; i < n; ++i) {
            double tmp = -ws.Qx[i] + ws.ATy_bar[i] - ws.c[i];
            double z_raw = ws.x_bar[i] + ws.sigma * tmp;
            ws.z_bar[i] = (ws.x_bar[i] - z_raw) / ws.sigma;
        }
        if (m > 0) {
            csc_mul(&A, ws.x_bar, ws.Ax);
            double f1 = ws.lambda_max_A * ws.sigma;
            for (int32_t i = 0; i < m; ++i)
                ws.s[i] = fmin_j(fmax_j(ws.Ax[i] - f1 * ws.y[i], ws.AL[i]), ws.AU[i]);
        }
    }

    restart_t r;
    memset(&r, 0, sizeof(r));
    r.first_restart = 1;
    r.save_gap = r.current_gap = r.last_gap = r.best_gap = INFINITY;
    r.best_sigma = 1.0;
    r.weighted_norm = INFINITY;
    r.best_sigma = ws.sigma;
    if (ws.Q_is_diag) update_Q_factors(&ws);

    /* noC uses the unscaled bounds (src/algorithm.jl:1969-1974) */
    int64_t n_free = 0;
    for (int32_t i = 0; i < n; ++i)
        if (l_in[i] == -INFINITY && u_in[i] == INFINITY) n_free++;
    ws.noC = (double)n_free > 0.8 * (double)n;
    out->noC = ws.noC;
    out->Q_is_diag = ws.Q_is_diag;

    resid_t res;
    memset(&res, 0, sizeof(res));
    int lasso = p->problem_type == 1;
    int64_t iter_4 = 0, iter_6 = 0;
    int first_4 = 1, first_6 = 1;
    int64_t nh = 0;
    int status = -1;
    int64_t iter;
    ws.to_check = 1;
    for (iter = 0; iter <= p->max_iter; ++iter) {
        int print_yes = should_print(iter, p, t_start);
        if (iter % p->check_iter == 0 || print_yes) {
            compute_residuals(&ws, &sc, &res, obj_constant, lasso, iter);
        }
        int st;
        /* check_break (src/algorithm.jl:558-576) */
        if (res.kkt < p->stoptol)
            st = ORACLE_OPTIMAL;
        else if (iter == p->max_iter)
            st = ORACLE_MAX_ITER;
        else if (now_s() - t_start > p->time_limit)
            st = ORACLE_TIME_LIMIT;
        else
            st = -1;
        check_restart(&r, iter, p->check_iter, ws.sigma);
        update_sigma(p, &r, &ws, &res);
        do_restart(&r, &ws);
        if ((iter % p->check_iter == 0 || print_yes) && hist && nh < hist_cap) {
            double *h = hist + nh * ORACLE_HIST_COLS;
            h[0] = (double)iter;
            h[1] = res.err_Rp;
            h[2] = res.err_Rd;
            h[3] = res.rel_gap;
            h[4] = res.primal_obj;
            h[5] = res.dual_obj;
            h[6] = ws.sigma;
            h[7] = (double)r.restart_flag;
            nh++;
        }
        if (res.kkt < 1e-4 && first_4) {
            iter_4 = iter;
            first_4 = 0;
        }
        if (res.kkt < 1e-6 && first_6) {
            iter_6 = iter;
            first_6 = 0;
        }
        if (st != -1) {
            status = st;
            break;
        }
        int64_t next = iter + 1;
        int64_t ps = p->print_frequency == -1 ? print_step(next) : p->print_frequency;
        ws.to_check = (next % p->check_iter == 0) || (r.restart_flag > 0) || (next % ps == 0);
        /* perform_iteration_step! (src/algorithm.jl:1881-1900) */
        main_update(&ws, &r);
        if (r.restart_flag > 0) r.last_gap = compute_M_norm(&ws, NULL);
        if ((iter + 1) % p->check_iter == 0) r.current_gap = compute_M_norm(&ws, NULL);
        r.inner++;
    }
    if (iter > p->max_iter) iter = p->max_iter;
    out->status = status;
    out->iter = iter;
    out->iter_4 = iter_4 == 0 ? iter : iter_4;
    out->iter_6 = iter_6 == 0 ? iter : iter_6;
    out->time = now_s() - t_start;
    out->primal_obj = res.primal_obj;
    out->dual_obj = res.dual_obj;
    out->residuals = res.kkt;
    out->gap = res.rel_gap;
    out->err_Rp = res.err_Rp;
    out->err_Rd = res.err_Rd;
    out->sigma = ws.sigma;
    out->restarts = (int32_t)r.times;
    out->n_history = nh;
    /* collect_results! (src/algorithm.jl:611-643) */
    for (int32_t i = 0; i < n; ++i) {
        if (x_out) x_out[i] = sc.b_scale * (ws.x_bar[i] / sc.col_norm[i]);
        if (w_out) w_out[i] = sc.b_scale * (ws.w_bar[i] / sc.col_norm[i]);
        if (z_out) z_out[i] = sc.c_scale * (ws.z_bar[i] * sc.col_norm[i]);
    }
    for (int32_t i = 0; i < m; ++i)
        if (y_out) y_out[i] = sc.c_scale * (ws.y_bar[i] / sc.row_norm[i]);

    ws_free(&ws);
    csc_free(&A);
    csc_free(&AT);
    qop_free(&Q);
    free(c);
    free(l);
    free(u);
    free(AL);
    free(AU);
    free(sc.row_norm);
    free(sc.col_norm);
    return 0;
}

/* ------------------------------------------------------------------ */
/* scaling! and compute_M_norm! on their own, for the pinning tests     */
/* ------------------------------------------------------------------ */

int oracle_scale(int32_t n, int32_t m, const oracle_qop *qin, int64_t a_nnz,
                 const int64_t *a_colptr, const int32_t *a_rowidx,
                 const double *a_val, const double *c_in, const double *AL_in,
                 const double *AU_in, const double *l_in, const double *u_in,
                 const oracle_params *p, double *Q_out, double *A_out,
                 double *c_out, double *AL_out, double *AU_out, double *l_out,
                 double *u_out, double *row_norm, double *col_norm,
                 double *norms) {
    (void)a_nnz;
    qop_t Q;
    if (qop_build(n, qin, &Q) != 0) return -1;
    csc_t A = csc_copy(m, n, a_colptr, a_rowidx, a_val);
    memcpy(c_out, c_in, sizeof(double) * (size_t)n);
    memcpy(l_out, l_in, sizeof(double) * (size_t)n);
    memcpy(u_out, u_in, sizeof(double) * (size_t)n);
    if (m > 0) {
        memcpy(AL_out, AL_in, sizeof(double) * (size_t)m);
        memcpy(AU_out, AU_in, sizeof(double) * (size_t)m);
    }
    scaling_t sc;
    scaling(n, m, &Q, &A, c_out, AL_out, AU_out, l_out, u_out, p, &sc);
    if (Q_out && Q.kind == ORACLE_Q_SPARSE) {
        memset(Q_out, 0, sizeof(double) * (size_t)n * (size_t)n);
        for (int32_t j = 0; j < n; ++j)
            for (int64_t k = Q.Q.cp[j]; k < Q.Q.cp[j + 1]; ++k)
                Q_out[Q.Q.ri[k] + (int64_t)j * n] = Q.Q.v[k];
    }
    if (A_out) {
        memset(A_out, 0, sizeof(double) * (size_t)m * (size_t)n);
        for (int32_t j = 0; j < n; ++j)
            for (int64_t k = A.cp[j]; k < A.cp[j + 1]; ++k) A_out[A.ri[k] + (int64_t)j * m] = A.v[k];
    }
    memcpy(row_norm, sc.row_norm, sizeof(double) * (size_t)m);
    memcpy(col_norm, sc.col_norm, sizeof(double) * (size_t)n);
    norms[0] = sc.norm_b;
    norms[1] = sc.norm_c;
    norms[2] = sc.norm_b_org;
    norms[3] = sc.norm_c_org;
    norms[4] = sc.b_scale;
    norms[5] = sc.c_scale;
    csc_free(&A);
    qop_free(&Q);
    free(sc.row_norm);
    free(sc.col_norm);
    return 0;
}

double oracle_mnorm(int32_t n, int32_t m, const oracle_qop *qin, int64_t a_nnz,
                    const int64_t *a_colptr, const int32_t *a_rowidx,
                    const double *a_val, const double *dx, const double *dy,
                    const double *dw, double sigma, double lambda_max_A,
                    double lambda_max_Q, int Q_is_diag, double *terms) {
    (void)a_nnz;
    qop_t Q;
    if (qop_build(n, qin, &Q) != 0) return NAN;
    csc_t A = csc_copy(m, n, a_colptr, a_rowidx, a_val);
    csc_t AT = csc_transpose(&A);
    ws_t ws;
    ws_alloc(&ws, n, m);
    ws.A = &A;
    ws.AT = &AT;
    ws.Q = &Q;
    ws.sigma = sigma;
    ws.lambda_max_A = lambda_max_A;
    ws.lambda_max_Q = lambda_max_Q;
    ws.Q_is_diag = Q_is_diag;
    ws.diag_Q = dalloc(n);
    if (Q_is_diag && Q.kind == ORACLE_Q_SPARSE)
        for (int32_t j = 0; j < n; ++j)
            for (int64_t k = Q.Q.cp[j]; k < Q.Q.cp[j + 1]; ++k)
                if (Q.Q.ri[k] == j) ws.diag_Q[j] = Q.Q.v[k];
    if (Q_is_diag) update_Q_factors(&ws);
    memcpy(ws.dx, dx, sizeof(double) * (size_t)n);
    memcpy(ws.dw, dw, sizeof(double) * (size_t)n);
    if (m > 0) memcpy(ws.dy, dy, sizeof(double) * (size_t)m);
    const double r = compute_M_norm(&ws, terms);
    ws_free(&ws);
    csc_free(&A);
    csc_free(&AT);
    qop_free(&Q);
    return r;
}

/* ------------------------------------------------------------------ */
/* one half-step on a given state (per-kernel check)                   */
/* ------------------------------------------------------------------ */

int oracle_halfstep(int kind, int32_t n, int32_t m, const oracle_qop *qin,
                    int64_t a_nnz, const int64_t *a_colptr,
                    const int32_t *a_rowidx, const double *a_val,
                    const double *c, const double *AL, const double *AU,
                    const double *l, const double *u, oracle_state *st) {
    (void)a_nnz;
    qop_t q;
    if (qop_build(n, qin, &q) != 0) return -1;
    csc_t A = csc_copy(m, n, a_colptr, a_rowidx, a_val);
    csc_t AT = csc_transpose(&A);
    ws_t ws;
    memset(&ws, 0, sizeof(ws));
    ws.n = n;
    ws.m = m;
    ws.w = st->w; ws.w_bar = st->w_bar; ws.dw = st->dw; ws.x = st->x;
    ws.x_hat = st->x_hat; ws.x_bar = st->x_bar; ws.dx = st->dx;
    ws.z_bar = st->z_bar; ws.ATy = st->ATy; ws.ATy_bar = st->ATy_bar;
    ws.ATdy = st->ATdy; ws.Qw = st->Qw; ws.Qw_bar = st->Qw_bar;
    ws.tempv = st->tempv; ws.last_x = st->last_x; ws.last_w = st->last_w;
    ws.last_ATy = st->last_ATy; ws.last_Qw = st->last_Qw;
    ws.fact1 = st->fact1; ws.fact2 = st->fact2; ws.fact = st->fact;
    ws.y = st->y; ws.y_bar = st->y_bar; ws.dy = st->dy; ws.s = st->s;
    ws.last_y = st->last_y;
    /* the step functions' own temporaries */
    ws.Ax = dalloc(m);
    ws.Rp = dalloc(m);
    ws.Rd = dalloc(n);
    ws.Qx = dalloc(n);
    ws.dQw = dalloc(n);
    ws.QATdy = dalloc(n);
    ws.fact_M = dalloc(n);
    ws.lambda = dalloc(n);
    for (int32_t i = 0; i < n; ++i) ws.lambda[i] = st->lambda;
    ws.c = c; ws.l = l; ws.u = u; ws.AL = AL; ws.AU = AU;
    ws.A = &A;
    ws.AT = &AT;
    ws.Q = &q;
    ws.sigma = st->sigma;
    ws.lambda_max_A = st->lambda_max_A;
    ws.lambda_max_Q = st->lambda_max_Q;
    ws.Q_is_diag = st->Q_is_diag;
    ws.to_check = st->full;
    int rc = 0;
    switch (kind) {
        case 1: step_zxw1(&ws, st->h1, st->h2); break;
        case 2: step_y(&ws, st->h1, st->h2); break;
        case 3: step_w2(&ws, st->h1, st->h2); break;
        case 4: step_noc_zxw(&ws, st->h1, st->h2); break;
        case 5: step_y_from(&ws, ws.tempv, st->h1, st->h2, 1); break;
        case 6: step_lasso(&ws, st->h1, st->h2); break;
        case 7: step_lp_zx(&ws, st->h1, st->h2); break;
        case 8: step_y_from(&ws, ws.x_hat, st->h1, st->h2, 0); break;
        default: rc = -2;
    }
    free(ws.Ax); free(ws.Rp); free(ws.Rd); free(ws.Qx); free(ws.dQw);
    free(ws.QATdy); free(ws.fact_M); free(ws.lambda);
    csc_free(&A);
    csc_free(&AT);
    qop_free(&q);
    return rc;
}
