/*
 * hprqp_oracle.h -- CPU FP64 restatement of the HPR-QP reference CPU path
 * (use_gpu=false).  TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this; the product library never does.
 *
 * Sparse matrices are CSC (column pointers, row indices, values), 0-based,
 * which is the layout of Julia's SparseMatrixCSC that the reference CPU path
 * multiplies with mul! (a column scatter; for each output entry it sums in
 * increasing index order, exactly like a sequential CSR row gather).
 */
#ifndef HPRQP_ORACLE_H
#define HPRQP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_Q_SPARSE = 0, ORACLE_Q_LASSO = 1, ORACLE_Q_QAP = 2, ORACLE_Q_LOWRANK = 3 };
enum { ORACLE_OPTIMAL = 0, ORACLE_MAX_ITER = 1, ORACLE_TIME_LIMIT = 2 };

typedef struct {
    int32_t kind;
    /* ORACLE_Q_SPARSE: n x n CSC */
    int64_t nnz;
    const int64_t *colptr;
    const int32_t *rowidx;
    const double *val;
    /* ORACLE_Q_LASSO: data matrix, rows x n, CSC */
    int32_t rows;
    int64_t l_nnz;
    const int64_t *l_colptr;
    const int32_t *l_rowidx;
    const double *l_val;
    /* ORACLE_Q_QAP: dense nq x nq column-major, Q(X) = 2(A X B - S X - X T) */
    int32_t nq;
    const double *qa, *qb, *qs, *qt;
    /* ORACLE_Q_LOWRANK: Q = diag(d) + U U^T, U n x k column-major (the
     * reference operator guide's DiagonalPlusLowRank custom operator) */
    int32_t lr_k;
    const double *lr_U, *lr_d;
} oracle_qop;

typedef struct {
    double stoptol;
    double sigma;          /* -1 = automatic */
    int64_t max_iter;
    int32_t sigma_fixed;
    double time_limit;
    double eig_factor;
    int64_t check_iter;
    int64_t print_frequency; /* -1 = adaptive */
    int32_t use_Ruiz_scaling;
    int32_t use_Pock_Chambolle_scaling;
    int32_t problem_type;    /* 0 QP, 1 LASSO, 2 QAP */
    int32_t pad;
    int64_t max_power_iter;  /* 0 = the reference's 5000; >0 bounds a timing sample */
} oracle_params;

typedef struct {
    int32_t status;
    int32_t Q_is_diag;
    int32_t noC;
    int32_t restarts;
    int64_t iter;
    int64_t iter_4;
    int64_t iter_6;
    double time;
    double power_time;
    double primal_obj;
    double dual_obj;
    double residuals;
    double gap;
    double err_Rp;
    double err_Rd;
    double lambda_max_A;
    double lambda_max_Q;
    double sigma;
    double sigma0;
    int64_t n_history;
    int64_t power_iter_A;
    int64_t power_iter_Q;
    double setup_time;
    double power_time_A;
} oracle_result;

/* History row written at every residual evaluation:
 * iter, err_Rp, err_Rd, rel_gap, primal_obj, dual_obj, sigma, restart_flag */
#define ORACLE_HIST_COLS 8

/*
 * Solve min 1/2<x,Qx> + <c,x> (+ lambda||x||_1 for LASSO)
 *   s.t. AL <= A x <= AU, l <= x <= u.
 * A is m x n CSC.  x0/y0 are optional warm starts in the original space.
 * x_out,y_out,z_out,w_out (may be NULL) receive unscaled results.
 * hist (may be NULL) receives up to hist_cap rows of ORACLE_HIST_COLS doubles.
 */
int oracle_solve(int32_t n, int32_t m, const oracle_qop *q,
                 int64_t a_nnz, const int64_t *a_colptr, const int32_t *a_rowidx,
                 const double *a_val, const double *c, const double *AL,
                 const double *AU, const double *l, const double *u,
                 double obj_constant, double lambda, const oracle_params *p,
                 const double *x0, const double *y0, double *x_out,
                 double *y_out, double *z_out, double *w_out, double *hist,
                 int64_t hist_cap, oracle_result *res);

/* y = M x for an nrow x ncol CSC matrix (the reference's CPU mul!). */
void oracle_csc_mul(int32_t nrow, int32_t ncol, const int64_t *colptr,
                    const int32_t *rowidx, const double *val, const double *x,
                    double *y);

/* Qx = Q(x) for any oracle_qop (the reference's CPU Qmap!). */
int oracle_qmap(int32_t n, const oracle_qop *q, const double *x, double *Qx);

/* Largest eigenvalue estimate by the reference's power method on A A^T
 * (which=0) or Q (which=1); start vector from oracle_randn(seed). */
double oracle_power_iteration(int32_t n, int32_t m, const oracle_qop *q,
                              int64_t a_nnz, const int64_t *a_colptr,
                              const int32_t *a_rowidx, const double *a_val,
                              int which, int64_t max_it, double tol,
                              int64_t *iters_out);

/* Deterministic N(0,1) stream used for the power-method start vector. */
void oracle_randn(uint64_t seed, int64_t count, double *out);

/* Golden-section searches (host sigma update). */
double oracle_golden(double a, double b, double c, double d, double lo,
                     double hi, double tol, int maxiter);
double oracle_golden_Q_diag(double a, double b, int32_t n, const double *Q,
                            const double *c, const double *d, double lo,
                            double hi, double tol, int maxiter);

/* scaling! on its own (src/utils.jl:350-804): the scaled Q (sparse kind
 * only) and A as dense column-major n x n and m x n arrays, the scaled c,
 * AL, AU, l, u, the cumulative row_norm (m) and col_norm (n), and norms[6] =
 * norm_b, norm_c (scaled data), norm_b_org, norm_c_org, b_scale, c_scale. */
int oracle_scale(int32_t n, int32_t m, const oracle_qop *q, int64_t a_nnz,
                 const int64_t *a_colptr, const int32_t *a_rowidx,
                 const double *a_val, const double *c, const double *AL,
                 const double *AU, const double *l, const double *u,
                 const oracle_params *p, double *Q_out, double *A_out,
                 double *c_out, double *AL_out, double *AU_out, double *l_out,
                 double *u_out, double *row_norm, double *col_norm,
                 double *norms);

/* compute_M_norm! on its own (src/algorithm.jl:689-739) for given dx (n),
 * dy (m), dw (n), sigma, lambda_max_A, lambda_max_Q and the diagonal-Q flag
 * (the factor fact_M = sigma^2 / (1 + sigma Q_ii) is formed as the solver
 * forms it).  Returns the M-norm; terms (may be NULL) receives the
 * unclamped M_1, M_2, M_3. */
double oracle_mnorm(int32_t n, int32_t m, const oracle_qop *q, int64_t a_nnz,
                    const int64_t *a_colptr, const int32_t *a_rowidx,
                    const double *a_val, const double *dx, const double *dy,
                    const double *dw, double sigma, double lambda_max_A,
                    double lambda_max_Q, int Q_is_diag, double *terms);

/* One half-step of the iteration on a given state (test infrastructure:
 * the check of each GPU half-step kernel on its own).  kind:
 *   1 z/x/w1 (update_zxw1_cpu!, src/kernels.jl:2301-2397)
 *   2 tempv and y (update_y_cpu!, :2682-2743)
 *   3 w2 (update_w2_cpu!, :2747-2832)
 *   4 noC z/x/w with the Halpern Qw update and tempv (:2117-2238)
 *   5 the noC y step on tempv (:2242-2298)
 *   6 LASSO z/x/w (update_zxw_LASSO_cpu!, :2400-2496)
 *   7 LP z/x (unified_update_zx_cpu!, :2504-2578)
 *   8 the LP y step on x_hat (:2582-2658; the CPU loop's partial write set,
 *     without y_bar)
 * The data are taken as given (no scaling); full selects the full write
 * set.  Every vector must be allocated with its length (n or m). */
typedef struct {
    double *w, *w_bar, *dw, *x, *x_hat, *x_bar, *dx, *z_bar, *ATy, *ATy_bar,
        *ATdy, *Qw, *Qw_bar, *tempv, *last_x, *last_w, *last_ATy, *last_Qw,
        *fact1, *fact2, *fact;        /* n */
    double *y, *y_bar, *dy, *s, *last_y; /* m */
    double sigma, lambda_max_A, lambda_max_Q, lambda, h1, h2;
    int32_t Q_is_diag, full;
} oracle_state;
int oracle_halfstep(int kind, int32_t n, int32_t m, const oracle_qop *q,
                    int64_t a_nnz, const int64_t *a_colptr,
                    const int32_t *a_rowidx, const double *a_val,
                    const double *c, const double *AL, const double *AU,
                    const double *l, const double *u, oracle_state *st);

#ifdef __cplusplus
}
#endif
#endif
