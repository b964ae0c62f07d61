/*
 * hprqp.h -- C ABI of libhprqp, the MI355X (gfx950) implementation of the
 * HPR-QP iteration (dual Halpern Peaceman-Rachford for convex composite QP):
 *
 *     min 1/2 <x,Qx> + <c,x> + phi(x)   s.t.  AL <= A x <= AU,  l <= x <= u
 *
 * phi is the box indicator, or lambda*||x||_1 for LASSO.
 *
 * Every entry point returns 0 on success or a negative HPRQP_E* code; the
 * message of the last failure on a context is hprqp_last_error(ctx).  No C++
 * exception crosses this boundary.  A context owns one HIP stream; work on a
 * context is asynchronous on that stream unless the entry point returns host
 * data.  A context is used by one host thread at a time.
 *
 * Sparse matrices are CSR with int32 indices (0- or 1-based on input, the
 * latter being Julia's SparseMatrixCSC transposed/CuSparseMatrixCSR 'O'
 * layout).  Host arrays are copied; nothing is retained after a call.
 *
 * Reference interfaces replaced (HPRQP.jl, Julia):
 *   optimize / solve ................. src/algorithm.jl:1511-1554, 1904-2039
 *   Qmap! plugin interface ........... src/Q_operators/Q_operator_interface.jl:5-92
 *   CSR/LASSO/QAP operators .......... src/Q_operators/{sparse_matrix,LASSO,QAP}_operator.jl
 *   A*x, A'*y (cuSPARSE) ............. src/kernels.jl:33-99, src/unified_operations.jl:117-143
 *   fused half-step kernels .......... src/kernels.jl:261-1426
 *   residual/KKT reductions .......... src/algorithm.jl:37-133, src/kernels.jl:114-194
 *   M-norm, sigma, restart ........... src/algorithm.jl:323-555, 689-739
 *   power method ..................... src/utils.jl:1335-1461
 *   scaling .......................... src/utils.jl:350-804
 */
#ifndef HPRQP_H
#define HPRQP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPRQP_ABI_VERSION 6  /* 2: hprqp_params gained the auto-save fields;
                                3: hprqp_batch_create takes the total nnz of A;
                                4: hprqp_batch_result gained interval_time;
                                5: hprqp_params.reserved0 became
                                   qap_recompute_Qw;
                                6: hprqp_params gained qap_family and
                                   hprqp_result.lanes_pad became qap_family */

/* ---- error codes ---------------------------------------------------- */
#define HPRQP_OK 0
#define HPRQP_E_ARG (-1)     /* invalid argument (dimension, NULL, range) */
#define HPRQP_E_HIP (-2)     /* HIP runtime error */
#define HPRQP_E_ALLOC (-3)   /* device or host allocation failed */
#define HPRQP_E_STATE (-4)   /* object used in the wrong state */
#define HPRQP_E_UNSUP (-5)   /* unsupported combination */
#define HPRQP_E_DATA (-6)    /* invalid problem data (NaN/Inf, l > u, ...) */

/* ---- status (HPRQP_results.status) ---------------------------------- */
#define HPRQP_STATUS_OPTIMAL 0
#define HPRQP_STATUS_MAX_ITER 1
#define HPRQP_STATUS_TIME_LIMIT 2

/* ---- problem types (HPRQP_parameters.problem_type) ------------------ */
#define HPRQP_PROBLEM_QP 0
#define HPRQP_PROBLEM_LASSO 1
#define HPRQP_PROBLEM_QAP 2

/* ---- precision modes ------------------------------------------------- */
#define HPRQP_PREC_FP64 0   /* everything FP64 (plain form) */
#define HPRQP_PREC_MIXED 1  /* FP32 matrix values and problem data (the bulk
                               byte streams); FP64 iterates, anchors,
                               accumulation and KKT */
#define HPRQP_PREC_FP32 2   /* FP32 matrices, iterates and accumulation; FP64
                               anchors and KKT; FP64 polish when FP32 stalls.
                               On the standard path with a non-diagonal sparse
                               Q the bulk iterates are FP32 offsets from the
                               FP64 Halpern anchors, and the anchors' products
                               are FP64, formed once per restart: that form
                               tracks the FP64 iteration to the tolerance
                               and switches to FP64 only on a long stall.
                               LASSO runs the HPRQP_PREC_MIXED form: its
                               sigma climbs to 1e4-1e5, where FP32 iterates
                               drift off the solution */

typedef struct hprqp_ctx hprqp_ctx;
typedef struct hprqp_buf hprqp_buf;
typedef struct hprqp_csr hprqp_csr;
typedef struct hprqp_qop hprqp_qop;
typedef struct hprqp_model hprqp_model;

/* State handed to the auto-save callback at print steps and at termination
 * (save_state_to_hdf5!, src/algorithm.jl:183-301): the current iterate and
 * the best one so far -- lowest max(err_Rp, err_Rd, rel_gap) over the
 * residual checks (src/algorithm.jl:97-112) -- in the original space with
 * the reference's save formulas (x, w: b_scale v / col_norm; y: c_scale v /
 * row_norm; z: c_scale v / col_norm).  The arrays are host memory valid only
 * during the call; y and best_y are empty when m = 0. */
typedef struct hprqp_save_state {
    int32_t n, m;
    int64_t iteration;
    double time_elapsed, sigma, err_Rp, err_Rd, primal_obj, dual_obj, rel_gap, kkt;
    const double *x, *y, *z, *w;
    int64_t best_iteration;
    double best_sigma, best_err_Rp, best_err_Rd, best_primal_obj, best_dual_obj, best_rel_gap;
    const double *best_x, *best_y, *best_z, *best_w;
} hprqp_save_state;
/* returns 0 on success; a non-zero return is reported (when verbose) and the
 * solve continues, as the reference catches a failed save
 * (src/algorithm.jl:2005-2012) */
typedef int (*hprqp_save_fn)(const hprqp_save_state *state, void *user);

/* Solver parameters.  Field meanings and defaults follow HPRQP_parameters
 * (src/structs.jl:195-275); hprqp_params_default fills the defaults. */
typedef struct {
    double stoptol;             /* 1e-6 */
    double sigma;               /* -1 = automatic */
    int64_t max_iter;           /* 2^31-1 */
    int32_t sigma_fixed;        /* 0 */
    int32_t problem_type;       /* HPRQP_PROBLEM_QP */
    double time_limit;          /* 3600 s */
    double eig_factor;          /* 1.05 */
    int64_t check_iter;         /* 100 */
    int64_t print_frequency;    /* -1 = adaptive cadence */
    int32_t use_Ruiz_scaling;   /* 1 */
    int32_t use_Pock_Chambolle_scaling; /* 1 */
    int32_t use_bc_scaling;     /* 0 (disabled in the reference as well) */
    int32_t precision;          /* HPRQP_PREC_FP64 */
    int32_t verbose;            /* 0 */
    int32_t spmv_lanes;         /* 0 = autotune; 1..64 (power of two) = lanes
                                   per row; 256 = one block per row;
                                   -64/-128/-256 = CSR-stream with that many
                                   rows per block; 1048576 = column-chunked
                                   (x staged in LDS 4,096 columns at a time);
                                   -1032/-1064/-1128/-1256 = value-table form;
                                   -2000 = the diagonal form for Q, which must
                                   have one, and for A and A^T where they have
                                   one (nonzeros on at most 8 diagonals j =
                                   floor(i ncol / nrow) + offset, each kept as
                                   an array by row; the others keep their
                                   default shapes).  The autotuner also times
                                   the diagonal form for such matrices */
    int32_t time_kernels;       /* 1 = time the dominant fused kernel with
                                   HIP events during the solve (result.kernel_ms);
                                   one launch in eight carries the events */
    int32_t polish_stall_checks; /* reduced precision: residual checks without
                                    the KKT residual halving before the FP64
                                    polish takes over, once the residual has
                                    been below 1e-3 (0 = automatic: 100 for
                                    HPRQP_PREC_MIXED, 40 for HPRQP_PREC_FP32,
                                    400 for its offset form) */
    double polish_switch;        /* reduced precision: the FP64 polish also
                                    takes over once the KKT residual drops
                                    below polish_switch * stoptol (default 5;
                                    0 = only on a stall; not used by the FP32
                                    offset form) */
    int32_t auto_save;           /* 0; 1 = track the best iterate and hand it
                                    with the current one to save_cb at print
                                    steps and at termination (auto_save) */
    int32_t qap_recompute_Qw;    /* 0: for the QAP operator with the relaxation's
                                    doubly stochastic constraints, one operator
                                    apply per iteration -- Q w and Q ATy are
                                    carried by their linear recurrences and
                                    Q A^T y_bar is formed from y_bar in O(nq^2);
                                    1: Q w is recomputed every iteration, as
                                    the reference does (two applies) */
    int32_t qap_family;          /* QAP operator, n >= 256: kernel family of
                                    its two GEMM passes.  0 = autotune (each
                                    family that can run is timed at the
                                    operator's first solve in each precision
                                    and the fastest kept); 2 = blocked 32 x 32
                                    tiles from global memory; 3 = LDS-staged
                                    64 x 64 blocks (diagonal S and T only);
                                    4 = LDS-staged 64 x 128 blocks (split-BF16
                                    passes, diagonal S and T, a grid of at
                                    least 512 blocks).  A family that cannot
                                    run on the operator falls back to the size
                                    rule (n < 256: one-tile passes, family 1;
                                    blocked below n = 1024; LDS-staged from
                                    n = 1024 with diagonal S and T) */
    hprqp_save_fn save_cb;       /* NULL: tracked but not handed over */
    void *save_user;             /* passed to save_cb */
} hprqp_params;

/* Results (HPRQP_results, src/structs.jl:292-309) plus diagnostics. */
typedef struct {
    int32_t status;
    int32_t restarts;
    int64_t iter;
    int64_t iter_4;
    int64_t iter_6;
    double time;          /* seconds, the reference's results.time window */
    double time_4;
    double time_6;
    double power_time;
    double setup_time;    /* scaling + allocation, outside results.time */
    double primal_obj;
    double dual_obj;
    double residuals;     /* max(err_Rp, err_Rd, rel_gap) */
    double gap;
    double err_Rp;
    double err_Rd;
    double lambda_max_A;
    double lambda_max_Q;
    double sigma;
    double sigma0;
    int32_t Q_is_diag;
    int32_t noC;
    int64_t n_history;
    int64_t polish_iter;  /* FP64 polish iterations (mixed mode) */
    int64_t power_iter_A; /* power-method iterations for lambda_max(A A^T) */
    int64_t power_iter_Q; /* power-method iterations for lambda_max(Q) */
    double kernel_ms;     /* average ms of the first fused step kernel when
                             params.time_kernels = 1 (0 otherwise) */
    int64_t kernel_calls; /* launches averaged into kernel_ms (every eighth) */
    int32_t lanes_Q;      /* row-product shape used for Q, A, A^T (spmv_lanes */
    int32_t lanes_A;      /* codes: 1..64 lanes per row, 256 a block per row, */
    int32_t lanes_AT;     /* -16 .. -256 CSR-stream rows per block, 1048576
                             column-chunked: x staged in LDS 4,096 columns at a
                             time, for long rows over many columns, -1000 - rb
                             value-table form, -2000 diagonal form) */
    int32_t qap_family;   /* QAP operator: family of the iteration's passes
                             (params.qap_family codes, 1 = one-tile; the
                             split-BF16 one in the reduced precisions) */
} hprqp_result;

/* User Q operator plugin (AbstractQOperator, Q_operator_interface.jl:33-92).
 * apply must enqueue Qx = Q x on `stream` (a hipStream_t) for device
 * pointers x, Qx of length n (FP64). */
typedef struct {
    int (*apply)(void *user, const double *x_dev, double *Qx_dev, void *stream);
    int64_t (*problem_size)(void *user);
    int64_t (*temp_size)(void *user);
    const char *(*name)(void *user);
} hprqp_qop_vtable;

/* ---- context -------------------------------------------------------- */
int hprqp_ctx_create(int device, hprqp_ctx **out);
int hprqp_ctx_destroy(hprqp_ctx *ctx);
const char *hprqp_last_error(const hprqp_ctx *ctx);
int hprqp_ctx_sync(hprqp_ctx *ctx);
int hprqp_abi_version(void);
void hprqp_params_default(hprqp_params *p);

/* ---- device buffers (workspace vectors are FP64 device arrays) ------- */
int hprqp_buf_alloc(hprqp_ctx *ctx, size_t bytes, hprqp_buf **out);
int hprqp_buf_free(hprqp_ctx *ctx, hprqp_buf *buf);
int hprqp_buf_upload(hprqp_ctx *ctx, hprqp_buf *buf, const void *host, size_t bytes);
int hprqp_buf_download(hprqp_ctx *ctx, const hprqp_buf *buf, void *host, size_t bytes);
void *hprqp_buf_device_ptr(hprqp_buf *buf);
size_t hprqp_buf_size(const hprqp_buf *buf);

/* ---- CSR matrices (CuSparseMatrixCSR + the stored transpose) ---------- */
/* replaces CuSparseMatrixCSR(A) / CuSparseMatrixCSR(A') built in
 * transfer_to_gpu (src/algorithm.jl:1284-1322). */
int hprqp_csr_create(hprqp_ctx *ctx, int32_t nrow, int32_t ncol, int64_t nnz,
                     const int32_t *rowptr, const int32_t *colidx,
                     const double *vals, int index_base, int build_transpose,
                     hprqp_csr **out);
int hprqp_csr_destroy(hprqp_ctx *ctx, hprqp_csr *a);
/* y = A x (transpose=0) or A^T x via the stored transpose (transpose=1);
 * replaces Amap!/ATmap!/unified_mul! (src/kernels.jl:33-99). */
int hprqp_csr_spmv(hprqp_ctx *ctx, const hprqp_csr *a, int transpose,
                   const hprqp_buf *x, hprqp_buf *y);

/* ---- Q operators (Qmap! plugin interface) ----------------------------- */
/* sparse Q (n x n CSR), src/Q_operators/sparse_matrix_operator.jl */
int hprqp_qop_create_csr(hprqp_ctx *ctx, int32_t n, int64_t nnz,
                         const int32_t *rowptr, const int32_t *colidx,
                         const double *vals, int index_base, hprqp_qop **out);
/* Q = A^T A for a rows x n data matrix, src/Q_operators/LASSO_operator.jl */
int hprqp_qop_create_lasso(hprqp_ctx *ctx, int32_t rows, int32_t n, int64_t nnz,
                           const int32_t *rowptr, const int32_t *colidx,
                           const double *vals, int index_base, hprqp_qop **out);
/* Q(X) = 2(A X B - S X - X T), dense nq x nq column-major inputs,
 * src/Q_operators/QAP_operator.jl */
int hprqp_qop_create_qap(hprqp_ctx *ctx, int32_t nq, const double *A,
                         const double *B, const double *S, const double *T,
                         hprqp_qop **out);
/* Q = diag(d) + U U^T with U n x k column-major: the "diagonal plus low
 * rank" custom operator of the reference's operator guide
 * (docs/src/guide/q_operators_overview.md:121-139), here built in */
int hprqp_qop_create_lowrank(hprqp_ctx *ctx, int32_t n, int32_t k,
                             const double *U, const double *diag,
                             hprqp_qop **out);
/* user operator through a vtable of C function pointers */
int hprqp_qop_create_user(hprqp_ctx *ctx, const hprqp_qop_vtable *vt,
                          void *user, hprqp_qop **out);
int hprqp_qop_destroy(hprqp_ctx *ctx, hprqp_qop *q);
/* Qmap!(x, Qx, Q): FP64 in, FP64 out, on the context stream */
int hprqp_qop_apply(hprqp_ctx *ctx, hprqp_qop *q, const hprqp_buf *x,
                    hprqp_buf *Qx);
/* The hprqp_qop_apply product with no allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (HIP graph
 * capture, tight serving loops).  The operator's temporaries -- LASSO A x
 * (rows doubles), QAP A X and S X (2 nq^2 doubles), low rank U^T x (k
 * doubles) -- go to a static device array of 1,048,576 doubles (8 MiB) that
 * each GPU holds from the first context created on it until the process
 * ends; operators needing more are refused with HPRQP_E_ARG before any
 * launch.  Every call on one GPU shares that array (every temporary is fully
 * written before it is read, so nothing carries over between calls), so
 * calls that overlap on one GPU -- concurrent streams, contexts or threads --
 * race.  Enqueues only: no hipMalloc, hipFree, memset node or host
 * synchronisation. */
int hprqp_qop_apply_noalloc(hprqp_ctx *ctx, hprqp_qop *q, const hprqp_buf *x,
                            hprqp_buf *Qx);
/* The same product in the precision the solver's iteration uses:
 * HPRQP_PREC_FP64 as hprqp_qop_apply; HPRQP_PREC_MIXED with FP32 matrix
 * data (CSR, LASSO, low rank) and FP64 accumulation, and for QAP the
 * split-BF16 three-term MFMA product accumulated in FP32.  x and Qx are FP64
 * buffers of n doubles.  No counterpart in the reference (FP64 only). */
int hprqp_qop_apply_prec(hprqp_ctx *ctx, hprqp_qop *q, const hprqp_buf *x,
                         hprqp_buf *Qx, int precision);
/* QAP operator: the kernel family its applies outside a solve take
 * (hprqp_params.qap_family codes; 0 = the size rule), and in *fp64_family /
 * *bf16_family (either may be NULL) the family its FP64 and split-BF16
 * applies then run, which is the size rule's where the one asked for cannot
 * run on the operator.  A solve replaces the setting with its own
 * (params.qap_family, or the autotuned family).  HPRQP_E_ARG for another
 * kind of operator or an unknown code.  No counterpart in the reference (one
 * CPU/GPU product). */
int hprqp_qop_set_qap_family(hprqp_ctx *ctx, hprqp_qop *q, int32_t family,
                             int32_t *fp64_family, int32_t *bf16_family);
/* get_problem_size / get_temp_size / get_operator_name */
int hprqp_qop_info(const hprqp_qop *q, int64_t *n, int64_t *temp_size,
                   const char **name);

/* ---- model (QP_info_cpu after qp_formulation, src/utils.jl:67-188) ---- */
/* The model keeps its own device copy of the raw (unscaled) data; Q is
 * referenced, not copied, and must outlive the model.  A is m x n CSR. */
int hprqp_model_create(hprqp_ctx *ctx, int32_t n, int32_t m, hprqp_qop *Q,
                       int64_t a_nnz, const int32_t *a_rowptr,
                       const int32_t *a_colidx, const double *a_vals,
                       int index_base, const double *c, const double *AL,
                       const double *AU, const double *l, const double *u,
                       double obj_constant, double lambda, hprqp_model **out);
int hprqp_model_destroy(hprqp_ctx *ctx, hprqp_model *model);

/* ---- solve (optimize -> solve, src/algorithm.jl:1511-1554, 1904-2039) -- */
/* x0 (length n) and y0 (length m) are optional warm starts in the original
 * space (params.initial_x / initial_y).  x/y/z/w_out are optional host
 * outputs in the original space.  hist (optional) receives up to hist_cap
 * rows of 8 doubles per residual evaluation: iter, err_Rp, err_Rd, rel_gap,
 * primal_obj, dual_obj, sigma, restart_flag. */
int hprqp_solve(hprqp_ctx *ctx, const hprqp_model *model,
                const hprqp_params *params, const double *x0, const double *y0,
                double *x_out, double *y_out, double *z_out, double *w_out,
                double *hist, int64_t hist_cap, hprqp_result *result);

/* ---- step-level interface: the host runs the iteration loop ------------
 * HPRQP.jl keeps its own loop (solve, src/algorithm.jl:1904-2039): the
 * restart tests, the sigma update and the termination tests run on the
 * host, and each device operation of an iteration is one call below.  A
 * workspace (HPRQP_workspace_gpu, src/structs.jl:394-459) holds the scaled
 * problem and every iteration vector on the device.  hprqp_solve runs the
 * same device operations in the same order, so a host loop that follows
 * solve()'s cadence reproduces its iterates exactly.  All calls are
 * asynchronous on the context stream except those returning host data. */
typedef struct hprqp_ws hprqp_ws;

#define HPRQP_PATH_STANDARD 0 /* zxw1, tempv (m > 0), y, w2 */
#define HPRQP_PATH_NOC 1      /* noc_zxw, noc_tempv, y  (> 80 % free variables) */
#define HPRQP_PATH_LASSO 2    /* lasso */
#define HPRQP_PATH_LP 3       /* lp_zx, lp_y  (Q = 0) */

typedef struct {
    int32_t n, m;
    int32_t Q_is_diag;    /* check_Q_diagonal after scaling */
    int32_t noC;          /* > 80 % free variables (unscaled bounds) */
    int32_t path;         /* HPRQP_PATH_*: the half-steps main_update runs */
    int32_t precision;    /* the workspace's HPRQP_PREC_* */
    double lambda_max_A;  /* power method x eig_factor, as solve uses it */
    double lambda_max_Q;  /* (diagonal Q: max Q_ii, no eig_factor) */
    double sigma0;        /* src/algorithm.jl:839-851 */
    double b_scale, c_scale, norm_b, norm_c;
    int64_t power_iter_A, power_iter_Q;
    double setup_time;    /* scaling + allocation */
    double power_time;
} hprqp_ws_info;

/* scalars of one half-step (main_update_gpu!, src/algorithm.jl:1250-1279):
 * h1 = 1/(inner + 2), h2 = 1 - h1; full = 1 writes the full set (dx, x_bar,
 * z_bar, dy, s, ...) that residual and restart steps read */
typedef struct {
    double sigma;
    double h1, h2;
    double lambda_max_A, lambda_max_Q;
    int32_t full;
    int32_t reserved;
} hprqp_step_params;

/* compute_residuals! (src/algorithm.jl:37-133), in the original space */
typedef struct {
    double err_Rp, err_Rd, rel_gap, primal_obj, dual_obj, kkt;
} hprqp_kkt_out;

/* the workspace vectors (HPRQP_workspace_gpu's fields, src/structs.jl:394-459)
 * hprqp_ws_download returns and hprqp_ws_upload sets (scaled space) */
#define HPRQP_VEC_X_BAR 0     /* n */
#define HPRQP_VEC_Y_BAR 1     /* m */
#define HPRQP_VEC_Z_BAR 2     /* n */
#define HPRQP_VEC_W_BAR 3     /* n */
#define HPRQP_VEC_ATDY 4      /* n: A^T dy after hprqp_sigma_dots */
#define HPRQP_VEC_QATDY 5     /* n: Q A^T dy after hprqp_sigma_dots */
#define HPRQP_VEC_DIAG_Q 6    /* n: diagonal of the scaled Q (download only) */
#define HPRQP_VEC_X 7         /* n */
#define HPRQP_VEC_W 8         /* n */
#define HPRQP_VEC_Y 9         /* m */
#define HPRQP_VEC_ATY 10      /* n: Halpern-averaged A^T y */
#define HPRQP_VEC_QW 11       /* n */
#define HPRQP_VEC_X_HAT 12    /* n */
#define HPRQP_VEC_DX 13       /* n */
#define HPRQP_VEC_DW 14       /* n */
#define HPRQP_VEC_DY 15       /* m */
#define HPRQP_VEC_S 16        /* m */
#define HPRQP_VEC_ATY_BAR 17  /* n */
#define HPRQP_VEC_TEMPV 18    /* n */
#define HPRQP_VEC_QW_BAR 19   /* n */
#define HPRQP_VEC_LAST_X 20   /* n: Halpern anchors */
#define HPRQP_VEC_LAST_Y 21   /* m */
#define HPRQP_VEC_LAST_W 22   /* n */
#define HPRQP_VEC_LAST_ATY 23 /* n */
#define HPRQP_VEC_LAST_QW 24  /* n */
#define HPRQP_VEC_FACT1 25    /* n: diagonal-Q factors (update_Q_factors) */
#define HPRQP_VEC_FACT2 26    /* n */
#define HPRQP_VEC_FACT 27     /* n */

/* Scaling, check_Q_diagonal, row-product shapes (autotuned as in solve),
 * allocation, power method, sigma_0, the noC decision, then
 * process_initial_points! with x0 / y0 (optional, original space):
 * everything solve() does before its loop (src/algorithm.jl:1904-1976). */
int hprqp_ws_create(hprqp_ctx *ctx, const hprqp_model *model,
                    const hprqp_params *params, const double *x0,
                    const double *y0, hprqp_ws **out, hprqp_ws_info *info);
int hprqp_ws_destroy(hprqp_ctx *ctx, hprqp_ws *ws);
/* one whole iteration along info.path (main_update_gpu!) */
int hprqp_step(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp);
/* the half-steps themselves (src/kernels.jl:261-1426) */
int hprqp_step_zxw1(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp);
int hprqp_step_tempv(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp);
int hprqp_step_y(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp);
int hprqp_step_w2(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp);
int hprqp_step_noc_zxw(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp);
int hprqp_step_noc_tempv(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp);
int hprqp_step_lasso(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp);
int hprqp_step_lp_zx(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp);
int hprqp_step_lp_y(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp);
/* residuals of the current (x_bar, y_bar, z_bar, s) at iteration `iter`
 * (iter 0 adds the bound violation, src/algorithm.jl:118-124); the one
 * readback of a check step */
int hprqp_kkt_reduce(hprqp_ctx *ctx, hprqp_ws *ws, int64_t iter,
                     hprqp_kkt_out *out);
/* compute_M_norm! (src/algorithm.jl:689-739) of the current (dx, dy, dw)
 * with sp's sigma and lambdas */
int hprqp_mnorm(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp,
                double *out);
/* the device half of update_sigma! (src/algorithm.jl:323-360): dx, dw, dy,
 * A^T dy from the anchors, Q dw, Q A^T dy, then out[6] = <dx,dx>,
 * <Q dw, A^T dy>, <Q dw, Q dw>, <dw, Q dw>, <A^T dy, Q A^T dy>, <dy,dy> */
int hprqp_sigma_dots(hprqp_ctx *ctx, hprqp_ws *ws, double *out);
/* the sigma estimate of a diagonal Q with constraints (the minimiser in
 * [1e-12, 1e12] of unified_golden_Q_diag's a x + b/x + x^2 sum_i c_i d_i /
 * (1 + x Q_i), src/unified_operations.jl:348-386) over A^T dy and Q A^T dy
 * of the last hprqp_sigma_dots, as hprqp_solve takes it: the golden-section
 * search on host copies for n < 2,048, from there a search for the sign
 * change of the derivative on the device (ten passes over 64 points).
 * a, b > 0; HPRQP_E_STATE when Q is not diagonal or m = 0. */
int hprqp_sigma_diag_estimate(hprqp_ctx *ctx, hprqp_ws *ws, double a, double b,
                              double *est);
/* unified_update_Q_factors! for a new sigma (diagonal Q only) */
int hprqp_update_Q_factors(hprqp_ctx *ctx, hprqp_ws *ws, double sigma);
/* the device half of do_restart! (src/algorithm.jl:522-555): anchors and
 * iterates from the averaged iterates, A^T y_bar */
int hprqp_restart(hprqp_ctx *ctx, hprqp_ws *ws);
/* collect_results! (src/algorithm.jl:611-643): x, y, z, w in the original
 * space (each optional) */
int hprqp_ws_collect(hprqp_ctx *ctx, hprqp_ws *ws, double *x, double *y,
                     double *z, double *w);
int hprqp_ws_download(hprqp_ctx *ctx, hprqp_ws *ws, int which, double *host,
                      int64_t len);
/* set a workspace vector (a host loop's ws.x .= ..., for instance); the
 * QAP one-apply form's carried products are formed again at the next step */
int hprqp_ws_upload(hprqp_ctx *ctx, hprqp_ws *ws, int which, const double *host,
                    int64_t len);

/* ---- batched solves (many instances, one launch per check interval) ----
 * Instance i: Q_i = diag(d_i) + U_i U_i^T (U_i n_i x k_i column-major),
 * sparse A_i (m_i x n_i CSR), c_i, l_i, u_i, AL_i, AU_i.  Arrays are packed
 * instance after instance: U at offsets sum_j n_j k_j, n-vectors at sum_j n_j,
 * m-vectors at sum_j m_j, row pointers at sum_j (m_j + 1) holding offsets
 * local to the instance's nonzeros (which are packed at sum_j nnz_j), column
 * indices local to the instance.  Each instance runs the reference's
 * iteration (solve, src/algorithm.jl:1904-2039) with its own restart and
 * sigma state and retires when its relative KKT residual < stoptol.  No
 * counterpart in the reference, which solves one instance per call. */
typedef struct hprqp_batch hprqp_batch;
typedef struct {
    double time;              /* wall seconds, power method to last retirement */
    double power_time;
    int64_t solved;           /* instances that reached stoptol */
    int64_t intervals;        /* check intervals launched */
    int64_t max_iter_reached; /* iteration count of the last interval */
    int64_t total_iter;       /* sum of per-instance iteration counts */
    double predict_time;      /* hprqp_batch_solve_warm: seconds computing the
                                 learned warm start (not part of time) */
    double interval_time;     /* seconds of the check-interval launches (HIP
                                 events around each replay: the interval
                                 kernel and its control / record copies) */
} hprqp_batch_result;

int hprqp_batch_create(hprqp_ctx *ctx, int32_t count, const int32_t *n,
                       const int32_t *m, const int32_t *k, const double *U,
                       const double *diag, int64_t a_nnz,
                       const int32_t *a_rowptr,
                       const int32_t *a_colidx, const double *a_vals,
                       const double *c, const double *l, const double *u,
                       const double *AL, const double *AU,
                       const double *obj_constant, hprqp_batch **out);
/* Batched instances whose Q is sparse (Qmap! for the CSR operator,
 * src/Q_operators/sparse_matrix_operator.jl:27-51): per instance b,
 *   min 1/2 x'Q_b x + c_b'x  s.t.  AL_b <= A_b x <= AU_b,  l_b <= x <= u_b,
 * Q_b given as n_b x n_b CSR (0-based, its rowptr starting at 0), packed back
 * to back in q_rowptr (n_b + 1 entries each) / q_colidx / q_vals (q_nnz in
 * all); A_b as in hprqp_batch_create.  Each solve scales every instance as
 * the single solve does (Ruiz and Pock-Chambolle, src/utils.jl:350-804;
 * (Q + Q')/2 after scaling) and runs the same block-per-instance check
 * intervals with CSR products of Q.  The batch runs the standard path only:
 * an instance with a diagonal or empty Q, or with more than 80% free
 * variables (the noC path), is refused with HPRQP_E_UNSUP.  No counterpart in
 * the reference (one instance per call); hprqp_batch_predict and
 * hprqp_batch_solve_warm take only the diagonal-plus-low-rank kind. */
int hprqp_batch_create_csr(hprqp_ctx *ctx, int32_t count, const int32_t *n,
                           const int32_t *m, int64_t q_nnz,
                           const int32_t *q_rowptr, const int32_t *q_colidx,
                           const double *q_vals, int64_t a_nnz,
                           const int32_t *a_rowptr, const int32_t *a_colidx,
                           const double *a_vals, const double *c,
                           const double *l, const double *u, const double *AL,
                           const double *AU, const double *obj_constant,
                           hprqp_batch **out);
int hprqp_batch_destroy(hprqp_ctx *ctx, hprqp_batch *batch);
/* precision: HPRQP_PREC_FP64 (all FP64) or HPRQP_PREC_MIXED (FP32 U, d, A,
 * c, l, u, AL, AU; FP64 iterates, anchors, accumulation and KKT).  x0/y0
 * optional packed warm starts; outputs optional, packed. */
int hprqp_batch_solve(hprqp_ctx *ctx, hprqp_batch *batch,
                      const hprqp_params *params, const double *x0,
                      const double *y0, double *x_out, double *y_out,
                      int32_t *status_out, int64_t *iter_out,
                      double *obj_out, double *kkt_out,
                      hprqp_batch_result *result);

/* ---- learned warm start: per-variable MLP n_in -> h -> h -> h -> 2 (ReLU,
 * FP32), shared across the variables of an instance.  Its predictions are
 * passed as initial_x / initial_y (process_initial_points!,
 * src/algorithm.jl:1783-1836); no counterpart in the reference, which only
 * takes user-supplied warm starts.  Training data are (groups x group) rows:
 * X row-major rows x n_in, tx one target per row (output 0), ty one target
 * per group of `group` consecutive rows (mean of output 1 over the group).
 * loss = mean_r (o_r0 - tx_r)^2 + wy * mean_g (mean_{r in g} o_r1 - ty_g)^2. */
typedef struct hprqp_mlp hprqp_mlp;
int hprqp_mlp_create(hprqp_ctx *ctx, int32_t n_in, int32_t hidden,
                     uint64_t seed, hprqp_mlp **out);
int hprqp_mlp_destroy(hprqp_ctx *ctx, hprqp_mlp *mlp);
int64_t hprqp_mlp_param_count(const hprqp_mlp *mlp);
/* parameters flattened as W1 (n_in x h, row-major), b1, W2, b2, W3, b3,
 * W4 (h x 2), b4 */
int hprqp_mlp_get_params(hprqp_ctx *ctx, const hprqp_mlp *mlp, float *host);
int hprqp_mlp_set_params(hprqp_ctx *ctx, hprqp_mlp *mlp, const float *host);
/* out: rows x 2 */
int hprqp_mlp_forward(hprqp_ctx *ctx, hprqp_mlp *mlp, int64_t rows,
                      const float *X, float *out);
/* loss and its gradient (flattened like the parameters) on one batch */
int hprqp_mlp_loss_grad(hprqp_ctx *ctx, hprqp_mlp *mlp, int64_t groups,
                        int32_t group, const float *X, const float *tx,
                        const float *ty, float wy, double *loss, float *grad);
/* Adam over shuffled minibatches of batch_groups groups; loss_history
 * (optional) receives the mean loss of each epoch */
int hprqp_mlp_train(hprqp_ctx *ctx, hprqp_mlp *mlp, int64_t groups,
                    int32_t group, const float *X, const float *tx,
                    const float *ty, float wy, int32_t batch_groups,
                    int32_t epochs, float lr, uint64_t seed,
                    double *loss_history);

/* Learned warm start of every instance of a batch, computed on the device:
 * the 12 per-variable features below, normalised as (f - feat_mean) /
 * feat_std, through the MLP (n_in = 12), then x0_i = clamp(o_i0, 0, 1) u_i
 * and y0 = mean_i(o_i1) / y_scale on each of the instance's rows (u_i := 1
 * where infinite).  Features of variable i (cm, cs: mean and population
 * standard deviation of the instance's c, cs += 1e-12; ubar: column means of
 * U; q_ii = d_i + |U_i.|^2; r_i = -c_i / q_ii; rank(v)_i = #{j: v_j < v_i}
 * + #{j < i: v_j == v_i}, over max(n - 1, 1)):
 *   (c_i - cm)/cs, 20 c_i, 10 d_i, 3 (q_ii - d_i), 30 U_i.ubar, 0.5 r_i,
 *   rank(c)_i, rank(-r)_i, 100/n, k/50, min(n u_i, 100)/100, 50 cs.
 * hprqp_batch_predict returns the packed (x0, y0) to the host (either may be
 * NULL); hprqp_batch_solve_warm feeds them to the solve without leaving the
 * device and reports the prediction time in result->predict_time.  No
 * counterpart in the reference. */
int hprqp_batch_predict(hprqp_ctx *ctx, hprqp_batch *batch, hprqp_mlp *mlp,
                        const float *feat_mean, const float *feat_std,
                        double y_scale, double *x0_out, double *y0_out);
int hprqp_batch_solve_warm(hprqp_ctx *ctx, hprqp_batch *batch,
                           hprqp_mlp *mlp, const float *feat_mean,
                           const float *feat_std, double y_scale,
                           const hprqp_params *params, double *x_out,
                           double *y_out, int32_t *status_out,
                           int64_t *iter_out, double *obj_out,
                           double *kkt_out, hprqp_batch_result *result);

/* ---- measurement (determine_spmv_mode's timing loop,
 *      src/algorithm.jl:1557-1746) ------------------------------------------
 * Prepares the solver state for `model` exactly as hprqp_solve does (scaling,
 * power method, autotuned SpMV lanes), then times each fused kernel of the
 * standard step with HIP events on the context stream, `reps` back-to-back
 * launches per kernel.  ms_out[0..5] = average ms of: z/x/w1 step (Q w),
 * tempv step (Q w_bar), y step (A tempv), w2 step (A^T y_bar), one whole
 * partial iteration, one residual evaluation.  *n_out = number of slots
 * filled (4 per-kernel slots are 0 when the problem does not take the
 * standard path).  Iterate values are not meaningful afterwards. */
int hprqp_profile_step(hprqp_ctx *ctx, const hprqp_model *model,
                       const hprqp_params *params, int reps, double *ms_out,
                       int *n_out);

#ifdef __cplusplus
}
#endif
#endif /* HPRQP_H */
