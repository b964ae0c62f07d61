// hprqp_internal.hpp -- host-side objects behind the C ABI (include/hprqp.h).
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "hprqp.h"

namespace hprqp {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HPRQP_HIP(expr)                                                              \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess)                                                        \
            throw ::hprqp::Error(HPRQP_E_HIP, std::string(#expr) + " failed: " +     \
                                                  hipGetErrorString(_e));            \
    } while (0)

#define HPRQP_REQUIRE(cond, code, msg)                           \
    do {                                                         \
        if (!(cond)) throw ::hprqp::Error((code), (msg));        \
    } while (0)

// RAII device array
template <class T>
struct DevVec {
    T *p = nullptr;
    int64_t n = 0;
    DevVec() = default;
    explicit DevVec(int64_t n_) { alloc(n_); }
    void alloc(int64_t n_) {
        release();
        n = n_;
        if (n > 0) {
            hipError_t e = hipMalloc((void **)&p, sizeof(T) * (size_t)n);
            if (e != hipSuccess) {
                p = nullptr;
                n = 0;
                throw Error(HPRQP_E_ALLOC, std::string("hipMalloc failed: ") + hipGetErrorString(e));
            }
        }
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevVec() { release(); }
    DevVec(const DevVec &) = delete;
    DevVec &operator=(const DevVec &) = delete;
    DevVec(DevVec &&o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DevVec &operator=(DevVec &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
};

// Device CSR (0-based, int32 indices).  vals: FP64; vals32: optional FP32
// copy for mixed precision.  lanes: rows-per-group choice of the SpMV.
struct DevCsr {
    int32_t nrow = 0, ncol = 0;
    int64_t nnz = 0;
    DevVec<int32_t> rp, ci;
    DevVec<double> v;
    DevVec<float> v32;
    int lanes = 1;
    // column-chunked product (built on its first use): per-chunk row segment
    // starts, (nchunk + 1) x nrow, and the per-chunk sums, nchunk x nrow
    mutable DevVec<int32_t> cseg;
    mutable DevVec<double> cpart;
    mutable int32_t nchunk = 0;
    // value table form (matrices with at most 256 distinct values before
    // scaling): one byte per nonzero indexing dict, the scaling applied as
    // the row factor rsc and the column factor csc (entry = dict[code] *
    // rsc[row] * csc[col]); dict_rb > 0 selects it, rows per block of the
    // CSR-stream shape
    int dict_rb = 0;
    int ndict = 0;
    DevVec<uint8_t> code;
    DevVec<double> dict;
    const double *rsc = nullptr, *csc = nullptr;
    // diagonal form (a matrix whose nonzeros lie on at most 8 diagonals j =
    // floor(i * ncol / nrow) + off): dia_nd > 0 selects it; dia_off the
    // offsets, ascending; dia_v (FP64) and dia_v32 (FP32) the diagonals,
    // dia_nd x nrow, zero where the pattern has no entry
    int dia_nd = 0;
    int32_t dia_off[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    DevVec<double> dia_v;
    DevVec<float> dia_v32;
};

// Host-side CSR (sorted column indices), used while building models.
struct HostCsr {
    int32_t nrow = 0, ncol = 0;
    std::vector<int32_t> rp, ci;
    std::vector<double> v;
};

// Kernel timing carried by the launches themselves (hipExtLaunchKernel):
// while armed, the first launch records `start` as it begins and every
// launch records `stop` as it ends, so start..stop spans exactly the kernels
// of one timed step -- the durations a kernel trace reports -- without the
// gaps of separately recorded events.
struct LaunchTiming {
    hipEvent_t start = nullptr, stop = nullptr;
    bool armed = false, first = true;
};
inline thread_local LaunchTiming g_launch_timing;

template <class... KA, class... PA>
inline void launch_k(void (*kernel)(KA...), dim3 grid, dim3 block, uint32_t shm, hipStream_t s, PA &&...args) {
    LaunchTiming &t = g_launch_timing;
    if (t.armed) {
        hipExtLaunchKernelGGL(kernel, grid, block, shm, s, t.first ? t.start : nullptr, t.stop, 0u,
                              static_cast<KA>(args)...);
        t.first = false;
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shm, s, static_cast<KA>(args)...);
    }
}

HostCsr host_csr_from(int32_t nrow, int32_t ncol, int64_t nnz, const int32_t *rowptr,
                      const int32_t *colidx, const double *vals, int index_base);
HostCsr host_csr_transpose(const HostCsr &a);
HostCsr host_csr_symmetric_pattern(const HostCsr &q);
bool dict_encode(const std::vector<double> &v, std::vector<double> &dict, std::vector<uint8_t> &code);
std::vector<int32_t> host_csr_transpose_positions(const HostCsr &s);
void upload_csr(hipStream_t s, const HostCsr &h, DevCsr &d);
int default_lanes(const HostCsr &h);

}  // namespace hprqp

struct hprqp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // a second stream (default priority) for work independent of the
    // solve's stream, and the event that orders it after that stream
    hipStream_t stream2 = nullptr;
    hipEvent_t fork = nullptr;
    std::string err;
    double *pinned = nullptr;  // small pinned staging area for scalar readback
    size_t pinned_count = 0;
    double *scratch = nullptr;  // this device's static scratch array (hprqp_qop_apply_noalloc)
};

struct hprqp_buf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct hprqp_csr {
    hprqp::DevCsr a;
    hprqp::DevCsr at;  // stored transpose (may be empty)
    bool has_t = false;
};

enum QopKind { QOP_CSR = 0, QOP_LASSO = 1, QOP_QAP = 2, QOP_USER = 3, QOP_LOWRANK = 4 };

struct hprqp_qop {
    int kind = QOP_CSR;
    int32_t n = 0;
    // CSR: Q (n x n) with the host copy of its pattern kept for scaling
    hprqp::HostCsr hq;
    hprqp::DevCsr q;
    // LASSO: data matrix and its transpose, temp of length rows
    hprqp::DevCsr la, lat;
    hprqp::DevVec<double> ltemp;
    hprqp::DevVec<float> ltemp32;
    // QAP: dense column-major nq x nq
    int32_t nq = 0;
    hprqp::DevVec<double> qa, qb, qs, qt, qtemp;
    hprqp::DevVec<float> qtemp32;
    // QAP operands of the MFMA kernels (built by qap_prepare): split-BF16
    // pieces padded to qnp = nq rounded up to 32, or to 128 where the
    // blocked passes run (nq >= 256) (A, S row-major; B, T
    // column-major); the diagonals of S / T when those are diagonal (the
    // full matrices are then not read); qtemp / qtemp32 hold the two n x n
    // intermediates of the FP64 / split-BF16 passes
    int32_t qnp = 0;
    bool qs_diag = false, qt_diag = false;
    bool qa_exact = false, qb_exact = false, qs_exact = false, qt_exact = false;
    hprqp::DevVec<double> qsd, qtd;
    hprqp::DevVec<uint16_t> qa0, qa1, qb0, qb1, qs0, qs1, qt0, qt1;
    // large-n split-BF16 passes: BF16 pieces of the data operands, each
    // qnp x qnp with zero padding, three pieces back to back: X transposed
    // (qxp), Y row-major (qyp, written by pass 1), X row-major (qxr, full T)
    hprqp::DevVec<uint16_t> qxp, qyp, qxr;
    // kernel family of the FP64 and split-BF16 passes (QapFamily): 0 by
    // size (n < 256 one-tile, then blocked, LDS-tiled from n = 1024 with
    // diagonal S and T).  qfam* is the family the launches take, set per
    // solve from the caller's choice or from qtuned*, the fastest family
    // the solver's timing found for this operator (0 = not timed yet)
    int qfam64 = 0, qfam16 = 0;
    mutable int qtuned64 = 0, qtuned16 = 0;
    // diagonal plus low rank: Q = diag(d) + U U^T, U n x k column-major
    int32_t lr_k = 0;
    hprqp::DevVec<double> lr_U, lr_d, lr_t;
    hprqp::DevVec<float> lr_U32, lr_d32;
    // user plugin
    hprqp_qop_vtable vt{};
    void *user = nullptr;
    std::string name;
};

struct hprqp_model {
    int32_t n = 0, m = 0;
    hprqp_qop *Q = nullptr;
    hprqp::HostCsr A;  // host copy (m x n), sorted
    std::vector<double> c, AL, AU, l, u;
    double obj_constant = 0.0, lambda = 0.0;
    // Q pattern symmetric completion and transpose position map (CSR Q only)
    hprqp::HostCsr Qsym;
    std::vector<int32_t> q_tpos;
    // SpMV lanes chosen by the autotuner (0 = not tuned yet), per precision
    // mode (HPRQP_PREC_FP64 / MIXED / FP32) for Q, A, A^T and a LASSO
    // operator's data matrix and its transpose
    mutable int tuned_lanes[3][5] = {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}};
    // value tables of A (A^T's codes in its own order) and of Q's
    // symmetrised values (Qsym, (q_ij + q_ji) / 2), empty when a matrix has
    // more than 256 distinct values; the FP64 value-table shape the
    // autotuner chose for Q, A, A^T (0 not tuned, -1 none, else rows per block)
    std::vector<double> a_dict, q_dict;
    std::vector<uint8_t> a_code, at_code, q_code;
    mutable int tuned_dict[3] = {0, 0, 0};
    // diagonal forms of Q (Qsym's pattern), A and A^T: dia_nd diagonals at
    // offsets dia_off (0 = not examined yet, -1 = more than 8 diagonals or
    // too much padding); tuned_dia per precision mode and matrix (0 not
    // tuned, 1 taken, -1 not taken)
    mutable int dia_nd[3] = {0, 0, 0};
    mutable int32_t dia_off[3][8] = {};
    mutable int tuned_dia[3][3] = {};
    // device copies of the raw data (A, its transpose, Qsym, the vectors
    // and the value-table codes), made at the first solve; each solve
    // copies them device to device into its scaled data instead of
    // uploading the host copies again (and transposing A on the host)
    struct Raw {
        bool ready = false;
        hprqp::DevCsr A, AT, Q;
        hprqp::DevVec<double> c, l, u, AL, AU, a_dict, q_dict;
        hprqp::DevVec<uint8_t> a_code, at_code, q_code;
    };
    mutable Raw raw;
};

struct hprqp_batch {
    int32_t count = 0;
    // qkind 0: Q_i = diag(d_i) + U_i U_i' (operator, no scaling); 1: Q_i
    // sparse CSR (scaled per instance at each solve, as the single solve)
    int qkind = 0;
    int64_t Ntot = 0, Mtot = 0, Utot = 0, nnzA = 0, nnzQ = 0;
    // qkind 1: the instances' Q with symmetric patterns, packed (row
    // pointers global, instance b's at qrp_off[b]) with each entry's
    // transpose position
    std::vector<int64_t> qrp_off;
    hprqp::DevVec<int64_t> dqrp_off;
    hprqp::DevVec<int32_t> q_rp, q_ci, q_tpos;
    hprqp::DevVec<double> q_v;
    std::vector<int32_t> n, m, k;
    std::vector<int64_t> noff, moff, uoff;
    std::vector<int64_t> auoff;  // A_i U_i (m_i x k_i) at sum_j m_j k_j
    int64_t AUtot = 0;
    hprqp::DevVec<int64_t> dauoff;
    std::vector<double> obj_constant, norm_b2, norm_c2, norm_b_inf, norm_c_inf;
    hprqp::DevVec<int32_t> dn, dm, dk;
    hprqp::DevVec<int64_t> dnoff, dmoff, duoff, darp_off, datrp_off;
    hprqp::DevVec<double> U, d, a_v, at_v, c, l, u, AL, AU;
    hprqp::DevVec<float> U32, d32, a_v32, at_v32, c32, l32, u32, AL32, AU32;
    hprqp::DevVec<int32_t> a_rp, a_ci, at_rp, at_ci;
    // dispatch order of the per-instance workgroups: instances sorted by
    // estimated per-iteration cost, largest first, so the big ones start
    // early and the small ones fill in behind them
    hprqp::DevVec<int32_t> dorder;
    hprqp::DevVec<double> dc0, dnbi, dnci;  // objective constants and residual norms
};

struct hprqp_mlp {
    int32_t nin = 0, hid = 0, nout = 2;
    int64_t np = 0;            // parameter count
    int64_t off_W[4], off_b[4];
    int32_t in_w[4], out_w[4];
    hprqp::DevVec<float> p, g, am, av;
    int64_t adam_t = 0;
    int64_t cap = 0;           // activation capacity in rows
    hprqp::DevVec<float> act[5], dact[5];
    hprqp::DevVec<double> part;
};

namespace hprqp {
// x0_dev / y0_dev: optional packed warm starts already on the device
// (taken instead of the host x0 / y0 when given)
void batch_solve(hprqp_ctx *ctx, hprqp_batch *B, const hprqp_params *p, const double *x0, const double *y0,
                 const double *x0_dev, const double *y0_dev, double *x_out, double *y_out, int32_t *status_out,
                 int64_t *iter_out, double *obj_out, double *kkt_out, hprqp_batch_result *res);
// MLP forward of `rows` device rows X (rows x nin) into device out (rows x 2)
void mlp_forward_dev(hipStream_t s, hprqp_mlp *M, const float *X, int64_t rows, float *out);
// learned warm start of every instance of a batch, computed on the device
// (hprqp_warm.hip); x0 (Ntot) and y0 (Mtot) are device arrays
void batch_predict(hprqp_ctx *ctx, hprqp_batch *B, hprqp_mlp *M, const float *feat_mean, const float *feat_std,
                   double y_scale, double *x0, double *y0);
constexpr int kWarmFeatures = 12;
// solver entry (hprqp_solver.hip)
void solve(hprqp_ctx *ctx, const hprqp_model *model, const hprqp_params *params, const double *x0,
           const double *y0, double *x_out, double *y_out, double *z_out, double *w_out,
           double *hist, int64_t hist_cap, hprqp_result *res);
// step-level workspace (hprqp_ws_*, hprqp_step_*; defined in hprqp_solver.hip)
hprqp_ws *ws_create(hprqp_ctx *ctx, const hprqp_model *md, const hprqp_params *p, const double *x0,
                    const double *y0);
void ws_info(const hprqp_ws *w, hprqp_ws_info *info);
hprqp_ctx *ws_ctx(const hprqp_ws *w);
void ws_step(hprqp_ws *w, int kind, const hprqp_step_params *sp);
void ws_kkt(hprqp_ws *w, int64_t iter, hprqp_kkt_out *out);
double ws_mnorm(hprqp_ws *w, const hprqp_step_params *sp);
void ws_sigma_dots(hprqp_ws *w, double *out);
double ws_sigma_diag(hprqp_ws *w, double a, double b);
void ws_q_factors(hprqp_ws *w, double sigma);
void ws_restart(hprqp_ws *w);
void ws_collect(hprqp_ws *w, double *x, double *y, double *z, double *wv);
void ws_download(hprqp_ws *w, int which, double *host, int64_t len);
void ws_upload(hprqp_ws *w, int which, const double *host, int64_t len);
void ws_destroy(hprqp_ws *w);
// per-kernel timing of the iteration (hprqp_profile_step)
void profile_step(hprqp_ctx *ctx, const hprqp_model *model, const hprqp_params *params, int reps,
                  double *ms_out, int *n_out);
// operator applies exposed through the ABI (FP64).  scratch (optional)
// replaces the operator's own temporaries (qop_scratch_doubles of them) and
// then no shape that builds its own buffers is used
void qop_apply_fp64(hprqp_ctx *ctx, hprqp_qop *q, const double *x, double *Qx, double *scratch = nullptr);
int64_t qop_scratch_doubles(const hprqp_qop *q);
void csr_spmv_fp64(hprqp_ctx *ctx, const DevCsr &a, const double *x, double *y);
// QAP dense products (hprqp_qap.hip).  qap_prepare builds the MFMA operands
// from qa/qb/qs/qt (host copies given); qap_apply_f64 is the FP64 product,
// qap_apply_bf16x3 the split-BF16 three-term product accumulated in FP32.
void qap_prepare(hipStream_t s, hprqp_qop *q, const double *A, const double *B, const double *S, const double *T);
// tmp (optional) replaces the operator's intermediates (2 n^2 doubles)
void qap_apply_f64(hipStream_t s, const hprqp_qop *q, const double *X, double *QX, double *tmp = nullptr);
template <class IT, class OT>
void qap_apply_bf16x3(hipStream_t s, const hprqp_qop *q, const IT *X, OT *QX);
// kernel family of the passes: the fastest of those that can run on the
// operator, timed (qap_tune), and the one its launches take now
// (qap_family_of; hprqp_params.qap_family codes)
int qap_tune(hipStream_t s, hprqp_qop *q, bool bf16, bool verbose);
int qap_family_of(const hprqp_qop *q, bool bf16);
// the reduced-precision operator apply of the MIXED path (FP32 matrix data;
// split-BF16 x3 MFMA for QAP), FP64 in and out
void qop_apply_mixed(hprqp_ctx *ctx, hprqp_qop *q, const double *x, double *Qx);
// diagonal plus low rank Qx = d.*x + U (U^T x)  (hprqp_lowrank.hip)
template <class MT, class XT, class YT>
void lowrank_apply(hipStream_t s, int32_t n, int32_t k, const MT *U, const MT *d, const XT *x, double *t, YT *Qx);
}  // namespace hprqp
