// hprqp_solver.hip -- the HPR-QP solve loop on one MI355X.
//
// Host side restates solve() (src/algorithm.jl:1904-2039): scaling, power
// method, warm start, then the iteration loop whose per-step device work is
// the fused kernels of hprqp_kernels.hpp.  Restart and sigma logic stay on
// the host and read back a few scalars per check step, as in the reference.
//
// Precision: Solver<VT, MT> keeps its bulk vectors in VT and reads matrix
// values and problem data in MT.  Solver<double> is the plain FP64 form.
// Solver<double, float> is the mixed form: FP32 matrix values and data, FP64
// vectors and FP64 accumulation.  Solver<float, float> is the FP32 form: FP32
// matrices, FP32 bulk vectors and FP32 accumulation in the row products and
// fused updates.  In every form the Halpern anchors (last_*) are FP64, and the
// residual/KKT evaluation and the dot products of the sigma/restart tests use
// FP64 matrices and FP64 accumulation.  A reduced-precision solve finishes
// with an FP64 polish that continues the same iteration in Solver<double>
// from its state once it stalls near its floor.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <type_traits>

#include "hprqp_internal.hpp"
#include "hprqp_kernels.hpp"
#include "hprqp_qap.hpp"
#include "hprqp_qapds.hpp"

namespace hprqp {

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static inline int nblk(int64_t n) { return (int)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------
template <class AT, class MT, class XT, class Epi>
static void launch_csr(hipStream_t s, int lanes, int32_t nrows, const int32_t *rp, const int32_t *ci,
                       const MT *v, const XT *x, const Epi &epi) {
    if (nrows <= 0) return;
    if (lanes > 64) {  // one block per row
        launch_k(k_csr_rowblock<MT, XT, AT, Epi>, dim3(nrows), dim3(kBlock), 0, s, nrows, rp, ci, v, x, epi);
        HPRQP_HIP(hipGetLastError());
        return;
    }
    if (lanes < 0) {  // CSR-stream: -lanes rows per block
        const int rb = -lanes;
        const int grid = (int)(((int64_t)nrows + rb - 1) / rb);
        launch_k(k_csr_stream<MT, XT, AT, Epi>, dim3(grid), dim3(kBlock), 0, s, nrows, rb, rp, ci, v, x, epi);
        HPRQP_HIP(hipGetLastError());
        return;
    }
    const int64_t threads = (int64_t)nrows * lanes;
    const dim3 grid((unsigned)((threads + kBlock - 1) / kBlock)), blk(kBlock);
    switch (lanes) {
        case 1: launch_k(k_csr<1, MT, XT, AT, Epi>, grid, blk, 0, s, nrows, rp, ci, v, x, epi); break;
        case 2: launch_k(k_csr<2, MT, XT, AT, Epi>, grid, blk, 0, s, nrows, rp, ci, v, x, epi); break;
        case 4: launch_k(k_csr<4, MT, XT, AT, Epi>, grid, blk, 0, s, nrows, rp, ci, v, x, epi); break;
        case 8: launch_k(k_csr<8, MT, XT, AT, Epi>, grid, blk, 0, s, nrows, rp, ci, v, x, epi); break;
        case 16: launch_k(k_csr<16, MT, XT, AT, Epi>, grid, blk, 0, s, nrows, rp, ci, v, x, epi); break;
        case 32: launch_k(k_csr<32, MT, XT, AT, Epi>, grid, blk, 0, s, nrows, rp, ci, v, x, epi); break;
        default: launch_k(k_csr<64, MT, XT, AT, Epi>, grid, blk, 0, s, nrows, rp, ci, v, x, epi); break;
    }
    HPRQP_HIP(hipGetLastError());
}

template <class MT>
static const MT *csr_vals(const DevCsr &a);
template <>
const double *csr_vals<double>(const DevCsr &a) { return a.v.p; }
template <>
const float *csr_vals<float>(const DevCsr &a) { return a.v32.p; }

// the column-chunked form: segment starts built on first use (the pattern
// is fixed for the matrix's life), then the chunk sums and their total
// the column-chunked shape's segment table, built on its first use
static void chunk_prepare(hipStream_t s, const DevCsr &a) {
    if (a.nrow <= 0 || a.nchunk != 0) return;
    const int32_t nc = (int32_t)((a.ncol + kChunkW - 1) / kChunkW);
    a.cseg.alloc((int64_t)(nc + 1) * a.nrow);
    a.cpart.alloc((int64_t)nc * a.nrow);
    const int64_t cnt = (int64_t)(nc + 1) * a.nrow;
    k_chunk_seg<<<(unsigned)((cnt + kBlock - 1) / kBlock), kBlock, 0, s>>>(a.nrow, nc, a.rp.p, a.ci.p, a.cseg.p);
    HPRQP_HIP(hipGetLastError());
    a.nchunk = nc;
}

template <class AT, class MT, class XT, class Epi>
static void launch_chunked(hipStream_t s, const DevCsr &a, const MT *v, const XT *x, const Epi &epi) {
    if (a.nrow <= 0) return;
    chunk_prepare(s, a);
    launch_k(k_csr_chunk<MT, XT, AT>, dim3((unsigned)(((a.nrow + kChunkRows - 1) / kChunkRows) * a.nchunk)),
             dim3(kBlock), 0, s,
             a.nrow, a.ncol, (const int32_t *)a.cseg.p, a.ci.p, v, x, a.cpart.p);
    HPRQP_HIP(hipGetLastError());
    launch_k(k_chunk_sum<AT, Epi>, dim3(nblk((int64_t)a.nrow * 16)), dim3(kBlock), 0, s, a.nrow, a.nchunk,
             (const double *)a.cpart.p, epi);
    HPRQP_HIP(hipGetLastError());
}

// A matrix without nonzeros has no value or index arrays (nothing is
// allocated for them): it takes the one-lane shape, whose row loop reads
// nothing when every row is empty, whatever shape was chosen (the chunked
// shape loads entry 0 of the value and index arrays speculatively).
template <class AT, class XT, class Epi>
static void launch_dict(hipStream_t s, const DevCsr &a, int rb, const XT *x, const Epi &epi) {
    const int grid = (int)(((int64_t)a.nrow + rb - 1) / rb);
    launch_k(k_csr_stream_dict<XT, AT, Epi>, dim3(grid), dim3(kBlock), 0, s, a.nrow, rb, a.rp.p, a.ci.p,
             (const uint8_t *)a.code.p, (const double *)a.dict.p, a.ndict, a.rsc, a.csc, x, epi);
    HPRQP_HIP(hipGetLastError());
}

// the diagonal form's product, in the values' precision MT; false when the
// matrix has no diagonal form in that precision
template <class AT, class MT, class XT, class Epi>
static bool launch_dia(hipStream_t s, const DevCsr &a, const XT *x, const Epi &epi) {
    const MT *dv;
    if constexpr (std::is_same<MT, double>::value)
        dv = a.dia_v.p;
    else
        dv = a.dia_v32.p;
    if (a.dia_nd <= 0 || a.nrow <= 0 || dv == nullptr) return false;
    DiaOff off;
    for (int d = 0; d < kDiaMax; ++d) off.o[d] = d < a.dia_nd ? a.dia_off[d] : 0;
    // one two-element load of x for neighbouring diagonals whose pair starts
    // at an even column in every row, x aligned to two elements
    unsigned pairs = 0;
    const bool even_slope = a.ncol % a.nrow == 0 && (a.ncol / a.nrow) % 2 == 0;
    if (even_slope && ((uintptr_t)x % (2 * sizeof(XT))) == 0)
        for (int d = 0; d + 1 < a.dia_nd; ++d)
            if (off.o[d] % 2 == 0 && off.o[d + 1] == off.o[d] + 1) {
                pairs |= 1u << d;
                ++d;
            }
    if (pairs)
        launch_k(k_dia<true, MT, XT, AT, Epi>, dim3(nblk(a.nrow)), dim3(kBlock), 0, s, a.nrow, a.ncol, a.dia_nd,
                 off, pairs, dv, x, epi);
    else
        launch_k(k_dia<false, MT, XT, AT, Epi>, dim3(nblk(a.nrow)), dim3(kBlock), 0, s, a.nrow, a.ncol, a.dia_nd,
                 off, 0u, dv, x, epi);
    HPRQP_HIP(hipGetLastError());
    return true;
}

template <class AT, class MT, class XT, class Epi>
static void csr_apply(hipStream_t s, const DevCsr &a, const XT *x, const Epi &epi) {
    // the diagonal form, where the autotuner (or the caller) chose it
    if (a.dia_nd > 0 && launch_dia<AT, MT>(s, a, x, epi)) return;
    if constexpr (std::is_same<MT, double>::value && std::is_same<AT, double>::value) {
        // the FP64 iteration's value-table form, where the autotuner chose it
        if (a.dict_rb > 0 && a.nnz > 0 && a.nrow > 0) {
            launch_dict<AT>(s, a, a.dict_rb, x, epi);
            return;
        }
    }
    if (a.nnz == 0)
        launch_csr<AT>(s, 1, a.nrow, a.rp.p, a.ci.p, csr_vals<MT>(a), x, epi);
    else if (a.lanes == kLanesChunked)
        launch_chunked<AT>(s, a, csr_vals<MT>(a), x, epi);
    else
        launch_csr<AT>(s, a.lanes, a.nrow, a.rp.p, a.ci.p, csr_vals<MT>(a), x, epi);
}

template <class F>
static void each(hipStream_t s, int64_t n, const F &f) {
    if (n <= 0) return;
    k_each<F><<<nblk(n), kBlock, 0, s>>>(n, f);
    HPRQP_HIP(hipGetLastError());
}

template <class ST, class AT, class Epi>
static void vec_apply(hipStream_t s, int32_t n, const ST *src, const Epi &epi) {
    if (n <= 0) return;
    launch_k(k_vec<ST, AT, Epi>, dim3(nblk(n)), dim3(kBlock), 0, s, n, src, epi);
    HPRQP_HIP(hipGetLastError());
}

static void ensure_v32(DevCsr &a) {
    if (a.nnz > 0 && a.v32.n != a.nnz) {
        a.v32.alloc(a.nnz);
        const double *src = a.v.p;
        float *dst = a.v32.p;
        each(nullptr, a.nnz, [=] __device__(int64_t k) { dst[k] = (float)src[k]; });
        HPRQP_HIP(hipDeviceSynchronize());
    }
}

// Reduction scratch and pinned readback
struct Reducer {
    hipStream_t s;
    DevVec<double> part, out;
    double *host;
    explicit Reducer(hprqp_ctx *ctx) : s(ctx->stream), host(ctx->pinned) {
        part.alloc((int64_t)kMaxReduceBlocks * 16);
        out.alloc(96);
    }
    // enqueue; results land in out[off .. off+NS+NM)
    template <int NS, int NM, class F>
    void enqueue(int64_t n, const F &f, int off) {
        const int nb = (int)std::min<int64_t>(kMaxReduceBlocks, std::max<int64_t>(1, (n + kBlock - 1) / kBlock));
        k_reduce<NS, NM, F><<<nb, kBlock, 0, s>>>(n, f, part.p);
        HPRQP_HIP(hipGetLastError());
        k_reduce_final<NS, NM><<<1, kBlock, 0, s>>>(nb, part.p, out.p + off);
        HPRQP_HIP(hipGetLastError());
    }
    // copy out[0..cnt) to host and wait
    const double *fetch(int cnt) {
        HPRQP_HIP(hipMemcpyAsync(host, out.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
        HPRQP_HIP(hipStreamSynchronize(s));
        return host;
    }
};

// HIP-event timing of one kernel class during a solve: events bracket each
// launch on the solve's stream; pairs are drained every `pairs` launches.
struct KernelTimer {
    hipStream_t s;
    std::vector<hipEvent_t> ev;
    int used = 0;
    double total_ms = 0.0;
    int64_t count = 0;
    // Only every kStride-th span is timed: a launch that carries events
    // costs the stream about 10 us of idle time (measured: random QP solve
    // 41 -> 53 ms with every z/x/w1 launch timed), so timing one in eight
    // keeps the measured solve within a few percent of an untimed one.
    static constexpr int kStride = 8;
    int64_t spans = 0;
    bool sampling = false;
    explicit KernelTimer(hipStream_t st, int pairs = 512) : s(st), ev(2 * pairs) {
        // timing-only events: no system-scope fence when they are recorded
        for (auto &e : ev) HPRQP_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    }
    ~KernelTimer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    // the kernels launched between begin() and end() carry the event pair
    void begin() {
        sampling = spans++ % kStride == 0;
        if (!sampling) return;
        if (used == (int)ev.size()) flush();
        g_launch_timing = LaunchTiming{ev[used], ev[used + 1], true, true};
    }
    void end() {
        if (!sampling) return;
        sampling = false;
        LaunchTiming &t = g_launch_timing;
        if (t.first) {  // nothing was launched: an empty span
            HPRQP_HIP(hipEventRecord(ev[used], s));
            HPRQP_HIP(hipEventRecord(ev[used + 1], s));
        }
        t = LaunchTiming{};
        used += 2;
    }
    void flush() {
        if (!used) return;
        HPRQP_HIP(hipEventSynchronize(ev[used - 1]));
        for (int k = 0; k < used; k += 2) {
            float ms = 0.f;
            HPRQP_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            total_ms += ms;
            ++count;
        }
        used = 0;
    }
};

// ---------------------------------------------------------------------
// scaled FP64 problem (output of scaling!, src/utils.jl:350-804)
// ---------------------------------------------------------------------
struct Scaled {
    int32_t n = 0, m = 0;
    int qkind = QOP_CSR;
    hprqp_qop *qop = nullptr;
    bool q_empty = true;
    DevCsr Q, A, AT;
    DevVec<double> c, l, u, AL, AU, row_norm, col_norm, diag_Q;
    DevVec<double> inv_rn, inv_cn;  // 1 / row_norm, 1 / col_norm (value-table products)
    std::vector<double> diag_Q_host;
    double b_scale = 1.0, c_scale = 1.0, norm_b = 0, norm_c = 0, norm_b_org = 0, norm_c_org = 0;
    bool Q_is_diag = false;
    double q_max_val = 0.0;
    // one bound value for every variable (0 or +-inf, which scaling keeps),
    // every constraint row an equality
    bool l_uniform = false, u_uniform = false, all_equal_rows = false;
    double l_value = 0.0, u_value = 0.0;
    // QAP operator with exactly the relaxation's doubly stochastic
    // constraints (rows: X 1 = 1, then X^T 1 = 1; every value 1)
    bool qap_ds = false;
};

// A is build_from_ABST's constraint matrix [kron(1', I); kron(I, 1')] for
// the operator's nq (src/utils.jl:1160-1163), values exactly 1
static bool qap_doubly_stochastic(const hprqp_model *md) {
    const int32_t nq = md->Q->nq;
    const HostCsr &A = md->A;
    if (nq <= 0 || A.nrow != 2 * nq || (int64_t)A.ncol != (int64_t)nq * nq ||
        (int64_t)A.v.size() != 2 * (int64_t)nq * nq || (int64_t)A.rp.size() != 2 * (int64_t)nq + 1)
        return false;
    for (int32_t r = 0; r < 2 * nq; ++r) {
        if (A.rp[r] != (int64_t)r * nq) return false;
        for (int32_t t = 0; t < nq; ++t) {
            const int64_t k = (int64_t)r * nq + t;
            const int64_t col = r < nq ? r + (int64_t)nq * t : (int64_t)(r - nq) * nq + t;
            if (A.ci[k] != col || A.v[k] != 1.0) return false;
        }
    }
    return A.rp[2 * nq] == 2 * (int64_t)nq * nq;
}

static void upload_vec(hipStream_t s, DevVec<double> &d, const std::vector<double> &h) {
    d.alloc((int64_t)h.size());
    if (!h.empty())
        HPRQP_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s));
}

template <class T>
static void copy_vec(hipStream_t s, DevVec<T> &d, const DevVec<T> &src) {
    d.alloc(src.n);
    if (src.n > 0) HPRQP_HIP(hipMemcpyAsync(d.p, src.p, sizeof(T) * src.n, hipMemcpyDeviceToDevice, s));
}

// a device-to-device copy of a CSR matrix (pattern, FP64 values, shape)
static void copy_csr(hipStream_t s, DevCsr &d, const DevCsr &src) {
    d.nrow = src.nrow;
    d.ncol = src.ncol;
    d.nnz = src.nnz;
    copy_vec(s, d.rp, src.rp);
    copy_vec(s, d.ci, src.ci);
    copy_vec(s, d.v, src.v);
    d.v32.release();
    d.lanes = src.lanes;
    d.cseg.release();  // a new pattern: chunk segments are rebuilt on use
    d.cpart.release();
    d.nchunk = 0;
}

// the model's raw data on the device, once per model
static void ensure_raw(hipStream_t s, const hprqp_model *md) {
    hprqp_model::Raw &r = md->raw;
    if (r.ready) return;
    upload_csr(s, md->A, r.A);
    upload_csr(s, host_csr_transpose(md->A), r.AT);
    if (md->Q->kind == QOP_CSR) upload_csr(s, md->Qsym, r.Q);
    upload_vec(s, r.c, md->c);
    upload_vec(s, r.l, md->l);
    upload_vec(s, r.u, md->u);
    upload_vec(s, r.AL, md->AL);
    upload_vec(s, r.AU, md->AU);
    upload_vec(s, r.a_dict, md->a_dict);
    upload_vec(s, r.q_dict, md->q_dict);
    auto up8 = [&](DevVec<uint8_t> &d, const std::vector<uint8_t> &h) {
        d.alloc((int64_t)h.size());
        if (!h.empty()) HPRQP_HIP(hipMemcpyAsync(d.p, h.data(), h.size(), hipMemcpyHostToDevice, s));
    };
    up8(r.a_code, md->a_code);
    up8(r.at_code, md->at_code);
    up8(r.q_code, md->q_code);
    HPRQP_HIP(hipStreamSynchronize(s));
    r.ready = true;
}

static void fill(hipStream_t s, double *p, int64_t n, double v) {
    each(s, n, [=] __device__(int64_t i) { p[i] = v; });
}

// bound statistic max(|AL_nInf|, |AU_nInf|): inf-norm (k=0) and 2-norm^2 (k=1)
struct BoundNorm {
    const double *AL, *AU;
    __device__ void operator()(int64_t i, double *s, double *mx) const {
        const double a = AL[i] == -INFINITY ? 0.0 : AL[i];
        const double b = AU[i] == INFINITY ? 0.0 : AU[i];
        const double t = dmax(fabs(a), fabs(b));
        s[0] += t * t;
        mx[0] = dmax(mx[0], t);
    }
};
struct VecNorm {
    const double *v;
    __device__ void operator()(int64_t i, double *s, double *mx) const {
        s[0] += v[i] * v[i];
        mx[0] = dmax(mx[0], fabs(v[i]));
    }
};

static void build_scaled_data(hprqp_ctx *ctx, const hprqp_model *md, const hprqp_params *p, Reducer &red,
                              Scaled &S) {
    hipStream_t s = ctx->stream;
    const int32_t n = md->n, m = md->m;
    S.n = n;
    S.m = m;
    S.qop = md->Q;
    S.qkind = md->Q->kind;
    auto uniform = [&](const std::vector<double> &v, double &val) {
        if (v.empty()) return false;
        const double a = v[0];
        if (!(a == 0.0 || std::isinf(a))) return false;
        for (double b : v)
            if (b != a) return false;
        val = a;
        return true;
    };
    S.l_uniform = uniform(md->l, S.l_value);
    S.u_uniform = uniform(md->u, S.u_value);
    S.all_equal_rows = m > 0;
    for (int32_t r = 0; r < m && S.all_equal_rows; ++r) S.all_equal_rows = md->AL[r] == md->AU[r];
    // raw data, from the model's device copies
    ensure_raw(s, md);
    const hprqp_model::Raw &raw = md->raw;
    copy_csr(s, S.A, raw.A);
    copy_csr(s, S.AT, raw.AT);
    if (S.qkind == QOP_CSR) {
        copy_csr(s, S.Q, raw.Q);
        S.q_empty = md->Qsym.v.empty();
        S.Q.lanes = md->Q->q.lanes;
    } else {
        S.q_empty = false;
    }
    copy_vec(s, S.c, raw.c);
    copy_vec(s, S.l, raw.l);
    copy_vec(s, S.u, raw.u);
    copy_vec(s, S.AL, raw.AL);
    copy_vec(s, S.AU, raw.AU);
    S.row_norm.alloc(m);
    S.col_norm.alloc(n);
    fill(s, S.row_norm.p, m, 1.0);
    fill(s, S.col_norm.p, n, 1.0);
    // original norms (src/utils.jl:404-408)
    red.enqueue<1, 1>(m, BoundNorm{S.AL.p, S.AU.p}, 0);
    red.enqueue<1, 1>(n, VecNorm{S.c.p}, 2);
    const double *h = red.fetch(4);
    S.norm_b_org = m > 0 ? h[1] : 0.0;
    S.norm_c_org = h[3];
    S.qap_ds = S.qkind == QOP_QAP && qap_doubly_stochastic(md);
    if (S.qkind != QOP_CSR) {
        S.norm_b = m > 0 ? std::sqrt(h[0]) : 0.0;
        S.norm_c = std::sqrt(h[2]);
        return;
    }
    DevVec<double> tr(m), tc(n);
    // lanes per row of the scaling kernels from the mean row length
    auto lanes_of = [](int64_t nnz, int64_t rows) { const int64_t mean = rows > 0 ? nnz / rows : 0;
                                                    return mean >= 32 ? 16 : mean >= 8 ? 4 : 1; };
    auto launch_rows = [&](int L, int64_t rows, auto &&f1, auto &&f4, auto &&f16) {
        const dim3 grid((unsigned)(((int64_t)rows * L + kBlock - 1) / kBlock)), blk(kBlock);
        if (L == 16)
            f16(grid, blk);
        else if (L == 4)
            f4(grid, blk);
        else
            f1(grid, blk);
        HPRQP_HIP(hipGetLastError());
    };
    const int Lq = lanes_of(S.Q.nnz, n), La = lanes_of(S.A.nnz, m), Lat = lanes_of(S.AT.nnz, n);
    const int Lc = std::max(Lq, Lat);
    const int nruiz = p->use_Ruiz_scaling ? 10 : 0;
    const int total = nruiz + (p->use_Pock_Chambolle_scaling ? 1 : 0);
    for (int pass = 0; pass < total; ++pass) {
        const int ruiz = pass < nruiz;
        {
            const int32_t *atrp = m > 0 ? S.AT.rp.p : nullptr, *qrp = S.Q.nnz > 0 ? S.Q.rp.p : nullptr;
            const double *atv = S.AT.v.p, *qv = S.Q.v.p;
            double *o = tc.p;
            launch_rows(
                Lc, n, [&](dim3 g, dim3 b) { k_col_stat<1><<<g, b, 0, s>>>(n, atrp, atv, qrp, qv, ruiz, o); },
                [&](dim3 g, dim3 b) { k_col_stat<4><<<g, b, 0, s>>>(n, atrp, atv, qrp, qv, ruiz, o); },
                [&](dim3 g, dim3 b) { k_col_stat<16><<<g, b, 0, s>>>(n, atrp, atv, qrp, qv, ruiz, o); });
        }
        if (m > 0) {
            const int32_t *arp = S.A.rp.p;
            const double *av = S.A.v.p;
            double *o = tr.p;
            launch_rows(
                La, m, [&](dim3 g, dim3 b) { k_row_stat<1><<<g, b, 0, s>>>(m, arp, av, ruiz, o); },
                [&](dim3 g, dim3 b) { k_row_stat<4><<<g, b, 0, s>>>(m, arp, av, ruiz, o); },
                [&](dim3 g, dim3 b) { k_row_stat<16><<<g, b, 0, s>>>(m, arp, av, ruiz, o); });
        }
        {
            double *cn = S.col_norm.p, *c = S.c.p, *l = S.l.p, *u = S.u.p;
            const double *t = tc.p;
            each(s, n, [=] __device__(int64_t j) {
                cn[j] *= t[j];
                c[j] /= t[j];
                l[j] *= t[j];
                u[j] *= t[j];
            });
        }
        if (m > 0) {
            double *rn = S.row_norm.p, *al = S.AL.p, *au = S.AU.p;
            const double *t = tr.p;
            each(s, m, [=] __device__(int64_t i) {
                rn[i] *= t[i];
                al[i] /= t[i];
                au[i] /= t[i];
            });
        }
        auto scale = [&](int L, int32_t rows, const DevCsr &a, const double *r, const double *c) {
            const int32_t *rp = a.rp.p, *ci = a.ci.p;
            double *v = a.v.p;
            launch_rows(
                L, rows, [&](dim3 g, dim3 b) { k_scale_csr<1><<<g, b, 0, s>>>(rows, rp, ci, v, r, c); },
                [&](dim3 g, dim3 b) { k_scale_csr<4><<<g, b, 0, s>>>(rows, rp, ci, v, r, c); },
                [&](dim3 g, dim3 b) { k_scale_csr<16><<<g, b, 0, s>>>(rows, rp, ci, v, r, c); });
        };
        if (S.Q.nnz > 0) scale(Lq, n, S.Q, tc.p, tc.p);
        if (m > 0 && S.A.nnz > 0) {
            scale(La, m, S.A, tr.p, tc.p);
            scale(Lat, n, S.AT, tc.p, tr.p);
        }
    }
    // Q = (Q + Q')/2 (src/utils.jl:772-777) over the symmetric pattern
    if (S.Q.nnz > 0) {
        DevVec<int32_t> tpos(S.Q.nnz);
        HPRQP_HIP(hipMemcpyAsync(tpos.p, md->q_tpos.data(), sizeof(int32_t) * S.Q.nnz,
                                 hipMemcpyHostToDevice, s));
        DevVec<double> v2(S.Q.nnz);
        const double *v = S.Q.v.p;
        const int32_t *tp = tpos.p;
        double *o = v2.p;
        each(s, S.Q.nnz, [=] __device__(int64_t k) { o[k] = (v[k] + v[tp[k]]) / 2.0; });
        HPRQP_HIP(hipStreamSynchronize(s));
        std::swap(S.Q.v, v2);
    }
    red.enqueue<1, 1>(m, BoundNorm{S.AL.p, S.AU.p}, 0);
    red.enqueue<1, 1>(n, VecNorm{S.c.p}, 2);
    h = red.fetch(4);
    S.norm_b = m > 0 ? std::sqrt(h[0]) : 0.0;
    S.norm_c = std::sqrt(h[2]);
}

// the value-table form of Q, A, A^T (model value tables, the scaling as
// row and column factors 1 / row_norm, 1 / col_norm); used where the
// autotuner finds it faster (apply_lanes)
static void attach_dicts(hipStream_t s, const hprqp_model *md, Scaled &S) {
    const int32_t n = S.n, m = S.m;
    S.inv_rn.alloc(m);
    S.inv_cn.alloc(n);
    {
        double *ir = S.inv_rn.p, *ic = S.inv_cn.p;
        const double *rn = S.row_norm.p, *cn = S.col_norm.p;
        each(s, m, [=] __device__(int64_t i) { ir[i] = 1.0 / rn[i]; });
        each(s, n, [=] __device__(int64_t j) { ic[j] = 1.0 / cn[j]; });
    }
    // the codes and tables from the model's device copies
    const hprqp_model::Raw &raw = md->raw;
    auto attach = [&](DevCsr &a, const DevVec<double> &dict, const DevVec<uint8_t> &code, const double *rsc,
                      const double *csc) {
        if (dict.n == 0 || code.n != a.nnz || a.nnz == 0) return;
        copy_vec(s, a.code, code);
        copy_vec(s, a.dict, dict);
        a.ndict = (int)dict.n;
        a.rsc = rsc;
        a.csc = csc;
    };
    if (m > 0) {
        attach(S.A, raw.a_dict, raw.a_code, S.inv_rn.p, S.inv_cn.p);
        attach(S.AT, raw.a_dict, raw.at_code, S.inv_cn.p, S.inv_rn.p);
    }
    if (S.qkind == QOP_CSR && !S.q_empty) attach(S.Q, raw.q_dict, raw.q_code, S.inv_cn.p, S.inv_cn.p);
    HPRQP_HIP(hipStreamSynchronize(s));
}

static void build_scaled(hprqp_ctx *ctx, const hprqp_model *md, const hprqp_params *p, Reducer &red, Scaled &S) {
    build_scaled_data(ctx, md, p, red, S);
    attach_dicts(ctx->stream, md, S);
}

// check_Q_diagonal after scaling (src/utils.jl:827-855)
struct DiagCheck {
    const int32_t *rp, *ci;
    const double *v;
    double *diag;
    __device__ void operator()(int64_t i, double *s, double *mx) const {
        double sum = 0.0, d = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            sum += fabs(v[k]);
            if (ci[k] == i) d = v[k];
            mx[1] = dmax(mx[1], v[k]);
        }
        diag[i] = d;
        const double ad = fabs(d);
        const bool ok = sum == ad || (isfinite(sum) && isfinite(ad) &&
                                      fabs(sum - ad) <= 1.4901161193847656e-8 * dmax(sum, ad));
        if (!ok) mx[0] = 1.0;
    }
};

static void check_Q_diagonal(Reducer &red, Scaled &S) {
    S.diag_Q.alloc(S.n);
    fill(red.s, S.diag_Q.p, S.n, 0.0);
    if (S.qkind != QOP_CSR) {
        S.Q_is_diag = false;
        S.diag_Q_host.assign(S.n, 0.0);
        return;
    }
    red.enqueue<0, 2>(S.n, DiagCheck{S.Q.rp.p, S.Q.ci.p, S.Q.v.p, S.diag_Q.p}, 0);
    const double *h = red.fetch(2);
    S.Q_is_diag = h[0] == 0.0;
    S.q_max_val = h[1];
    S.diag_Q_host.resize(S.n);
    HPRQP_HIP(hipMemcpy(S.diag_Q_host.data(), S.diag_Q.p, sizeof(double) * S.n, hipMemcpyDeviceToHost));
}

// deterministic N(0,1) stream for the power-method start vector; the same
// counter-based generator is used by the CPU oracle so both start alike.
static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static void randn_stream(uint64_t seed, int64_t count, double *out) {
    const double two_pi = 6.283185307179586476925286766559;
    const uint64_t ks = mix64(seed + 0x9E3779B97F4A7C15ULL);
    for (int64_t i = 0; i < count; ++i) {
        const uint64_t k0 = 2 * (uint64_t)i, k1 = k0 + 1;
        const double a = (double)(mix64(ks ^ (k0 * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL)) >> 11) *
                         (1.0 / 9007199254740992.0);
        const double b = (double)(mix64(ks ^ (k1 * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL)) >> 11) *
                         (1.0 / 9007199254740992.0);
        out[i] = std::sqrt(-2.0 * std::log(1.0 - a)) * std::cos(two_pi * b);
    }
}

// residuals of one evaluation (compute_residuals!, src/algorithm.jl:37-133)
struct Resid {
    double err_Rp = 0, err_Rd = 0, kkt = 0, primal_obj = 0, rel_gap = 0, dual_obj = 0;
};

// ---------------------------------------------------------------------
// Solver<VT>
// ---------------------------------------------------------------------
template <class VT, class MT = VT>
struct Solver {
    hprqp_ctx *ctx;
    hipStream_t s;
    Scaled &S;
    Reducer &red;
    const hprqp_params *p;
    int32_t n, m;
    // bulk vectors
    DevVec<VT> w, w_bar, dw, x, x_hat, x_bar, dx, y, y_bar, dy, z_bar;
    DevVec<VT> ATy, ATy_bar, ATdy, QATdy, s_, Qw, Qw_bar, dQw, tempv;
    DevVec<VT> fact1, fact2, fact, fact_M;
    DevVec<VT> qtemp;  // operator temp (LASSO rows / QAP n)
    // anchors (FP64)
    DevVec<double> last_x, last_y, last_Qw, last_w, last_ATy;
    // FP64 residual temporaries
    DevVec<double> Qx_d, ATdy_d, Ax_d, xd_tmp, qtemp_d, lr_t;
    DevVec<double> sig_buf;  // the diagonal-Q sigma search: bracket and partials
    // FP32 iterates: Q dw and Q A^T dy of the sigma update and the M-norm in
    // FP64 (<dw, Q dw> cancels below FP32 resolution near convergence when
    // Q has a large null space, as LASSO's A^T A does)
    DevVec<double> sg_Qw, sg_QATdy;
    void ensure_sig64() {
        if (sg_Qw.n < n) {
            sg_Qw.alloc(n);
            sg_QATdy.alloc(n);
        }
    }
    // data in VT (aliases of the FP64 arrays when VT == double)
    DevVec<MT> c_own, l_own, u_own, AL_own, AU_own;
    const MT *cV, *lV, *uV, *ALV, *AUV;
    // what the half-step epilogues read: null when every variable has the
    // same bound (then lS / uS) or every row is an equality (AU = AL)
    const MT *lE = nullptr, *uE = nullptr, *AUE = nullptr;
    double lS = 0.0, uS = 0.0;
    double sigma = 1.0, lambda_max_A = 0.0, lambda_max_Q = 0.0, lambda = 0.0;
    bool noC = false;
    KernelTimer *kt = nullptr;  // times the first fused kernel of partial steps
    // QAP with doubly stochastic constraints (qds): one apply per iteration.
    // qv_ok: Qw, QATy, last_Qw, last_QATy are the products of w, ATy,
    // last_w, last_ATy; pend: the last step's second w update is still to
    // run (it runs fused with the next z/x/w_bar update, or in flush())
    // qds_small (nq < 256): the y update is pending too and runs in the
    // fused kernel from pass 2's tile sums (tsum); y_alt is the other buffer
    // of the y pair
    bool qds = false, qds_small = false, qv_ok = false, pend = false, pend_full = false, flushed_full = false;
    double pend_h1 = 0.0, pend_h2 = 0.0, pend_fs = 0.0, pend_fact1 = 1.0;
    DevVec<VT> QATy, Qwb_new, Qg, y_alt, qdv;
    DevVec<double> last_QATy, qvec, qBt, qTt, tsum;
    bool timing_now = false;
    void kbegin() {
        if (timing_now) kt->begin();
    }
    void kend() {
        if (timing_now) kt->end();
    }

    Solver(hprqp_ctx *c, Scaled &sc, Reducer &r, const hprqp_params *pp)
        : ctx(c), s(c->stream), S(sc), red(r), p(pp), n(sc.n), m(sc.m) {}

    void alloc() {
        for (DevVec<VT> *v : {&w, &w_bar, &dw, &x, &x_hat, &x_bar, &dx, &z_bar, &ATy, &ATy_bar, &ATdy,
                              &QATdy, &Qw, &Qw_bar, &dQw, &tempv, &fact1, &fact2, &fact, &fact_M}) {
            v->alloc(n);
            HPRQP_HIP(hipMemsetAsync(v->p, 0, sizeof(VT) * n, s));
        }
        for (DevVec<VT> *v : {&y, &y_bar, &dy, &s_}) {
            v->alloc(m);
            if (m) HPRQP_HIP(hipMemsetAsync(v->p, 0, sizeof(VT) * m, s));
        }
        for (DevVec<double> *v : {&last_x, &last_Qw, &last_w, &last_ATy, &Qx_d, &ATdy_d, &xd_tmp}) {
            v->alloc(n);
            HPRQP_HIP(hipMemsetAsync(v->p, 0, sizeof(double) * n, s));
        }
        for (DevVec<double> *v : {&last_y, &Ax_d}) {
            v->alloc(m);
            if (m) HPRQP_HIP(hipMemsetAsync(v->p, 0, sizeof(double) * m, s));
        }
        if (S.qkind == QOP_LASSO) {
            qtemp.alloc(S.qop->la.nrow);
            qtemp_d.alloc(S.qop->la.nrow);
        } else if (S.qkind == QOP_QAP) {
            qtemp.alloc(n);
            qtemp_d.alloc(n);
        } else if (S.qkind == QOP_LOWRANK) {
            lr_t.alloc(std::max(1, S.qop->lr_k));
            if constexpr (!std::is_same<MT, double>::value) {
                hprqp_qop *q = S.qop;
                const int64_t nk = (int64_t)n * q->lr_k;
                if (q->lr_U32.n != nk) {
                    q->lr_U32.alloc(nk);
                    q->lr_d32.alloc(n);
                    const double *u = q->lr_U.p, *d = q->lr_d.p;
                    float *u32 = q->lr_U32.p, *d32 = q->lr_d32.p;
                    each(s, nk, [=] __device__(int64_t i) { u32[i] = (float)u[i]; });
                    each(s, n, [=] __device__(int64_t i) { d32[i] = (float)d[i]; });
                }
            }
        }
        if constexpr (std::is_same<MT, double>::value) {
            cV = S.c.p;
            lV = S.l.p;
            uV = S.u.p;
            ALV = S.AL.p;
            AUV = S.AU.p;
        } else {
            auto cvt = [&](DevVec<MT> &dst, const DevVec<double> &src) {
                dst.alloc(src.n);
                const double *a = src.p;
                MT *b = dst.p;
                each(s, src.n, [=] __device__(int64_t i) { b[i] = (MT)a[i]; });
            };
            cvt(c_own, S.c);
            cvt(l_own, S.l);
            cvt(u_own, S.u);
            cvt(AL_own, S.AL);
            cvt(AU_own, S.AU);
            cV = c_own.p;
            lV = l_own.p;
            uV = u_own.p;
            ALV = AL_own.p;
            AUV = AU_own.p;
            if (S.qkind == QOP_CSR) ensure_v32(S.Q);
            ensure_v32(S.A);
            ensure_v32(S.AT);
            if (S.qkind == QOP_LASSO) {
                ensure_v32(S.qop->la);
                ensure_v32(S.qop->lat);
            }
        }
        set_epilogue_bounds();
        if (S.qap_ds && !p->qap_recompute_Qw) alloc_qds();
    }

    // the carried products and the constants of Q A^T y (hprqp_qapds.hpp)
    void alloc_qds() {
        hprqp_qop *q = S.qop;
        const int32_t nq = q->nq;
        const int64_t nn = (int64_t)nq * nq;
        qds = true;
        for (DevVec<VT> *v : {&QATy, &Qwb_new, &Qg, &qdv}) {
            v->alloc(nn);
            HPRQP_HIP(hipMemsetAsync(v->p, 0, sizeof(VT) * nn, s));
        }
        last_QATy.alloc(nn);
        HPRQP_HIP(hipMemsetAsync(last_QATy.p, 0, sizeof(double) * nn, s));
        qvec.alloc(8 * (int64_t)nq);
        auto transpose = [&](DevVec<double> &dst, const double *src) {
            dst.alloc(nn);
            double *d = dst.p;
            each(s, nn, [=] __device__(int64_t o) { d[o] = src[(o / nq) + (o % nq) * (int64_t)nq]; });
        };
        transpose(qBt, q->qb.p);
        if (!q->qt_diag) transpose(qTt, q->qt.p);
        double *c = qvec.p + 4 * (int64_t)nq;  // rA, cB, rS, cT
        qap_ds_gemv<double>(s, nq, q->qa.p, qBt.p, q->qs_diag ? nullptr : q->qs.p, q->qt_diag ? nullptr : qTt.p,
                            nullptr, c, c + nq, c + 2 * nq, c + 3 * nq);
        qds_small = nq < kQapBlockedMin;
        if (qds_small) {
            y_alt.alloc(m);
            tsum.alloc(2 * (int64_t)((nq + 15) / 16) * nq);
        }
        qv_ok = false;
        pend = false;
    }

    // the products of the iterates by applies (Q w in the mode's precision,
    // the anchors' in FP64)
    void qds_init() {
        q_apply(w.p, Qw.p);
        q_apply(ATy.p, QATy.p);
        qop_fp64(last_w.p, last_Qw.p);
        qop_fp64(last_ATy.p, last_QATy.p);
        qv_ok = true;
    }

    QapDsStep<VT, MT> qds_args() {
        hprqp_qop *q = S.qop;
        const int32_t nq = q->nq;
        const double *v = qvec.p, *c = qvec.p + 4 * (int64_t)nq;
        QapDsStep<VT, MT> e{};
        e.n = nq;
        e.yb = y_bar.p;
        e.P = v;
        e.R = v + nq;
        e.Sa = q->qs_diag ? nullptr : v + 2 * nq;
        e.Tb = q->qt_diag ? nullptr : v + 3 * nq;
        e.rA = c;
        e.cB = c + nq;
        e.rS = c + 2 * nq;
        e.cT = c + 3 * nq;
        e.sd = q->qsd.p;
        e.td = q->qtd.p;
        e.gA = q->qa.p;
        e.gBt = qBt.p;
        e.gS = q->qs_diag ? nullptr : q->qs.p;
        e.gTt = q->qt_diag ? nullptr : qTt.p;
        e.last_w = last_w.p;
        e.last_ATy = last_ATy.p;
        e.last_Qw = last_Qw.p;
        e.last_QATy = last_QATy.p;
        e.w = w.p;
        e.w_bar = w_bar.p;
        e.ATy = ATy.p;
        e.Qw = Qw.p;
        e.QATy = QATy.p;
        e.Qw_bar = Qw_bar.p;
        e.dw = dw.p;
        e.ATdy = ATdy.p;
        e.ATy_bar = ATy_bar.p;
        e.Qwb_new = Qwb_new.p;
        e.Qg = Qg.p;
        e.fs = pend_fs;
        e.wh1 = pend_h1;
        e.wh2 = pend_h2;
        e.c = cV;
        e.l = lE;
        e.u = uE;
        e.lS = lS;
        e.uS = uS;
        e.last_x = last_x.p;
        e.x = x.p;
        e.x_hat = x_hat.p;
        e.dx = dx.p;
        e.x_bar = x_bar.p;
        e.z_bar = z_bar.p;
        e.d = qdv.p;
        e.sigma = sigma;
        e.f2s = 1.0 / (1.0 + sigma * lambda_max_Q);
        e.f1s = 1.0 - e.f2s;
        return e;
    }

    // the pending y update (EpiY's scalars of its step)
    QapDsY<VT, MT> qds_yargs() {
        QapDsY<VT, MT> q{};
        q.tsum = tsum.p;
        q.AL = ALV;
        q.AU = AUE;
        q.last_y = last_y.p;
        q.y = y.p;
        q.y_next = y_alt.p;
        q.y_bar = y_bar.p;
        q.s = s_.p;
        q.dy = dy.p;
        q.fact1 = pend_fact1;
        q.fact2 = 1.0 / pend_fact1;
        q.h1 = pend_h1;
        q.h2 = pend_h2;
        return q;
    }

    void qds_gemv() {
        hprqp_qop *q = S.qop;
        const int32_t nq = q->nq;
        double *v = qvec.p;
        qap_ds_gemv<VT>(s, nq, q->qa.p, qBt.p, q->qs_diag ? nullptr : q->qs.p, q->qt_diag ? nullptr : qTt.p,
                        y_bar.p, v, v + nq, v + 2 * nq, v + 3 * nq);
    }

    // run the pending second w update on its own
    void flush() {
        if (!pend) return;
        QapDsStep<VT, MT> e = qds_args();
        if (qds_small) {
            if (pend_full)
                qap_ds_tile<VT, MT, 2, 0>(s, e, qds_yargs());
            else
                qap_ds_tile<VT, MT, 1, 0>(s, e, qds_yargs());
            std::swap(y, y_alt);
        } else {
            qds_gemv();
            if (pend_full)
                qap_ds_step<VT, MT, 2, 0>(s, e);
            else
                qap_ds_step<VT, MT, 1, 0>(s, e);
        }
        flushed_full = pend_full;
        pend = false;
    }

    // the iterates were set from outside the iteration: products by applies
    // at the next step
    void invalidate_products() {
        flush();
        qv_ok = false;
    }

    // one HPR step of the QAP relaxation: [second w update of the pending
    // step +] z/x/w_bar update on the carried Q w, then Q w_bar with the
    // tempv epilogue (the only apply), then the y update; the second w
    // update is left pending
    template <bool FULL>
    void qds_update(double h1, double h2) {
        if (!qv_ok) {
            flush();
            qds_init();
        }
        if (pend && pend_full) flush();
        QapDsStep<VT, MT> e = qds_args();
        e.h1 = h1;
        e.h2 = h2;
        if (qds_small) {
            if (pend) {
                flushed_full = false;
                qap_ds_tile<VT, MT, 1, FULL ? 2 : 1>(s, e, qds_yargs());
                std::swap(y, y_alt);
            } else {
                qap_ds_tile<VT, MT, 0, FULL ? 2 : 1>(s, e, qds_yargs());
            }
        } else if (pend) {
            qds_gemv();
            flushed_full = false;
            qap_ds_step<VT, MT, 1, FULL ? 2 : 1>(s, e);
        } else {
            qap_ds_step<VT, MT, 0, FULL ? 2 : 1>(s, e);
        }
        pend = false;
        // the apply, of d = w - x_hat (its two MFMA passes are the timed
        // kernels of the step)
        kbegin();
        if (qds_small) {
            // pass 2 leaves tempv's tile sums; the y update waits for the
            // next fused kernel
            EpiTempvDiff<VT, true> et{x_hat.p, Qw.p, Qw_bar.p, nullptr, tsum.p, sigma, e.f2s};
            q_fused(qdv.p, et, Qw_bar.p);
            kend();
        } else {
            EpiTempvDiff<VT, false> et{x_hat.p, Qw.p, Qw_bar.p, tempv.p, nullptr, sigma, e.f2s};
            q_fused(qdv.p, et, Qw_bar.p);
            kend();
            step_y<FULL>(tempv.p, h1, h2);
        }
        pend = true;
        pend_full = FULL;
        pend_h1 = h1;
        pend_h2 = h2;
        pend_fs = sigma / (1.0 + sigma * lambda_max_Q);
        pend_fact1 = lambda_max_A * sigma;
    }

    void set_epilogue_bounds() {
        lE = S.l_uniform ? nullptr : lV;
        uE = S.u_uniform ? nullptr : uV;
        AUE = S.all_equal_rows ? nullptr : AUV;
        lS = S.l_value;
        uS = S.u_value;
    }

    // ---------------- Q applies -------------------------------------
    // Qx = Q x in VT precision
    void q_apply(const VT *xin, VT *out) {
        switch (S.qkind) {
            case QOP_CSR:
                if (S.q_empty) {
                    HPRQP_HIP(hipMemsetAsync(out, 0, sizeof(VT) * n, s));
                } else {
                    csr_apply<VT, MT>(s, S.Q, xin, EpiStore<VT>{out});
                }
                break;
            case QOP_LASSO:
                csr_apply<VT, MT>(s, S.qop->la, xin, EpiStore<VT>{qtemp.p});
                csr_apply<VT, MT>(s, S.qop->lat, (const VT *)qtemp.p, EpiStore<VT>{out});
                break;
            case QOP_QAP:
                // FP64 solves: FP64 MFMA; reduced precision: split-BF16 x3 MFMA
                if constexpr (std::is_same<MT, double>::value)
                    qap_apply_f64(s, S.qop, xin, out);
                else
                    qap_apply_bf16x3<VT, VT>(s, S.qop, xin, out);
                break;
            case QOP_LOWRANK:
                if constexpr (std::is_same<MT, double>::value)
                    lowrank_apply<double, VT, VT>(s, n, S.qop->lr_k, S.qop->lr_U.p, S.qop->lr_d.p, xin,
                                                  lr_t.p, out);
                else
                    lowrank_apply<float, VT, VT>(s, n, S.qop->lr_k, S.qop->lr_U32.p, S.qop->lr_d32.p, xin,
                                                 lr_t.p, out);
                break;
            default: {
                if constexpr (std::is_same<VT, double>::value) {
                    int rc = S.qop->vt.apply(S.qop->user, xin, out, (void *)s);
                    HPRQP_REQUIRE(rc == 0, HPRQP_E_ARG, "user Q operator apply failed");
                } else {
                    double *xd = xd_tmp.p, *od = Qx_d.p;
                    each(s, n, [=] __device__(int64_t i) { xd[i] = (double)xin[i]; });
                    int rc = S.qop->vt.apply(S.qop->user, xd, od, (void *)s);
                    HPRQP_REQUIRE(rc == 0, HPRQP_E_ARG, "user Q operator apply failed");
                    each(s, n, [=] __device__(int64_t i) { out[i] = (VT)od[i]; });
                }
            }
        }
    }

    // Q x with a fused per-row epilogue; qbuf receives the product when Q
    // is a dense/user operator (then the epilogue runs elementwise).
    template <class Epi>
    void q_fused(const VT *xin, const Epi &epi, VT *qbuf) {
        if (S.qkind == QOP_CSR && !S.q_empty) {
            csr_apply<VT, MT>(s, S.Q, xin, epi);
        } else if (S.qkind == QOP_LASSO) {
            csr_apply<VT, MT>(s, S.qop->la, xin, EpiStore<VT>{qtemp.p});
            csr_apply<VT, MT>(s, S.qop->lat, (const VT *)qtemp.p, epi);
        } else if (S.qkind == QOP_QAP) {
            // the half-step epilogue runs in the second MFMA pass
            if constexpr (std::is_same<MT, double>::value)
                qap_launch_f64(s, S.qop, xin, epi);
            else
                qap_launch_bf16x3<VT>(s, S.qop, xin, epi);
        } else {
            q_apply(xin, qbuf);
            vec_apply<VT, VT>(s, n, (const VT *)qbuf, epi);
        }
    }

    // Qx = Q x in FP64 from a VT input (residuals)
    void q_apply_d(const VT *xin, double *out) {
        switch (S.qkind) {
            case QOP_CSR:
                if (S.q_empty)
                    HPRQP_HIP(hipMemsetAsync(out, 0, sizeof(double) * n, s));
                else
                    csr_apply<double, double>(s, S.Q, xin, EpiStore<double>{out});
                break;
            case QOP_LASSO:
                csr_apply<double, double>(s, S.qop->la, xin, EpiStore<double>{qtemp_d.p});
                csr_apply<double, double>(s, S.qop->lat, (const double *)qtemp_d.p, EpiStore<double>{out});
                break;
            default: {
                const double *xd;
                if constexpr (std::is_same<VT, double>::value) {
                    xd = xin;
                } else {
                    double *t = xd_tmp.p;
                    each(s, n, [=] __device__(int64_t i) { t[i] = (double)xin[i]; });
                    xd = t;
                }
                if (S.qkind == QOP_QAP) {
                    qap_apply_f64(s, S.qop, xd, out);
                } else if (S.qkind == QOP_LOWRANK) {
                    lowrank_apply<double, double, double>(s, n, S.qop->lr_k, S.qop->lr_U.p, S.qop->lr_d.p, xd,
                                                          lr_t.p, out);
                } else {
                    int rc = S.qop->vt.apply(S.qop->user, xd, out, (void *)s);
                    HPRQP_REQUIRE(rc == 0, HPRQP_E_ARG, "user Q operator apply failed");
                }
            }
        }
    }

    // ---------------- power method (src/utils.jl:1335-1461) ---------
    struct SumSq {
        const double *v;
        __device__ void operator()(int64_t i, double *s, double *) const { s[0] += v[i] * v[i]; }
    };

    struct Dot2Sq {
        const double *a, *b;
        __device__ void operator()(int64_t i, double *s, double *) const {
            const double ai = a[i];
            s[0] += ai * b[i];
            s[1] += ai * ai;
        }
    };

    // The power method's row product over a: the matrix's own row-product
    // shape (CSR-stream rows per block, or lanes per row; a block per row and
    // the column-chunked shape take 64 lanes; the FP64 value-table form where
    // the autotuner chose it), a block per row group
    struct PowShape {
        int L = 64, rb = 0;
        bool dict = false;
        int32_t ngroups = 0;
    };
    static PowShape pow_shape(const DevCsr &a) {
        PowShape ps;
        if (a.dict_rb > 0 && a.nnz > 0) {
            ps.L = 0;
            ps.rb = a.dict_rb;
            ps.dict = true;
            ps.ngroups = (int32_t)(((int64_t)a.nrow + ps.rb - 1) / ps.rb);
        } else if (a.lanes < 0) {
            ps.L = 0;
            ps.rb = -a.lanes;
            ps.ngroups = (int32_t)(((int64_t)a.nrow + ps.rb - 1) / ps.rb);
        } else {
            const int L = a.lanes;
            ps.L = L >= 1 && L <= 64 && (L & (L - 1)) == 0 ? L : 64;
            ps.ngroups = (int32_t)(((int64_t)a.nrow * ps.L + kBlock - 1) / kBlock);
        }
        ps.ngroups = std::max<int32_t>(1, ps.ngroups);
        return ps;
    }
    // the fused form: every block of the next launch totals the product's
    // pairs, so they must be few (<= kPowGrid, and pairs x next grid <= 2^20).
    // Larger matrices take the solve's own products: on the 1414^2 grid the
    // fused launches were slower than the plain row products by 8-10 us each
    // (A^T 20.5 -> 30.2 us, A 20.1 -> 28.0 us, Q 31.9 -> 57.4 us; power
    // method 0.193 -> 0.219 s with the pairs summed by one block first), and
    // dealing the row groups to 1,024 blocks was slower still.
    static bool pow_fits(const PowShape &prod, const PowShape &next) {
        return prod.ngroups <= kPowGrid && (int64_t)prod.ngroups * next.ngroups <= (int64_t)1 << 20;
    }

    template <int MODE>
    static void pow_launch(hipStream_t ps, const DevCsr &a, const PowShape &sh, const double *x, const double *t,
                           double *y, const PowArgs &pa) {
        const dim3 grid((unsigned)sh.ngroups), blk(kBlock);
        const PowMat M{a.nrow,  sh.rb,          sh.ngroups, a.rp.p, a.ci.p, a.v.p, a.code.p, a.dict.p,
                       a.ndict, a.rsc,          a.csc};
        if (sh.dict) {
            k_pow<0, MODE, true><<<grid, blk, 0, ps>>>(M, x, t, y, pa);
            HPRQP_HIP(hipGetLastError());
            return;
        }
        switch (sh.L) {
            case 0: k_pow<0, MODE, false><<<grid, blk, 0, ps>>>(M, x, t, y, pa); break;
            case 1: k_pow<1, MODE, false><<<grid, blk, 0, ps>>>(M, x, t, y, pa); break;
            case 2: k_pow<2, MODE, false><<<grid, blk, 0, ps>>>(M, x, t, y, pa); break;
            case 4: k_pow<4, MODE, false><<<grid, blk, 0, ps>>>(M, x, t, y, pa); break;
            case 8: k_pow<8, MODE, false><<<grid, blk, 0, ps>>>(M, x, t, y, pa); break;
            case 16: k_pow<16, MODE, false><<<grid, blk, 0, ps>>>(M, x, t, y, pa); break;
            case 32: k_pow<32, MODE, false><<<grid, blk, 0, ps>>>(M, x, t, y, pa); break;
            default: k_pow<64, MODE, false><<<grid, blk, 0, ps>>>(M, x, t, y, pa); break;
        }
        HPRQP_HIP(hipGetLastError());
    }

    // One power-method run: which = 0: A A^T (size m); which = 1: Q (size
    // n).  FP64 throughout.  The reference's iteration
    // (src/utils.jl:1357-1376): q = z / |z|, z' = A A^T q (or Q q), lambda =
    // q.z', err = |z' - lambda q| / (|z'| + lambda), stop at the first err <
    // tol; |z' - lambda q|^2 is expanded as |z'|^2 - lambda^2 (2 - |q|^2),
    // with |q|^2 = |z|^2 / |z|^2 as rounded.  Fused form (A A^T and a sparse
    // Q whose products make few pairs, pow_fits; k_pow): each product totals
    // the partial pairs of the one before it, so A A^T takes two launches per
    // iteration (A^T with the stop test, A with its pairs) and Q one, the
    // test recorded in res = {stopped, lambda, iterations, err}.  The other
    // form (larger matrices, other operators) takes the solve's own products
    // (A^T scaled by 1/|z| in its epilogue, then A; or the operator on q =
    // z/|z|), a two-pass reduction and a scalar update, with the scalars in
    // d.  Iterations are enqueued in chunks with a copy of the scalars behind
    // each; in the fused form the host reads one chunk's copy while the next
    // chunk (8 iterations) runs, and the launches of iterations past the stop
    // store nothing; the other form is read after every chunk of 16.
    // (Replaying the 16 from a hipGraph was slower: the random QP's 197
    // iterations took 4.1 ms with the capture and instantiation each call
    // needs, 2.8 ms as launches; one persistent kernel would need grid-wide
    // barriers, 7.5 us each at 64 blocks and 25 us at 256, against 2.6 us for
    // a dependent launch.)
    struct PowerRun {
        int which = 0;
        int64_t len = 0, it = 0, its = 0;
        hipStream_t ps = nullptr;
        DevVec<double> za, zb, q, ATq;
        bool fused = false;
        PowShape sp, sp_at;  // Q or A (the product with pairs); A^T
        // fused form: pairs in two slots of slot_pairs, and qaux = {0,
        // |z_0|^2 | zz ring of 4 | res of 4}
        DevVec<double> fpart, qaux;
        int64_t slot_pairs = 0;
        int slot = 32;                         // other form: scalars at red.out + slot
        double *d = nullptr, *host = nullptr;  // two pinned copies at host, host + 16
        double *zo = nullptr, *zn = nullptr;
        double lambda = 1.0;
        bool done = false;
        // chunks enqueued and read; iterations enqueued up to each copy; the
        // event behind each copy
        int64_t launched = 0, polled = 0, it_at[2] = {0, 0};
        hipEvent_t ev[2] = {nullptr, nullptr};
        PowerRun() = default;
        PowerRun(const PowerRun &) = delete;
        ~PowerRun() {
            for (hipEvent_t e : ev)
                if (e) (void)hipEventDestroy(e);
        }
    };

    // z0 (the reference's randn(1) stream + 1e-8) and |z0|^2, enqueued on s
    void power_start(PowerRun &r, int which, int slot, double *host) {
        r.which = which;
        r.len = which == 0 ? m : n;
        r.ps = s;
        r.slot = slot;
        r.host = host;
        std::vector<double> z0(r.len);
        randn_stream(1, r.len, z0.data());
        for (auto &v : z0) v += 1e-8;
        r.za.alloc(r.len);
        r.zb.alloc(r.len);
        if (which == 0) r.ATq.alloc(n);
        HPRQP_HIP(hipMemcpyAsync(r.za.p, z0.data(), sizeof(double) * r.len, hipMemcpyHostToDevice, s));
        // d: 0 |z|^2, 1 1/|z|, 2 z'.q, 3 |z'|^2, 4 done, 5 lambda at the
        // stop, 6 stop iteration, 7 error, 8 iterations run (the other form;
        // the fused form takes |z_0|^2 from d[0])
        double *d = red.out.p + slot;
        r.d = d;
        HPRQP_HIP(hipMemsetAsync(d, 0, sizeof(double) * 9, s));
        red.enqueue<1, 0>(r.len, SumSq{r.za.p}, slot);
        k_each<<<1, 1, 0, s>>>(1, [=] __device__(int64_t) { d[1] = 1.0 / sqrt(d[0]); });
        HPRQP_HIP(hipGetLastError());
        if (which == 0) {
            r.sp = pow_shape(S.A);
            r.sp_at = pow_shape(S.AT);
            r.fused = pow_fits(r.sp, r.sp_at);
        } else if (S.qkind == QOP_CSR && !S.q_empty) {
            r.sp = pow_shape(S.Q);
            r.fused = pow_fits(r.sp, r.sp);
        }
        if (r.fused) {
            r.slot_pairs = r.sp.ngroups;
            r.fpart.alloc(4 * r.slot_pairs);
            r.qaux.alloc(10);
            double *qa = r.qaux.p;
            HPRQP_HIP(hipMemsetAsync(qa, 0, sizeof(double) * 10, s));
            k_each<<<1, 1, 0, s>>>(1, [=] __device__(int64_t) { qa[1] = d[0]; });
            HPRQP_HIP(hipGetLastError());
        } else if (which == 1) {
            r.q.alloc(r.len);
        }
        r.zo = r.za.p;
        r.zn = r.zb.p;
        for (hipEvent_t &e : r.ev) HPRQP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }

    // up to one chunk of iterations and the scalars' readback, enqueued on
    // r.ps.  Only the fused form may run on a stream other than s: the other
    // uses the shared reduction scratch and the operator's own temporaries.
    void power_chunk(PowerRun &r, int64_t max_it, double tol) {
        const int64_t kPoll = r.fused ? 8 : 16;
        const int64_t chunk = std::min<int64_t>(kPoll, max_it - r.it);
        const hipStream_t ps = r.ps;
        const double *src;
        if (r.fused) {
            double *qa = r.qaux.p, *fp = r.fpart.p;
            const int64_t sp2 = 2 * r.slot_pairs;
            // iteration t's product writes its pairs to slot t & 1
            auto args = [&](int64_t t, bool pairs_out) {
                const double *prev = t == 0 ? (const double *)qa : (const double *)fp + sp2 * ((t - 1) & 1);
                const int pnb = t == 0 ? 1 : r.sp.ngroups;
                return PowArgs{prev, pnb, pairs_out ? fp + sp2 * (t & 1) : nullptr, qa + 2, qa + 6, t, tol,
                               r.which == 0 ? 1 : 0};
            };
            for (int64_t k = 0; k < chunk; ++k) {
                const int64_t t = r.it + k;
                if (r.which == 1) {
                    pow_launch<kPowQ>(ps, S.Q, r.sp, r.zo, nullptr, r.zn, args(t, true));
                } else {
                    pow_launch<kPowAT>(ps, S.AT, r.sp_at, r.zo, nullptr, r.ATq.p, args(t, false));
                    pow_launch<kPowA>(ps, S.A, r.sp, r.ATq.p, r.zo, r.zn, args(t, true));
                }
                std::swap(r.zo, r.zn);
            }
            r.it += chunk;
            // at the limit, the last iteration's test has no next launch
            if (r.it >= max_it) {
                k_pow_tail<<<1, kBlock, 0, ps>>>(args(r.it, false));
                HPRQP_HIP(hipGetLastError());
            }
            src = qa + 6;
        } else {
            HPRQP_REQUIRE(ps == s, HPRQP_E_STATE, "the power method's other form off the solve's stream");
            double *d = r.d;
            const double tl = tol;
            const bool scaled = r.which == 0;
            // tot = {z'.z (A A^T) or z'.q (Q), |z'|^2}; d[0], d[1], d[4], d[8]
            // were written by earlier launches
            auto post = [=] __device__(const double *tot) {
                const double zz = d[0], rz = d[1], done = d[4], itn = d[8] + 1.0;
                const double dot = tot[0], nn = tot[1];
                const double lam = scaled ? dot * rz : dot;
                const double qn = zz * rz * rz;
                const double e2 = fmax(nn - lam * lam * (2.0 - qn), 0.0);
                if (done == 0.0) {
                    const double err = sqrt(e2) / (sqrt(nn) + lam);
                    d[5] = lam;
                    d[6] = itn;
                    d[7] = err;
                    if (err < tl) d[4] = 1.0;
                }
                d[2] = dot;
                d[3] = nn;
                d[0] = nn;
                d[1] = 1.0 / sqrt(nn);
                d[8] = itn;
            };
            for (int64_t k = 0; k < chunk; ++k) {
                const double *zo = r.zo;
                double *zn = r.zn;
                if (r.which == 0) {
                    csr_apply<double, double>(s, S.AT, zo, EpiStoreScaled<double>{r.ATq.p, d + 1});
                    csr_apply<double, double>(s, S.A, (const double *)r.ATq.p, EpiStore<double>{zn});
                    red.enqueue<2, 0>(r.len, Dot2Sq{zn, zo}, r.slot + 2);
                } else {
                    double *qq = r.q.p;
                    each(s, r.len, [=] __device__(int64_t i) { qq[i] = zo[i] * d[1]; });
                    qop_fp64(r.q.p, zn);
                    red.enqueue<2, 0>(r.len, Dot2Sq{zn, r.q.p}, r.slot + 2);
                }
                k_each<<<1, 1, 0, s>>>(1, [=] __device__(int64_t) { post(d + 2); });
                HPRQP_HIP(hipGetLastError());
                std::swap(r.zo, r.zn);
            }
            r.it += chunk;
            src = d + 4;
        }
        const int k = (int)(r.launched & 1);
        HPRQP_HIP(hipMemcpyAsync(r.host + 16 * k, src, sizeof(double) * 4, hipMemcpyDeviceToHost, ps));
        HPRQP_HIP(hipEventRecord(r.ev[k], ps));
        r.it_at[k] = r.it;
        r.launched++;
    }

    // read the oldest unread chunk's copy (waiting for it); sets done once
    // the run has stopped or reached max_it
    void power_poll(PowerRun &r, int64_t max_it) {
        const int k = (int)(r.polled & 1);
        HPRQP_HIP(hipEventSynchronize(r.ev[k]));
        r.polled++;
        // {stopped, lambda, iterations, err}
        const double *h = r.host + 16 * k;
        r.lambda = h[1];
        r.its = (int64_t)h[2];
        if (h[0] != 0.0) r.done = true;
        if (!r.done && r.it_at[k] >= max_it) {
            r.done = true;
            if (p->verbose)
                std::printf("Power iteration did not converge within %lld iterations\n", (long long)max_it);
        }
    }

    // the next chunk, if fewer than two are unread and the limit is not
    // reached
    void power_step(PowerRun &r, int64_t max_it, double tol) {
        if (r.launched - r.polled < 2 && r.it < max_it) power_chunk(r, max_it, tol);
    }

    // lambda_max(A A^T) and lambda_max(Q) (estimate_eigenvalues,
    // src/algorithm.jl:1407-1440), by one run each.  The two runs are
    // independent: when at least one is in the fused form (which uses no
    // shared scratch) that one goes on the context's second stream, its
    // chunks interleaved with the other run's on s, so the small launches of
    // the two overlap on the GPU.  Either may be skipped (need_A / need_Q
    // false).
    void power_pair(bool need_A, bool need_Q, int64_t max_it, double tol, double *lam_A, int64_t *its_A,
                    double *lam_Q, int64_t *its_Q) {
        PowerRun ra, rq;
        if (need_A) power_start(ra, 0, 32, red.host);
        if (need_Q) power_start(rq, 1, 48, red.host + 128);
        // a run in the other form enqueues nothing past its stop test that
        // would return at once: it is read after every chunk
        auto advance = [&](PowerRun &r) {
            power_step(r, max_it, tol);
            if (r.fused) power_step(r, max_it, tol);
        };
        if (need_A && need_Q && ctx->stream2 && (ra.fused || rq.fused)) {
            PowerRun &side = ra.fused ? ra : rq;
            side.ps = ctx->stream2;
            HPRQP_HIP(hipEventRecord(ctx->fork, s));
            HPRQP_HIP(hipStreamWaitEvent(side.ps, ctx->fork, 0));
            while (!ra.done || !rq.done) {
                for (PowerRun *r : {&ra, &rq})
                    if (!r->done) advance(*r);
                for (PowerRun *r : {&ra, &rq})
                    if (!r->done) power_poll(*r, max_it);
            }
            HPRQP_HIP(hipStreamSynchronize(ra.ps));
            HPRQP_HIP(hipStreamSynchronize(rq.ps));
        } else {
            for (PowerRun *r : {&ra, &rq}) {
                if (!(r == &ra ? need_A : need_Q)) continue;
                while (!r->done) {
                    advance(*r);
                    power_poll(*r, max_it);
                }
                HPRQP_HIP(hipStreamSynchronize(r->ps));
            }
        }
        if (need_A) {
            *lam_A = ra.lambda;
            *its_A = ra.its;
        }
        if (need_Q) {
            *lam_Q = rq.lambda;
            *its_Q = rq.its;
        }
    }

    // FP64 operator apply on FP64 vectors (power method)
    void qop_fp64(const double *xin, double *out) {
        switch (S.qkind) {
            case QOP_CSR:
                csr_apply<double, double>(s, S.Q, xin, EpiStore<double>{out});
                break;
            case QOP_LASSO:
                csr_apply<double, double>(s, S.qop->la, xin, EpiStore<double>{qtemp_d.p});
                csr_apply<double, double>(s, S.qop->lat, (const double *)qtemp_d.p, EpiStore<double>{out});
                break;
            case QOP_QAP:
                qap_apply_f64(s, S.qop, xin, out);
                break;
            case QOP_LOWRANK:
                lowrank_apply<double, double, double>(s, n, S.qop->lr_k, S.qop->lr_U.p, S.qop->lr_d.p, xin,
                                                      lr_t.p, out);
                break;
            default: {
                int rc = S.qop->vt.apply(S.qop->user, xin, out, (void *)s);
                HPRQP_REQUIRE(rc == 0, HPRQP_E_ARG, "user Q operator apply failed");
            }
        }
    }

    // ---------------- steps ------------------------------------------
    template <bool FULL, bool DIAG>
    void step_zxw1(double h1, double h2) {
        EpiZxw1<VT, MT, FULL, DIAG> e;
        e.Qw = Qw.p;
        e.ATy = ATy.p;
        e.c = cV;
        e.l = lE;
        e.u = uE;
        e.lS = lS;
        e.uS = uS;
        e.w = w.p;
        e.fact1 = fact1.p;
        e.fact2 = fact2.p;
        e.last_x = last_x.p;
        e.x = x.p;
        e.x_hat = x_hat.p;
        e.w_bar = w_bar.p;
        e.dx = dx.p;
        e.x_bar = x_bar.p;
        e.z_bar = z_bar.p;
        e.sigma = sigma;
        e.f2s = 1.0 / (1.0 + sigma * lambda_max_Q);
        e.f1s = 1.0 - e.f2s;
        e.h1 = h1;
        e.h2 = h2;
        kbegin();
        q_fused(w.p, e, Qw.p);
        kend();
    }
    void step_tempv() {
        EpiTempv<VT> e{x_hat.p, Qw.p, tempv.p, nullptr, sigma};
        q_fused(w_bar.p, e, Qw_bar.p);
    }
    template <bool FULL>
    void step_y(const VT *t, double h1, double h2) {
        if (m == 0) return;
        EpiY<VT, MT, FULL> e;
        e.AL = ALV;
        e.AU = AUE;
        e.last_y = last_y.p;
        e.y = y.p;
        e.y_bar = y_bar.p;
        e.s = s_.p;
        e.dy = dy.p;
        e.fact1 = lambda_max_A * sigma;
        e.fact2 = 1.0 / e.fact1;
        e.h1 = h1;
        e.h2 = h2;
        csr_apply<VT, MT>(s, S.A, t, e);
    }
    template <bool FULL, bool DIAG>
    void step_w2(double h1, double h2) {
        EpiW2<VT, FULL, DIAG> e;
        e.fact = fact.p;
        e.last_w = last_w.p;
        e.last_ATy = last_ATy.p;
        e.w = w.p;
        e.w_bar = w_bar.p;
        e.ATy = ATy.p;
        e.dw = dw.p;
        e.ATdy = ATdy.p;
        e.ATy_bar = ATy_bar.p;
        e.fs = sigma / (1.0 + sigma * lambda_max_Q);
        e.h1 = h1;
        e.h2 = h2;
        csr_apply<VT, MT>(s, S.AT, (const VT *)y_bar.p, e);
    }
    // noC half-steps: z/x/w from A^T y, then Q w_bar with the Halpern Qw
    // update and tempv, then the y step on tempv
    template <bool FULL, bool DIAG>
    void step_noc(double h1, double h2) {
        step_noc_zxw<FULL, DIAG>(h1, h2);
        step_noc_tempv(h1, h2);
        step_y<FULL>(tempv.p, h1, h2);
    }
    template <bool FULL, bool DIAG>
    void step_noc_zxw(double h1, double h2) {
        EpiNocZxw<VT, MT, FULL, DIAG> e;
        e.ATy = ATy.p;
        e.Qw = Qw.p;
        e.c = cV;
        e.l = lE;
        e.u = uE;
        e.lS = lS;
        e.uS = uS;
        e.fact1 = fact1.p;
        e.fact2 = fact2.p;
        e.last_x = last_x.p;
        e.last_w = last_w.p;
        e.x = x.p;
        e.x_hat = x_hat.p;
        e.w = w.p;
        e.w_bar = w_bar.p;
        e.dx = dx.p;
        e.x_bar = x_bar.p;
        e.z_bar = z_bar.p;
        e.dw = dw.p;
        e.sigma = sigma;
        e.f2s = 1.0 / (1.0 + sigma * lambda_max_Q);
        e.f1s = 1.0 - e.f2s;
        e.h1 = h1;
        e.h2 = h2;
        kbegin();
        csr_apply<VT, MT>(s, S.AT, (const VT *)y.p, e);
        kend();
    }
    void step_noc_tempv(double h1, double h2) {
        EpiNocTempv<VT> e2{x_hat.p, last_Qw.p, Qw.p, Qw_bar.p, tempv.p, sigma, h1, h2};
        q_fused(w_bar.p, e2, Qw_bar.p);
    }
    template <bool FULL>
    void step_lasso(double h1, double h2) {
        EpiLasso<VT, MT, FULL> e;
        e.Qw = Qw.p;
        e.ATy = ATy.p;
        e.c = cV;
        e.last_x = last_x.p;
        e.last_w = last_w.p;
        e.x = x.p;
        e.x_hat = x_hat.p;
        e.w = w.p;
        e.w_bar = w_bar.p;
        e.dx = dx.p;
        e.x_bar = x_bar.p;
        e.z_bar = z_bar.p;
        e.dw = dw.p;
        e.sigma = sigma;
        e.lambda_sigma = lambda * sigma;
        e.f2s = 1.0 / (1.0 + sigma * lambda_max_Q);
        e.f1s = 1.0 - e.f2s;
        e.h1 = h1;
        e.h2 = h2;
        kbegin();
        q_fused(w.p, e, Qw.p);
        kend();
    }
    // LP half-steps: z/x from A^T y, then the y step on x_hat
    template <bool FULL>
    void step_lp(double h1, double h2) {
        step_lp_zx<FULL>(h1, h2);
        step_y<FULL>(x_hat.p, h1, h2);
    }
    template <bool FULL>
    void step_lp_zx(double h1, double h2) {
        EpiLpZx<VT, MT, FULL> e;
        e.ATy = ATy.p;
        e.c = cV;
        e.l = lE;
        e.u = uE;
        e.lS = lS;
        e.uS = uS;
        e.last_x = last_x.p;
        e.x = x.p;
        e.x_hat = x_hat.p;
        e.dx = dx.p;
        e.x_bar = x_bar.p;
        e.z_bar = z_bar.p;
        e.sigma = sigma;
        e.h1 = h1;
        e.h2 = h2;
        kbegin();
        csr_apply<VT, MT>(s, S.AT, (const VT *)y.p, e);
        kend();
    }

    template <bool FULL, bool DIAG>
    void update_impl(double h1, double h2) {
        if (qds) {
            qds_update<FULL>(h1, h2);
            return;
        }
        if (S.qkind == QOP_LASSO) {
            step_lasso<FULL>(h1, h2);
            return;
        }
        if (!S.q_empty) {
            if (noC) {
                step_noc<FULL, DIAG>(h1, h2);
            } else {
                step_zxw1<FULL, DIAG>(h1, h2);
                if (m > 0) step_tempv();
                step_y<FULL>(tempv.p, h1, h2);
                step_w2<FULL, DIAG>(h1, h2);
            }
        } else {
            step_lp<FULL>(h1, h2);
        }
    }

    // main_update_gpu! (src/algorithm.jl:1250-1279)
    void main_update(int64_t inner, bool full) {
        const double h1 = 1.0 / ((double)inner + 2.0);
        const double h2 = 1.0 - h1;
        const bool diag = S.Q_is_diag;
        timing_now = kt != nullptr && !full;
        if (full)
            diag ? update_impl<true, true>(h1, h2) : update_impl<true, false>(h1, h2);
        else
            diag ? update_impl<false, true>(h1, h2) : update_impl<false, false>(h1, h2);
    }

    // ---------------- residuals (src/algorithm.jl:37-133) ------------
    struct ResN {
        const double *c, *Qx, *ATdy, *cn, *l, *u;
        const VT *xb, *zb;
        double c_scale, b_scale, lam;
        int has_m, iter0;
        __device__ void operator()(int64_t i, double *s, double *mx) const {
            const double x = (double)xb[i], q = Qx[i], z = (double)zb[i], ci = c[i];
            s[0] += ci * x;
            s[1] += x * q;
            s[2] += z * x;
            s[3] += fabs(lam * x);
            const double sf = cn[i] * c_scale;
            const double a = has_m ? ATdy[i] : 0.0;
            mx[0] = dmax(mx[0], fabs((q + ci - a - z) * sf));
            mx[1] = dmax(mx[1], fabs(a * sf));
            mx[2] = dmax(mx[2], fabs(q * sf));
            if (iter0) {
                const double v = (dmax(l[i] - x, 0.0) + dmax(x - u[i], 0.0)) * (b_scale / cn[i]);
                mx[3] = dmax(mx[3], fabs(v));
            }
        }
    };
    struct ResM {
        const double *AL, *AU, *Ax, *rn;
        const VT *yb, *sv;
        double b_scale;
        __device__ void operator()(int64_t i, double *s, double *mx) const {
            s[0] += (double)yb[i] * (double)sv[i];
            const double ax = Ax[i];
            const double corr = proj(ax, AL[i], AU[i]) - ax;
            const double sf = rn[i] * b_scale;
            mx[0] = dmax(mx[0], fabs(corr * sf));
            mx[1] = dmax(mx[1], fabs(ax * sf));
        }
    };

    using Resid = hprqp::Resid;

    // the residual products and reductions, results at red.out + slot
    void residuals_enqueue(int64_t iter, int slot) {
        flush();
        q_apply_d(x_bar.p, Qx_d.p);
        if (m > 0) {
            csr_apply<double, double>(s, S.AT, (const VT *)y_bar.p, EpiStore<double>{ATdy_d.p});
            csr_apply<double, double>(s, S.A, (const VT *)x_bar.p, EpiStore<double>{Ax_d.p});
        }
        red.enqueue<4, 4>(n, ResN{S.c.p, Qx_d.p, ATdy_d.p, S.col_norm.p, S.l.p, S.u.p, x_bar.p, z_bar.p,
                                  S.c_scale, S.b_scale, lambda, m > 0, iter == 0},
                          slot);
        red.enqueue<1, 2>(m, ResM{S.AL.p, S.AU.p, Ax_d.p, S.row_norm.p, y_bar.p, s_.p, S.b_scale}, slot + 8);
    }

    void compute_residuals(int64_t iter, double obj_constant, bool lasso, Resid &r) {
        residuals_enqueue(iter, 0);
        residuals_from(red.fetch(11), iter, obj_constant, lasso, r);
    }

    // A print step's residuals without the wait (between two check
    // iterations only the termination test acts on them): the reductions go
    // to their own slot and pinned staging, an event marks them, and the
    // iteration goes on; residuals_wait finishes them at the next iteration.
    // The iterate of that step stays in x_bar and z_bar until the next full
    // step; y_bar and w_bar, which the partial step rewrites, are copied
    // aside (with x_bar and z_bar, for the paths whose partial steps write
    // those) for a solve that stops there.
    static constexpr int kAsyncSlot = 64, kAsyncHost = 224;
    hipEvent_t res_ev = nullptr;
    DevVec<VT> snap_x, snap_z, snap_w, snap_y;
    bool async_ok() const { return true; }
    void residuals_start(int64_t iter) {
        residuals_enqueue(iter, kAsyncSlot);
        HPRQP_HIP(hipMemcpyAsync(red.host + kAsyncHost, red.out.p + kAsyncSlot, sizeof(double) * 11,
                                 hipMemcpyDeviceToHost, s));
        if (!res_ev) HPRQP_HIP(hipEventCreateWithFlags(&res_ev, hipEventDisableTiming));
        HPRQP_HIP(hipEventRecord(res_ev, s));
        if (snap_x.n != n) {
            snap_x.alloc(n);
            snap_z.alloc(n);
            snap_w.alloc(n);
            snap_y.alloc(m);
        }
        HPRQP_HIP(hipMemcpyAsync(snap_x.p, x_bar.p, sizeof(VT) * n, hipMemcpyDeviceToDevice, s));
        HPRQP_HIP(hipMemcpyAsync(snap_z.p, z_bar.p, sizeof(VT) * n, hipMemcpyDeviceToDevice, s));
        HPRQP_HIP(hipMemcpyAsync(snap_w.p, w_bar.p, sizeof(VT) * n, hipMemcpyDeviceToDevice, s));
        if (m > 0) HPRQP_HIP(hipMemcpyAsync(snap_y.p, y_bar.p, sizeof(VT) * m, hipMemcpyDeviceToDevice, s));
    }
    void residuals_wait(int64_t iter, double obj_constant, bool lasso, Resid &r) {
        HPRQP_HIP(hipEventSynchronize(res_ev));
        residuals_from(red.host + kAsyncHost, iter, obj_constant, lasso, r);
    }
    // the solve stops at the print step: its iterate back in place.  The
    // QAP one-apply form's update left pending by the step after it (the
    // second w update, and for nq < 256 the y update) belongs to that later
    // step and is dropped, as the waiting form never ran it.
    void restore_snapshot() {
        pend = false;
        HPRQP_HIP(hipMemcpyAsync(x_bar.p, snap_x.p, sizeof(VT) * n, hipMemcpyDeviceToDevice, s));
        HPRQP_HIP(hipMemcpyAsync(z_bar.p, snap_z.p, sizeof(VT) * n, hipMemcpyDeviceToDevice, s));
        HPRQP_HIP(hipMemcpyAsync(w_bar.p, snap_w.p, sizeof(VT) * n, hipMemcpyDeviceToDevice, s));
        if (m > 0) HPRQP_HIP(hipMemcpyAsync(y_bar.p, snap_y.p, sizeof(VT) * m, hipMemcpyDeviceToDevice, s));
        HPRQP_HIP(hipStreamSynchronize(s));
    }
    ~Solver() {
        if (res_ev) (void)hipEventDestroy(res_ev);
    }

    void residuals_from(const double *h, int64_t iter, double obj_constant, bool lasso, Resid &r) const {
        const double bc = S.b_scale * S.c_scale;
        r.primal_obj = bc * (h[0] + 0.5 * h[1]) + obj_constant;
        r.dual_obj = bc * (-0.5 * h[1] + h[2]) + obj_constant;
        if (m > 0) r.dual_obj += bc * h[8];
        if (lasso) {
            r.primal_obj += bc * h[3];
            r.dual_obj += bc * h[3];
        }
        r.rel_gap = std::fabs(r.primal_obj - r.dual_obj) /
                    (1.0 + std::max(std::fabs(r.primal_obj), std::fabs(r.dual_obj)));
        r.err_Rd = h[4] / (1.0 + std::max(std::max(S.norm_c_org, h[5]), h[6]));
        r.err_Rp = m > 0 ? h[9] / (1.0 + std::max(S.norm_b_org, h[10])) : 0.0;
        if (iter == 0) r.err_Rp = std::max(r.err_Rp, h[7]);
        r.kkt = std::max(std::max(r.err_Rp, r.err_Rd), r.rel_gap);
    }

    // ---------------- M-norm (src/algorithm.jl:689-739) --------------
    template <class P>
    struct MnN {
        const VT *dx;
        const P *dQw, *ATdy, *QATdy;
        const VT *dw, *fM;
        int diag;
        __device__ void operator()(int64_t i, double *s, double *) const {
            const double a = (double)dx[i], b = (double)dQw[i], c = ATdy ? (double)ATdy[i] : 0.0,
                         d = QATdy ? (double)QATdy[i] : 0.0, e = (double)dw[i];
            s[0] += a * a;
            s[1] += b * a;
            s[2] += b * c;
            s[3] += c * a;
            s[4] += (diag ? c * (double)fM[i] : c) * d;
            s[5] += b * b;
            s[6] += e * b;
        }
    };
    struct SumSqV {
        const VT *v;
        __device__ void operator()(int64_t i, double *s, double *) const {
            const double a = (double)v[i];
            s[0] += a * a;
        }
    };

    double compute_M_norm() {
        flush();
        if constexpr (std::is_same<VT, double>::value) {
            q_apply(dw.p, dQw.p);
            if (m > 0) {
                csr_apply<VT, MT>(s, S.AT, (const VT *)dy.p, EpiStore<VT>{ATdy.p});
                q_apply(ATdy.p, QATdy.p);
            }
            red.enqueue<7, 0>(n,
                              MnN<VT>{dx.p, dQw.p, m > 0 ? ATdy.p : nullptr, m > 0 ? QATdy.p : nullptr, dw.p,
                                      fact_M.p, S.Q_is_diag},
                              0);
        } else {
            ensure_sig64();
            q_apply_d(dw.p, sg_Qw.p);
            if (m > 0) {
                csr_apply<double, double>(s, S.AT, (const VT *)dy.p, EpiStore<double>{ATdy_d.p});
                qop_fp64(ATdy_d.p, sg_QATdy.p);
            }
            red.enqueue<7, 0>(n,
                              MnN<double>{dx.p, sg_Qw.p, m > 0 ? ATdy_d.p : nullptr,
                                          m > 0 ? sg_QATdy.p : nullptr, dw.p, fact_M.p, S.Q_is_diag},
                              0);
        }
        red.enqueue<1, 0>(m, SumSqV{dy.p}, 7);
        const double *h = red.fetch(8);
        double M1 = 0.0, M3 = 0.0;
        double M2 = (1.0 / sigma) * h[0];
        M2 -= 2.0 * h[1];
        if (m > 0) {
            M1 = sigma * lambda_max_A * h[7];
            M1 -= 2.0 * sigma * h[2];
            M2 += 2.0 * h[3];
            if (S.Q_is_diag) {
                M3 = h[4];
                M1 += sigma * h[5];
            } else {
                M3 = (sigma * sigma) / (1.0 + sigma * lambda_max_Q) * h[4];
                M1 += sigma * lambda_max_Q * h[6];
            }
        } else if (!S.Q_is_diag) {
            M1 = sigma * lambda_max_Q * h[6];
        }
        M2 += std::max(M1, 0.0);
        const double Mn = std::max(M2, 0.0) + std::max(M3, 0.0);
        return std::sqrt(Mn);
    }

    // ---------------- diagonal-Q factors ------------------------------
    void update_Q_factors() {
        const double sg = sigma, s2 = sigma * sigma;
        const double *dq = S.diag_Q.p;
        VT *f2 = fact2.p, *f = fact.p, *f1 = fact1.p, *fm = fact_M.p;
        each(s, n, [=] __device__(int64_t i) {
            const double t = 1.0 / (1.0 + sg * dq[i]);
            f2[i] = (VT)t;
            f[i] = (VT)(sg * t);
            f1[i] = (VT)(sg * dq[i] * t);
            fm[i] = (VT)(s2 * t);
        });
    }
};

// ---------------------------------------------------------------------
// restart bookkeeping (host, src/algorithm.jl:441-555, 1154-1171)
// ---------------------------------------------------------------------
struct Restart {
    int restart_flag = 0;
    bool first_restart = true;
    double last_gap = INFINITY, current_gap = INFINITY, save_gap = INFINITY;
    int64_t inner = 0, sufficient = 0, necessary = 0, long_ = 0, times = 0;
    double weighted_norm = INFINITY, best_gap = INFINITY, best_sigma = 1.0;

    void check(int64_t iter, int64_t check_iter, double sigma) {
        restart_flag = 0;
        if (first_restart) {
            if (iter == check_iter) {
                first_restart = false;
                restart_flag = 1;
                weighted_norm = current_gap;
                best_gap = current_gap;
                best_sigma = sigma;
            }
        } else if (iter % check_iter == 0) {
            if (current_gap < 0) current_gap = 1e-6;
            if (current_gap <= 0.2 * last_gap) {
                sufficient++;
                restart_flag = 1;
            }
            if (current_gap <= 0.8 * last_gap && current_gap > 1.00 * save_gap) {
                necessary++;
                restart_flag = 2;
            }
            const double fact = (current_gap / weighted_norm > 1e-1) ? 0.5 : 0.2;
            if ((double)inner >= fact * (double)iter) {
                long_++;
                restart_flag = 3;
            }
            if (best_gap > current_gap) {
                best_gap = current_gap;
                best_sigma = sigma;
            }
            save_gap = current_gap;
        }
    }
};

// golden-section searches (src/kernels.jl:1449-1486; src/unified_operations.jl:348-386)
static double golden(double ap, double bp, double cp, double dp, double lo, double hi, double tol, int maxiter) {
    const double phi = (std::sqrt(5.0) - 1.0) / 2.0;
    auto f = [&](double x) { return ap * x + bp / x + cp * x * x / (1 + dp * x); };
    double a = lo, b = hi;
    double c = b - phi * (b - a), d = a + phi * (b - a);
    double fc = f(c), fd = f(d);
    for (int i = 1; i <= maxiter; ++i) {
        if (fd < fc) {
            a = c;
            c = d;
            fc = fd;
            d = a + phi * (b - a);
            fd = f(d);
        } else {
            b = d;
            d = c;
            fd = fc;
            c = b - phi * (b - a);
            fc = f(c);
        }
        if ((b - a) < tol) break;
    }
    return 0.5 * (a + b);
}

static double golden_Q_diag(double a, double b, const std::vector<double> &Q, const std::vector<double> &c,
                            const std::vector<double> &d, double lo, double hi, double tol, int maxiter) {
    const double phi = (std::sqrt(5.0) - 1.0) / 2.0;
    auto f = [&](double x) {
        double acc = 0.0;
        for (size_t i = 0; i < Q.size(); ++i) acc += c[i] * (d[i] / (1.0 + x * Q[i]));
        return a * x + b / x + x * x * acc;
    };
    double x1 = hi - phi * (hi - lo), x2 = lo + phi * (hi - lo);
    double f1 = f(x1), f2 = f(x2);
    int iter = 0;
    while (std::fabs(hi - lo) > tol * std::max(1.0, std::fabs(lo)) && iter < maxiter) {
        if (f1 > f2) {
            lo = x1;
            x1 = x2;
            f1 = f2;
            x2 = lo + phi * (hi - lo);
            f2 = f(x2);
        } else {
            hi = x2;
            x2 = x1;
            f2 = f1;
            x1 = hi - phi * (hi - lo);
            f1 = f(x1);
        }
        iter++;
    }
    return (f1 < f2) ? x1 : x2;
}

// The same minimisation on the device, for large n: f(x) = a x + b/x +
// x^2 sum_i c_i (d_i / (1 + x Q_i)) with c_i d_i = Q_i c_i^2 >= 0 is convex
// on x > 0, so its minimiser in [1e-12, 1e12] is where f'(x) = a - b/x^2 +
// sum_i c_i d_i x (2 + x Q_i) / (1 + x Q_i)^2 changes sign (or an end of the
// interval).  Each pass evaluates f' at 64 points spaced evenly in log x over
// the bracket (one point per lane: a wave loads 64 elements at once and
// walks them by shuffles) and keeps the two points around the first
// non-negative value.  A pass shrinks the bracket's log width by 63; ten
// take [1e-12, 1e12] (log width 55) below 1e-13 relative, the golden
// search's tolerance.  The sign of f' is fixed to rounding where f itself
// is flat (near the minimum f varies by ~x^2 f''), so the point found moves
// by about 1e-16 relative under a rounding change of the inputs, where
// comparisons of f values move by about 1e-8.  br = {lo, hi, best}.
constexpr int kSigPoints = 64;
constexpr int kSigPasses = 10;
// n from which the device search runs (below it the golden search on two
// host copies of length n costs less than the 20 launches)
constexpr int32_t kSigmaDeviceMin = 2048;

__device__ __forceinline__ double sig_point(double llo, double lhi, int k) {
    return exp(llo + (lhi - llo) * ((double)k / (double)(kSigPoints - 1)));
}

template <class VT, class P>
__global__ __launch_bounds__(kBlock) void k_sigma_terms(int32_t n, const double *__restrict__ Q,
                                                        const VT *__restrict__ c, const P *__restrict__ d,
                                                        const double *__restrict__ br, double *__restrict__ part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double x = sig_point(log(br[0]), log(br[1]), lane);
    const int64_t tiles = ((int64_t)n + 63) / 64, waves = (int64_t)gridDim.x * (kBlock / 64);
    double acc = 0.0;
    for (int64_t tile = (int64_t)blockIdx.x * (kBlock / 64) + wave; tile < tiles; tile += waves) {
        const int64_t i = tile * 64 + lane;
        const double qi = i < n ? Q[i] : 0.0;
        const double cd = i < n ? (double)c[i] * (double)d[i] : 0.0;
        const int cnt = (int)min<int64_t>(64, (int64_t)n - tile * 64);
        for (int j = 0; j < cnt; ++j) {
            const double q = __shfl(qi, j), w = __shfl(cd, j);
            const double r = 1.0 / (1.0 + x * q);
            acc += w * (x * (2.0 + x * q) * r * r);
        }
    }
    __shared__ double sh[kBlock / 64][kSigPoints];
    sh[wave][lane] = acc;
    __syncthreads();
    if (threadIdx.x < kSigPoints)
        part[(int64_t)blockIdx.x * kSigPoints + lane] = (sh[0][lane] + sh[1][lane]) + (sh[2][lane] + sh[3][lane]);
}

// the sums of the nb blocks' partials per point, f' at the 64 points, and
// the new bracket (one block of 1,024 threads; wave 0 picks)
__global__ __launch_bounds__(1024) void k_sigma_pick(int nb, const double *__restrict__ part, double a, double b,
                                                     double *__restrict__ br) {
    const int k = threadIdx.x & 63, j = threadIdx.x >> 6;
    double s = 0.0;
    for (int blk = j; blk < nb; blk += 16) s += part[(int64_t)blk * kSigPoints + k];
    __shared__ double sh[16][kSigPoints];
    sh[j][k] = s;
    __syncthreads();
    if (threadIdx.x >= 64) return;  // wave 0 only from here
    double S = 0.0;
    for (int w = 0; w < 16; ++w) S += sh[w][k];
    const double llo = log(br[0]), lhi = log(br[1]);
    const double x = sig_point(llo, lhi, k);
    const double fp = a - b / (x * x) + S;
    // the first point where f' >= 0 (a NaN counts as >= 0); 64 if none
    const unsigned long long up = __ballot(!(fp < 0.0));
    const int k0 = up ? __ffsll((long long)up) - 1 : kSigPoints;
    // every lane of the wave has read br before lane 0 writes it
    if (k == 0) {
        double nlo, nhi;
        if (k0 == 0) {
            nlo = nhi = sig_point(llo, lhi, 0);  // f' >= 0 at the lower end
        } else if (k0 == kSigPoints) {
            nlo = nhi = sig_point(llo, lhi, kSigPoints - 1);  // f' < 0 up to the upper end
        } else {
            nlo = sig_point(llo, lhi, k0 - 1);
            nhi = sig_point(llo, lhi, k0);
        }
        br[0] = nlo;
        br[1] = nhi;
        br[2] = exp(0.5 * (log(nlo) + log(nhi)));
    }
}

template <class VT, class P>
struct SigmaDots {
    const VT *dx;
    const P *dQw;
    const VT *ATdy;
    const P *QATdy;
    const VT *dw;
    __device__ void operator()(int64_t i, double *s, double *) const {
        const double a = (double)dx[i], b = (double)dQw[i], c = ATdy ? (double)ATdy[i] : 0.0,
                     d = QATdy ? (double)QATdy[i] : 0.0, e = (double)dw[i];
        s[0] += a * a;
        s[1] += b * c;
        s[2] += b * b;
        s[3] += e * b;
        s[4] += c * d;
    }
};

// the device half of update_sigma! (src/algorithm.jl:323-360): the
// differences to the anchors, Q dw and Q A^T dy, and six dot products:
// h = <dx,dx>, <Q dw, A^T dy>, <Q dw, Q dw>, <dw, Q dw>, <A^T dy, Q A^T dy>,
// <dy,dy>
template <class VT, class MT>
static void sigma_dots(Solver<VT, MT> &W, double *h) {
    W.flush();
    hipStream_t s = W.s;
    const int32_t n = W.n, m = W.m;
    {
        VT *dx = W.dx.p, *dw = W.dw.p;
        const VT *xb = W.x_bar.p, *wb = W.w_bar.p;
        const double *lx = W.last_x.p, *lw = W.last_w.p;
        each(s, n, [=] __device__(int64_t i) {
            dx[i] = (VT)((double)xb[i] - lx[i]);
            dw[i] = (VT)((double)wb[i] - lw[i]);
        });
    }
    if (m > 0) {
        VT *dy = W.dy.p, *atdy = W.ATdy.p;
        const VT *yb = W.y_bar.p, *atyb = W.ATy_bar.p;
        const double *ly = W.last_y.p, *laty = W.last_ATy.p;
        each(s, m, [=] __device__(int64_t i) { dy[i] = (VT)((double)yb[i] - ly[i]); });
        each(s, n, [=] __device__(int64_t i) { atdy[i] = (VT)((double)atyb[i] - laty[i]); });
    }
    if constexpr (std::is_same<VT, double>::value) {
        W.q_apply(W.dw.p, W.dQw.p);
        if (m > 0) W.q_apply(W.ATdy.p, W.QATdy.p);
        W.red.template enqueue<5, 0>(n,
                                     SigmaDots<VT, VT>{W.dx.p, W.dQw.p, m > 0 ? W.ATdy.p : nullptr,
                                                       m > 0 ? W.QATdy.p : nullptr, W.dw.p},
                                     0);
    } else {
        W.ensure_sig64();
        W.q_apply_d(W.dw.p, W.sg_Qw.p);
        if (m > 0) W.q_apply_d(W.ATdy.p, W.sg_QATdy.p);
        W.red.template enqueue<5, 0>(n,
                                     SigmaDots<VT, double>{W.dx.p, W.sg_Qw.p, m > 0 ? W.ATdy.p : nullptr,
                                                           m > 0 ? W.sg_QATdy.p : nullptr, W.dw.p},
                                     0);
    }
    W.red.template enqueue<1, 0>(m, typename Solver<VT, MT>::SumSqV{W.dy.p}, 5);
    const double *f = W.red.fetch(6);
    for (int k = 0; k < 6; ++k) h[k] = f[k];
}

// The sigma estimate of a diagonal Q with constraints: the minimiser of
// unified_golden_Q_diag's function (src/unified_operations.jl:348-386) over
// A^T dy and Q A^T dy as sigma_dots left them -- the reference's golden
// search on host copies below kSigmaDeviceMin, the device's search on the
// sign of f' from there.
template <class VT, class MT>
static double sigma_estimate_diag(Solver<VT, MT> &W, double a, double b) {
    hipStream_t s = W.s;
    const int32_t n = W.n;
    if (n >= kSigmaDeviceMin) {
        const int64_t tiles = ((int64_t)n + 63) / 64;
        const int nb = (int)std::min<int64_t>(1024, (tiles + 3) / 4);
        if (W.sig_buf.n < 3 + (int64_t)nb * kSigPoints) W.sig_buf.alloc(3 + (int64_t)nb * kSigPoints);
        double *br = W.sig_buf.p, *part = W.sig_buf.p + 3;
        k_each<<<1, 1, 0, s>>>(1, [=] __device__(int64_t) {
            br[0] = 1e-12;
            br[1] = 1e12;
            br[2] = 0.0;
        });
        HPRQP_HIP(hipGetLastError());
        for (int pass = 0; pass < kSigPasses; ++pass) {
            if constexpr (std::is_same<VT, double>::value)
                k_sigma_terms<VT, VT><<<nb, kBlock, 0, s>>>(n, W.S.diag_Q.p, W.ATdy.p, W.QATdy.p, br, part);
            else
                k_sigma_terms<VT, double><<<nb, kBlock, 0, s>>>(n, W.S.diag_Q.p, W.ATdy.p, W.sg_QATdy.p, br, part);
            k_sigma_pick<<<1, 1024, 0, s>>>(nb, part, a, b, br);
        }
        HPRQP_HIP(hipGetLastError());
        HPRQP_HIP(hipMemcpyAsync(W.red.host, br + 2, sizeof(double), hipMemcpyDeviceToHost, s));
        HPRQP_HIP(hipStreamSynchronize(s));
        return W.red.host[0];
    }
    std::vector<VT> hc(n);
    std::vector<double> dd(n);
    HPRQP_HIP(hipMemcpyAsync(hc.data(), W.ATdy.p, sizeof(VT) * n, hipMemcpyDeviceToHost, s));
    if constexpr (std::is_same<VT, double>::value) {
        HPRQP_HIP(hipMemcpyAsync(dd.data(), W.QATdy.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    } else {
        HPRQP_HIP(hipMemcpyAsync(dd.data(), W.sg_QATdy.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    }
    HPRQP_HIP(hipStreamSynchronize(s));
    std::vector<double> cc(hc.begin(), hc.end());
    return golden_Q_diag(a, b, W.S.diag_Q_host, cc, dd, 1e-12, 1e12, 1e-13, 200);
}

// update_sigma! (src/algorithm.jl:323-423)
template <class VT, class MT>
static void update_sigma(Solver<VT, MT> &W, Restart &r, double err_Rd, double err_Rp, double rel_gap) {
    const hprqp_params *p = W.p;
    if (p->sigma_fixed || r.restart_flag < 1) return;
    hipStream_t s = W.s;
    const int32_t n = W.n, m = W.m;
    const double sigma_old = W.sigma;
    double h[6];
    sigma_dots(W, h);
    const double dxdx = h[0], dQwATdy = h[1], dQwdQw = h[2], dwdQw = h[3], ATdyQATdy = h[4], dydy = h[5];
    double a = 0.0, b = dxdx, c = 0.0, d = 0.0;
    if (m > 0) a = W.lambda_max_A * dydy - 2 * dQwATdy;
    if (W.S.Q_is_diag) {
        a += dQwdQw;
    } else {
        a += W.lambda_max_Q * dwdQw;
        if (m > 0) {
            c = ATdyQATdy;
            d = W.lambda_max_Q;
        }
    }
    a = std::max(a, 1e-12);
    b = std::max(b, 1e-12);
    double est;
    if (W.S.Q_is_diag) {
        est = m > 0 ? sigma_estimate_diag(W, a, b) : std::sqrt(b / a);
    } else {
        est = m > 0 ? golden(a, b, c, d, 1e-12, 1e12, 1e-13, 200) : std::sqrt(b / a);
    }
    const double fact = std::exp(-0.05 * (r.current_gap / r.best_gap));
    const double temp_1 = std::max(std::min(err_Rd, err_Rp), std::min(rel_gap, r.current_gap));
    const double cand = std::exp(fact * std::log(est) + (1 - fact) * std::log(r.best_sigma));
    auto clampj = [](double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); };
    double kappa;
    if (temp_1 > 9e-10)
        kappa = 1.0;
    else if (temp_1 > 5e-10)
        kappa = clampj(std::sqrt(err_Rd / err_Rp), 1e-2, 100.0);
    else
        kappa = clampj(err_Rd / err_Rp, 1e-2, 100.0);
    W.sigma = kappa * cand;
    if (W.S.Q_is_diag && std::fabs(sigma_old - W.sigma) > 1e-15) W.update_Q_factors();
}

// the device half of do_restart! (src/algorithm.jl:522-555): the anchors
// and iterates from the averaged iterates, A^T y_bar
template <class VT, class MT>
static void restart_vectors(Solver<VT, MT> &W) {
    W.flush();
    hipStream_t s = W.s;
    const int32_t n = W.n, m = W.m;
    if (m > 0) {
        csr_apply<VT, MT>(s, W.S.AT, (const VT *)W.y_bar.p, EpiStore<VT>{W.ATy_bar.p});
        VT *y = W.y.p;
        const VT *yb = W.y_bar.p;
        double *ly = W.last_y.p;
        each(s, m, [=] __device__(int64_t i) {
            y[i] = yb[i];
            ly[i] = (double)yb[i];
        });
    }
    {
        VT *x = W.x.p, *w = W.w.p, *aty = W.ATy.p, *qw = W.Qw.p;
        const VT *xb = W.x_bar.p, *wb = W.w_bar.p, *atyb = W.ATy_bar.p, *qwb = W.Qw_bar.p;
        double *lx = W.last_x.p, *lw = W.last_w.p, *laty = W.last_ATy.p, *lqw = W.last_Qw.p;
        const bool hm = m > 0, noc = W.noC;
        each(s, n, [=] __device__(int64_t i) {
            x[i] = xb[i];
            w[i] = wb[i];
            lx[i] = (double)xb[i];
            lw[i] = (double)wb[i];
            if (hm) {
                laty[i] = (double)atyb[i];
                aty[i] = atyb[i];
            }
            if (noc) {
                qw[i] = qwb[i];
                lqw[i] = (double)qwb[i];
            }
        });
    }
    // QAP one-apply form: w = w_bar' and ATy = A^T y_bar, whose products the
    // full step's second w update has formed
    if (W.qds) {
        if (W.qv_ok && W.flushed_full) {
            VT *qw = W.Qw.p, *qaty = W.QATy.p;
            const VT *qwn = W.Qwb_new.p, *qg = W.Qg.p;
            double *lqw = W.last_Qw.p, *lqaty = W.last_QATy.p;
            each(s, n, [=] __device__(int64_t i) {
                qw[i] = qwn[i];
                lqw[i] = (double)qwn[i];
                qaty[i] = qg[i];
                lqaty[i] = (double)qg[i];
            });
        } else {
            W.qv_ok = false;
        }
    }
}

// do_restart! (src/algorithm.jl:522-555)
template <class VT, class MT>
static void do_restart(Solver<VT, MT> &W, Restart &r) {
    if (r.restart_flag <= 0) return;
    restart_vectors(W);
    r.last_gap = r.current_gap;
    r.save_gap = INFINITY;
    r.times++;
    r.inner = 0;
}

// process_initial_points! (src/algorithm.jl:1783-1836)
template <class VT, class MT>
static void process_initial_points(Solver<VT, MT> &W, const double *x0, const double *y0) {
    W.invalidate_products();
    hipStream_t s = W.s;
    const int32_t n = W.n, m = W.m;
    const double bs = W.S.b_scale, cs = W.S.c_scale;
    if (x0) {
        DevVec<double> t(n);
        HPRQP_HIP(hipMemcpyAsync(t.p, x0, sizeof(double) * n, hipMemcpyHostToDevice, s));
        const double *tp = t.p, *cn = W.S.col_norm.p;
        VT *x = W.x.p, *xb = W.x_bar.p, *w = W.w.p, *wb = W.w_bar.p;
        double *lx = W.last_x.p, *lw = W.last_w.p;
        each(s, n, [=] __device__(int64_t i) {
            const double v = tp[i] * cn[i] / bs;
            x[i] = xb[i] = w[i] = wb[i] = (VT)v;
            lx[i] = lw[i] = (double)(VT)v;
        });
        HPRQP_HIP(hipStreamSynchronize(s));
    }
    if (y0) {
        if (m > 0) {
            DevVec<double> t(m);
            HPRQP_HIP(hipMemcpyAsync(t.p, y0, sizeof(double) * m, hipMemcpyHostToDevice, s));
            const double *tp = t.p, *rn = W.S.row_norm.p;
            VT *y = W.y.p, *yb = W.y_bar.p;
            double *ly = W.last_y.p;
            each(s, m, [=] __device__(int64_t i) {
                const double v = tp[i] * rn[i] / cs;
                y[i] = yb[i] = (VT)v;
                ly[i] = (double)(VT)v;
            });
            csr_apply<VT, MT>(s, W.S.AT, (const VT *)W.y_bar.p, EpiStore<VT>{W.ATy_bar.p});
            VT *aty = W.ATy.p;
            const VT *atyb = W.ATy_bar.p;
            double *laty = W.last_ATy.p;
            each(s, n, [=] __device__(int64_t i) {
                aty[i] = atyb[i];
                laty[i] = (double)atyb[i];
            });
            HPRQP_HIP(hipStreamSynchronize(s));
        }
        W.q_apply_d(W.x_bar.p, W.Qx_d.p);
        {
            const double *qx = W.Qx_d.p;
            const VT *xb = W.x_bar.p, *atyb = W.ATy_bar.p;
            const MT *c = W.cV;
            VT *zb = W.z_bar.p;
            const double sg = W.sigma;
            each(s, n, [=] __device__(int64_t i) {
                const double z_raw = (double)xb[i] + sg * (-qx[i] + (double)atyb[i] - (double)c[i]);
                zb[i] = (VT)(((double)xb[i] - z_raw) / sg);
            });
        }
        if (m > 0) {
            csr_apply<double, double>(s, W.S.A, (const VT *)W.x_bar.p, EpiStore<double>{W.Ax_d.p});
            const double *ax = W.Ax_d.p;
            const VT *y = W.y.p;
            const MT *al = W.ALV, *au = W.AUV;
            VT *sv = W.s_.p;
            const double f1 = W.lambda_max_A * W.sigma;
            each(s, m, [=] __device__(int64_t i) {
                sv[i] = (VT)dmin(dmax(ax[i] - f1 * (double)y[i], (double)al[i]), (double)au[i]);
            });
        }
        HPRQP_HIP(hipStreamSynchronize(s));
    }
}

static int64_t print_step(int64_t k) {
    if (k <= 0) return 10;
    int64_t p = 1;
    while (p <= k / 10) p *= 10;
    return std::max<int64_t>(p / 10, 10);
}

// copy the FP32 state of a mixed solve into the FP64 solver (polish start)
template <class VT, class MT>
static void promote_state(Solver<VT, MT> &F, Solver<double> &D) {
    F.flush();
    D.invalidate_products();
    hipStream_t s = D.s;
    auto cp = [&](DevVec<double> &dst, const DevVec<VT> &src) {
        const VT *a = src.p;
        double *b = dst.p;
        each(s, src.n, [=] __device__(int64_t i) { b[i] = (double)a[i]; });
    };
    cp(D.w, F.w);
    cp(D.w_bar, F.w_bar);
    cp(D.dw, F.dw);
    cp(D.x, F.x);
    cp(D.x_hat, F.x_hat);
    cp(D.x_bar, F.x_bar);
    cp(D.dx, F.dx);
    cp(D.y, F.y);
    cp(D.y_bar, F.y_bar);
    cp(D.dy, F.dy);
    cp(D.z_bar, F.z_bar);
    cp(D.ATy, F.ATy);
    cp(D.ATy_bar, F.ATy_bar);
    cp(D.s_, F.s_);
    cp(D.Qw, F.Qw);
    cp(D.Qw_bar, F.Qw_bar);
    cp(D.tempv, F.tempv);
    auto cpd = [&](DevVec<double> &dst, const DevVec<double> &src) {
        if (src.n) HPRQP_HIP(hipMemcpyAsync(dst.p, src.p, sizeof(double) * src.n, hipMemcpyDeviceToDevice, s));
    };
    cpd(D.last_x, F.last_x);
    cpd(D.last_y, F.last_y);
    cpd(D.last_w, F.last_w);
    cpd(D.last_ATy, F.last_ATy);
    cpd(D.last_Qw, F.last_Qw);
    D.sigma = F.sigma;
    D.lambda_max_A = F.lambda_max_A;
    D.lambda_max_Q = F.lambda_max_Q;
    D.lambda = F.lambda;
    D.noC = F.noC;
    if (D.S.Q_is_diag) D.update_Q_factors();
    HPRQP_HIP(hipStreamSynchronize(s));
}

// The iteration loop of solve() (src/algorithm.jl:1977-2038).  Returns the
// status, or -1 when a mixed-precision run hands over to the FP64 polish.
struct LoopState {
    Restart r;
    int64_t iter = 0;
    int64_t iter_4 = 0, iter_6 = 0;
    double time_4 = 0.0, time_6 = 0.0;
    bool first_4 = true, first_6 = true;
    int64_t nh = 0;
    double best_kkt = INFINITY;
    int stall = 0;
    // auto-save (src/algorithm.jl:97-112): the best iterate so far by
    // max(err_Rp, err_Rd, rel_gap), scaled space, FP64
    DevVec<double> sx, sy, sz, sw;
    bool saved = false;
    int64_t save_iter = 0;
    double save_sigma = 0.0, save_Rp = 0.0, save_Rd = 0.0, save_pobj = 0.0, save_dobj = 0.0, save_gap = 0.0;
};

template <class T>
static void save_copy(hipStream_t s, const T *src, double *dst, int64_t len) {
    if (len > 0) each(s, len, [=] __device__(int64_t i) { dst[i] = (double)src[i]; });
}

// at each residual evaluation: keep the iterate if it is the best so far
template <class VT, class MT>
static void auto_save_track(Solver<VT, MT> &W, LoopState &L, int64_t iter, const Resid &res) {
    W.flush();
    if (L.saved && !(res.kkt < std::max(std::max(L.save_Rp, L.save_Rd), L.save_gap))) return;
    const int32_t n = W.n, m = W.m;
    if (!L.saved) {
        L.sx.alloc(n);
        L.sz.alloc(n);
        L.sw.alloc(n);
        L.sy.alloc(m);
    }
    save_copy(W.s, (const VT *)W.x_bar.p, L.sx.p, n);
    save_copy(W.s, (const VT *)W.z_bar.p, L.sz.p, n);
    save_copy(W.s, (const VT *)W.w_bar.p, L.sw.p, n);
    save_copy(W.s, (const VT *)W.y_bar.p, L.sy.p, m);
    L.saved = true;
    L.save_iter = iter;
    L.save_sigma = W.sigma;
    L.save_Rp = res.err_Rp;
    L.save_Rd = res.err_Rd;
    L.save_pobj = res.primal_obj;
    L.save_dobj = res.dual_obj;
    L.save_gap = res.rel_gap;
}

// v -> sc * v / nrm (the reference's save formulas), downloaded into out
template <class T>
static void save_download(hipStream_t s, const T *src, const double *nrm, int64_t len, double sc, double *tmp,
                          double *out) {
    if (len <= 0) return;
    each(s, len, [=] __device__(int64_t i) { tmp[i] = sc * ((double)src[i] / nrm[i]); });
    HPRQP_HIP(hipMemcpyAsync(out, tmp, sizeof(double) * len, hipMemcpyDeviceToHost, s));
    HPRQP_HIP(hipStreamSynchronize(s));
}

// at print steps and at termination: the current and the best iterate in
// the original space to the caller's callback (save_state_to_hdf5!,
// src/algorithm.jl:183-301)
template <class VT, class MT>
static void auto_save_write(Solver<VT, MT> &W, LoopState &L, int64_t iter, const Resid &res, double elapsed) {
    W.flush();
    const hprqp_params *p = W.p;
    if (!p->save_cb || !L.saved) return;
    hipStream_t s = W.s;
    const int32_t n = W.n, m = W.m;
    const double bs = W.S.b_scale, cs = W.S.c_scale;
    const double *cn = W.S.col_norm.p, *rn = W.S.row_norm.p;
    DevVec<double> t(std::max<int64_t>(1, std::max(n, m)));
    std::vector<double> cur(3 * (size_t)n + (size_t)m), best(3 * (size_t)n + (size_t)m);
    double *cx = cur.data(), *cz = cx + n, *cw = cz + n, *cy = cw + n;
    double *bx = best.data(), *bz = bx + n, *bw = bz + n, *by = bw + n;
    save_download(s, (const VT *)W.x_bar.p, cn, n, bs, t.p, cx);
    save_download(s, (const VT *)W.w_bar.p, cn, n, bs, t.p, cw);
    save_download(s, (const VT *)W.z_bar.p, cn, n, cs, t.p, cz);
    save_download(s, (const VT *)W.y_bar.p, rn, m, cs, t.p, cy);
    save_download(s, (const double *)L.sx.p, cn, n, bs, t.p, bx);
    save_download(s, (const double *)L.sw.p, cn, n, bs, t.p, bw);
    save_download(s, (const double *)L.sz.p, cn, n, cs, t.p, bz);
    save_download(s, (const double *)L.sy.p, rn, m, cs, t.p, by);
    hprqp_save_state st{};
    st.n = n;
    st.m = m;
    st.iteration = iter;
    st.time_elapsed = elapsed;
    st.sigma = W.sigma;
    st.err_Rp = res.err_Rp;
    st.err_Rd = res.err_Rd;
    st.primal_obj = res.primal_obj;
    st.dual_obj = res.dual_obj;
    st.rel_gap = res.rel_gap;
    st.kkt = res.kkt;
    st.x = cx;
    st.y = cy;
    st.z = cz;
    st.w = cw;
    st.best_iteration = L.save_iter;
    st.best_sigma = L.save_sigma;
    st.best_err_Rp = L.save_Rp;
    st.best_err_Rd = L.save_Rd;
    st.best_primal_obj = L.save_pobj;
    st.best_dual_obj = L.save_dobj;
    st.best_rel_gap = L.save_gap;
    st.best_x = bx;
    st.best_y = by;
    st.best_z = bz;
    st.best_w = bw;
    // a failed save is reported and the solve goes on (the reference catches
    // the failure of save_state_to_hdf5!, src/algorithm.jl:2005-2012)
    const int rc = p->save_cb(&st, p->save_user);
    if (rc != 0 && p->verbose)
        std::printf("Warning: auto-save callback failed (code %d) at iteration %lld\n", rc, (long long)iter);
}

// ---------------------------------------------------------------------
// FP32 bulk in offset form (HPRQP_PREC_FP32 on the standard path with a
// non-diagonal sparse Q): the bulk iterates are FP32 offsets from the FP64
// Halpern anchors, the row products act on offsets only (FP32 values and
// accumulation), and the anchors' products are FP64 vectors formed once per
// restart (see the EpiOff* epilogues).  D, an FP64 Solver, holds the
// anchors and the FP64 iterates the full steps write, and does the
// residuals, the M-norm, the sigma update and the restarts exactly as in
// the FP64 form.
// ---------------------------------------------------------------------

// G0 = -Q W0 + T0 - c
struct EpiAnchorG {
    double *G0;
    const double *T0, *c;
    template <class AT>
    __device__ void operator()(int32_t i, AT acc) const { G0[i] = -(double)acc + T0[i] - c[i]; }
};
// H0 = A X0 - fa Y0
struct EpiAnchorH {
    double *H0;
    const double *Y0;
    double fa;
    template <class AT>
    __device__ void operator()(int32_t i, AT acc) const { H0[i] = (double)acc - fa * Y0[i]; }
};

struct OffsetSolver {
    Solver<double> &D;
    hipStream_t s;
    const hprqp_params *p;
    Scaled &S;
    int32_t n, m;
    double &sigma;
    DevVec<float> dX, dXh, dW, dWb, dT, qd, dTv, dY, dYb, D0;
    DevVec<double> G0, H0;
    KernelTimer *kt = nullptr;

    explicit OffsetSolver(Solver<double> &d) : D(d), s(d.s), p(d.p), S(d.S), n(d.n), m(d.m), sigma(d.sigma) {}

    void alloc() {
        for (DevVec<float> *v : {&dX, &dXh, &dW, &dWb, &dT, &qd, &dTv}) v->alloc(n);
        for (DevVec<float> *v : {&dY, &dYb}) v->alloc(m);
        G0.alloc(n);
        D0.alloc(n);
        H0.alloc(m);
        ensure_v32(S.Q);
        ensure_v32(S.A);
        ensure_v32(S.AT);
    }

    // after the anchors (or sigma) change: the offsets from the new anchors
    // are zero, and the anchors' products are formed in FP64
    void reanchor() {
        auto zero = [&](DevVec<float> &v) {
            float *q = v.p;
            each(s, v.n, [=] __device__(int64_t i) { q[i] = 0.f; });
        };
        for (DevVec<float> *v : {&dX, &dW, &dT, &dWb, &dY}) zero(*v);
        csr_apply<double, double>(s, S.Q, (const double *)D.last_w.p, EpiAnchorG{G0.p, D.last_ATy.p, S.c.p});
        if (m > 0)
            csr_apply<double, double>(s, S.A, (const double *)D.last_x.p,
                                      EpiAnchorH{H0.p, D.last_y.p, D.lambda_max_A * sigma});
        float *d0 = D0.p;
        const double *x0 = D.last_x.p, *w0 = D.last_w.p;
        each(s, n, [=] __device__(int64_t i) { d0[i] = (float)(x0[i] - w0[i]); });
    }

    template <bool FULL>
    void steps(double h2) {
        const double f2s = 1.0 / (1.0 + sigma * D.lambda_max_Q), f1s = 1.0 - f2s;
        EpiOffZxw1<FULL> e;
        e.qd = qd.p;
        e.dW = dW.p;
        e.dT = dT.p;
        e.G0 = G0.p;
        e.X0 = D.last_x.p;
        e.D0 = D0.p;
        e.l = D.lE;
        e.u = D.uE;
        e.lS = D.lS;
        e.uS = D.uS;
        e.dX = dX.p;
        e.dXh = dXh.p;
        e.dWb = dWb.p;
        e.x_bar = D.x_bar.p;
        e.z_bar = D.z_bar.p;
        e.dx = D.dx.p;
        e.sigma = sigma;
        e.f1s = f1s;
        e.f2s = f2s;
        e.h2 = h2;
        const bool timed = kt && !FULL;
        if (timed) kt->begin();
        csr_apply<float, float>(s, S.Q, (const float *)dW.p, e);
        if (timed) kt->end();
        if (m > 0) {
            csr_apply<float, float>(s, S.Q, (const float *)dWb.p, EpiOffTempv{dXh.p, qd.p, dTv.p, sigma});
            EpiOffY<FULL> ey;
            ey.H0 = H0.p;
            ey.Y0 = D.last_y.p;
            ey.AL = S.AL.p;
            ey.AU = S.all_equal_rows ? nullptr : S.AU.p;
            ey.dY = dY.p;
            ey.dYb = dYb.p;
            ey.y_bar = D.y_bar.p;
            ey.s = D.s_.p;
            ey.dy = D.dy.p;
            ey.fact1 = D.lambda_max_A * sigma;
            ey.fact2 = 1.0 / ey.fact1;
            ey.h2 = h2;
            csr_apply<float, float>(s, S.A, (const float *)dTv.p, ey);
        }
        EpiOffW2<FULL> ew;
        ew.W0 = D.last_w.p;
        ew.T0 = D.last_ATy.p;
        ew.dW = dW.p;
        ew.dWb = dWb.p;
        ew.dT = dT.p;
        ew.w_bar = D.w_bar.p;
        ew.dw = D.dw.p;
        ew.ATy_bar = D.ATy_bar.p;
        ew.ATdy = D.ATdy.p;
        ew.fs = sigma / (1.0 + sigma * D.lambda_max_Q);
        ew.h2 = h2;
        csr_apply<float, float>(s, S.AT, (const float *)dYb.p, ew);
    }

    // main_update_gpu! (src/algorithm.jl:1250-1279) in offset form
    void main_update(int64_t inner, bool full) {
        const double h1 = 1.0 / ((double)inner + 2.0), h2 = 1.0 - h1;
        full ? steps<true>(h2) : steps<false>(h2);
    }
    void compute_residuals(int64_t iter, double obj_constant, bool lasso, Resid &r) {
        D.compute_residuals(iter, obj_constant, lasso, r);
    }
    bool async_ok() const { return D.async_ok(); }
    void residuals_start(int64_t iter) { D.residuals_start(iter); }
    void residuals_wait(int64_t iter, double obj_constant, bool lasso, Resid &r) {
        D.residuals_wait(iter, obj_constant, lasso, r);
    }
    void restore_snapshot() { D.restore_snapshot(); }
    double compute_M_norm() { return D.compute_M_norm(); }

    // hand the iteration to the FP64 form: its bulk vectors from the
    // anchors and offsets (the full-step vectors are FP64 already)
    void promote() {
        auto rebuild = [&](DevVec<double> &v, const DevVec<double> &a, const DevVec<float> &d) {
            double *o = v.p;
            const double *ap = a.p;
            const float *dp = d.p;
            each(s, v.n, [=] __device__(int64_t i) { o[i] = ap[i] + (double)dp[i]; });
        };
        rebuild(D.x, D.last_x, dX);
        rebuild(D.w, D.last_w, dW);
        rebuild(D.ATy, D.last_ATy, dT);
        if (m > 0) rebuild(D.y, D.last_y, dY);
        HPRQP_HIP(hipStreamSynchronize(s));
    }
};

static void update_sigma(OffsetSolver &O, Restart &r, double err_Rd, double err_Rp, double rel_gap) {
    update_sigma(O.D, r, err_Rd, err_Rp, rel_gap);
}
static void do_restart(OffsetSolver &O, Restart &r) {
    if (r.restart_flag <= 0) return;
    do_restart(O.D, r);
    O.reanchor();
}
static void auto_save_track(OffsetSolver &O, LoopState &L, int64_t iter, const Resid &res) {
    auto_save_track(O.D, L, iter, res);
}
static void auto_save_write(OffsetSolver &O, LoopState &L, int64_t iter, const Resid &res, double elapsed) {
    auto_save_write(O.D, L, iter, res, elapsed);
}

// W is a Solver<VT, MT> or the FP32 offset form (OffsetSolver).  mixed: a
// reduced-precision run that may hand over to the FP64 polish, on a stall
// and, with use_switch, once the KKT residual is within polish_switch of the
// tolerance.
template <class WS>
static int run_loop(WS &W, LoopState &L, int64_t iter_begin, double t_start, double obj_constant, bool lasso,
                    double *hist, int64_t hist_cap, bool mixed, Resid &res, bool has_res, bool use_switch = true) {
    const hprqp_params *p = W.p;
    Restart &r = L.r;
    int64_t pend = -1;  // a print step whose residuals are still on their way
    for (int64_t iter = iter_begin; iter <= p->max_iter; ++iter) {
        L.iter = iter;
        if (pend >= 0) {
            // the previous print step's residuals: its record, line and
            // termination test, as if they had been waited for there
            const int64_t k = pend;
            pend = -1;
            W.residuals_wait(k, obj_constant, lasso, res);
            if (hist && L.nh < hist_cap) {
                double *h = hist + L.nh * 8;
                h[0] = (double)k;
                h[1] = res.err_Rp;
                h[2] = res.err_Rd;
                h[3] = res.rel_gap;
                h[4] = res.primal_obj;
                h[5] = res.dual_obj;
                h[6] = W.sigma;
                h[7] = (double)r.restart_flag;
                L.nh++;
            }
            if (p->verbose)
                std::printf("%7lld  %3.2e  %3.2e  %+.8e  %+.8e  %3.2e  %3.2e  %7.2f\n", (long long)k, res.err_Rp,
                            res.err_Rd, res.primal_obj, res.dual_obj, res.rel_gap, W.sigma, now_s() - t_start);
            if (res.kkt < 1e-4 && L.first_4) {
                L.time_4 = now_s() - t_start;
                L.iter_4 = k;
                L.first_4 = false;
            }
            if (res.kkt < 1e-6 && L.first_6) {
                L.time_6 = now_s() - t_start;
                L.iter_6 = k;
                L.first_6 = false;
            }
            if (res.kkt < p->stoptol) {
                W.restore_snapshot();
                L.iter = k;
                return HPRQP_STATUS_OPTIMAL;
            }
        }
        const double tnow = now_s() - t_start;
        const int64_t ps = p->print_frequency == -1 ? print_step(iter) : p->print_frequency;
        const bool print_yes = (iter % ps == 0) || iter == p->max_iter || tnow > p->time_limit;
        const bool do_res = (iter % p->check_iter == 0) || print_yes || !has_res;
        // a print step between check iterations whose next step is partial:
        // its residuals run without a wait (only the termination test acts
        // on them, at the next iteration)
        const int64_t nx = iter + 1;
        const int64_t ps_nx = p->print_frequency == -1 ? print_step(nx) : p->print_frequency;
        const bool async = do_res && has_res && !p->auto_save && W.async_ok() && iter % p->check_iter != 0 &&
                           iter < p->max_iter && tnow <= p->time_limit && nx % p->check_iter != 0 &&
                           nx % ps_nx != 0;
        if (async) {
            W.residuals_start(iter);
            pend = iter;
        } else if (do_res) {
            W.compute_residuals(iter, obj_constant, lasso, res);
            has_res = true;
            if (p->auto_save) auto_save_track(W, L, iter, res);
            // reduced precision: hand over to the FP64 polish once the
            // FP64-measured KKT residual is within polish_switch of the
            // tolerance (the FP32 data / iterates floor lies near there), or
            // once the FP32 iteration stops improving it.  The offset form
            // (use_switch false) has no such floor: its error shrinks with
            // the distance to the anchors, and on slow grids the FP64
            // iteration itself runs 40 checks without halving the residual,
            // so its stall window is 400 checks
            if (mixed && res.kkt >= p->stoptol && (iter % p->check_iter == 0)) {
                if (use_switch && res.kkt < p->polish_switch * p->stoptol) {
                    L.iter = iter;
                    return -1;
                }
                if (res.kkt < 0.5 * L.best_kkt) {
                    L.best_kkt = res.kkt;
                    L.stall = 0;
                } else if (++L.stall >= (p->polish_stall_checks > 0 ? p->polish_stall_checks
                                         : !use_switch                    ? 400
                                         : (p->precision == HPRQP_PREC_MIXED ? 100 : 40)) &&
                           L.best_kkt < 1e-3) {
                    L.iter = iter;
                    return -1;
                }
            }
        }
        int st = -2;
        if (async) {
            // decided at the next iteration, from this step's residuals
        } else if (res.kkt < p->stoptol)
            st = HPRQP_STATUS_OPTIMAL;
        else if (iter == p->max_iter)
            st = HPRQP_STATUS_MAX_ITER;
        else if (now_s() - t_start > p->time_limit)
            st = HPRQP_STATUS_TIME_LIMIT;
        r.check(iter, p->check_iter, W.sigma);
        update_sigma(W, r, res.err_Rd, res.err_Rp, res.rel_gap);
        do_restart(W, r);
        if (do_res && !async && hist && L.nh < hist_cap) {
            double *h = hist + L.nh * 8;
            h[0] = (double)iter;
            h[1] = res.err_Rp;
            h[2] = res.err_Rd;
            h[3] = res.rel_gap;
            h[4] = res.primal_obj;
            h[5] = res.dual_obj;
            h[6] = W.sigma;
            h[7] = (double)r.restart_flag;
            L.nh++;
        }
        if (p->verbose && !async && (print_yes || st != -2))
            std::printf("%7lld  %3.2e  %3.2e  %+.8e  %+.8e  %3.2e  %3.2e  %7.2f\n", (long long)iter, res.err_Rp,
                        res.err_Rd, res.primal_obj, res.dual_obj, res.rel_gap, W.sigma, now_s() - t_start);
        if (p->auto_save && (print_yes || st != -2)) auto_save_write(W, L, iter, res, now_s() - t_start);
        if (!async && res.kkt < 1e-4 && L.first_4) {
            L.time_4 = now_s() - t_start;
            L.iter_4 = iter;
            L.first_4 = false;
        }
        if (!async && res.kkt < 1e-6 && L.first_6) {
            L.time_6 = now_s() - t_start;
            L.iter_6 = iter;
            L.first_6 = false;
        }
        if (st != -2) return st;
        const int64_t next = iter + 1;
        const int64_t ps2 = p->print_frequency == -1 ? print_step(next) : p->print_frequency;
        const bool to_check = (next % p->check_iter == 0) || (r.restart_flag > 0) || (next % ps2 == 0);
        // perform_iteration_step! (src/algorithm.jl:1881-1900)
        W.main_update(r.inner, to_check);
        if (r.restart_flag > 0) r.last_gap = W.compute_M_norm();
        if ((iter + 1) % p->check_iter == 0) r.current_gap = W.compute_M_norm();
        r.inner += 1;
    }
    return HPRQP_STATUS_MAX_ITER;
}

template <class VT, class MT>
static void collect(Solver<VT, MT> &W, double *x_out, double *y_out, double *z_out, double *w_out) {
    W.flush();
    hipStream_t s = W.s;
    const int32_t n = W.n, m = W.m;
    const double bs = W.S.b_scale, cs = W.S.c_scale;
    DevVec<double> t(std::max(n, m));
    auto dl = [&](const VT *src, const double *nrm, int64_t len, double sc, int mode, double *out) {
        if (!out || len == 0) return;
        double *tp = t.p;
        each(s, len, [=] __device__(int64_t i) {
            const double v = (double)src[i];
            tp[i] = mode == 0 ? sc * (v / nrm[i]) : sc * (v * nrm[i]);
        });
        HPRQP_HIP(hipMemcpyAsync(out, tp, sizeof(double) * len, hipMemcpyDeviceToHost, s));
        HPRQP_HIP(hipStreamSynchronize(s));
    };
    dl(W.x_bar.p, W.S.col_norm.p, n, bs, 0, x_out);
    dl(W.w_bar.p, W.S.col_norm.p, n, bs, 0, w_out);
    dl(W.z_bar.p, W.S.col_norm.p, n, cs, 1, z_out);
    dl(W.y_bar.p, W.S.row_norm.p, m, cs, 0, y_out);
}

template <class VT, class MT>
static void init_solver(Solver<VT, MT> &W, const hprqp_model *md) {
    W.alloc();
    W.lambda = md->lambda;
}

static void apply_lanes(hipStream_t s, const hprqp_model *md, const hprqp_params *p, Scaled &S, bool mixed);

// The row-product shapes of a LASSO operator's matrices live in the operator,
// which other solves and hprqp_qop_apply share: a solve sets them (forced or
// tuned) for its own duration, and this guard puts back what they were when
// the solve ends, however it ends.
struct OperatorLanesGuard {
    hprqp_qop *q;
    int la = 0, lat = 0;
    explicit OperatorLanesGuard(hprqp_qop *op) : q(op->kind == QOP_LASSO ? op : nullptr) {
        if (q) {
            la = q->la.lanes;
            lat = q->lat.lanes;
        }
    }
    ~OperatorLanesGuard() {
        if (q) {
            q->la.lanes = la;
            q->lat.lanes = lat;
        }
    }
};

// estimate_eigenvalues (src/algorithm.jl:1407-1440): lambda_max(A A^T) when
// there are constraints; lambda_max(Q) by the power method for a non-empty,
// non-diagonal Q or any other operator (the largest entry for a diagonal Q,
// 0 for an empty one); both scaled by eig_factor
template <class VT, class MT>
static void estimate_lambdas(Solver<VT, MT> &D, const hprqp_model *md, double eig_factor, int64_t *its_A,
                             int64_t *its_Q) {
    const Scaled &S = D.S;
    const bool need_A = md->m > 0;
    const bool need_Q = S.qkind != QOP_CSR || (!S.q_empty && !S.Q_is_diag);
    double lam_A = 0.0, lam_Q = 0.0;
    D.power_pair(need_A, need_Q, 5000, 1e-4, &lam_A, its_A, &lam_Q, its_Q);
    D.lambda_max_A = need_A ? lam_A * eig_factor : 0.0;
    if (need_Q)
        D.lambda_max_Q = lam_Q * eig_factor;
    else
        D.lambda_max_Q = S.qkind == QOP_CSR && !S.q_empty ? S.q_max_val : 0.0;
}

void solve(hprqp_ctx *ctx, const hprqp_model *md, const hprqp_params *p, const double *x0, const double *y0,
           double *x_out, double *y_out, double *z_out, double *w_out, double *hist, int64_t hist_cap,
           hprqp_result *res) {
    std::memset(res, 0, sizeof(*res));
    OperatorLanesGuard lanes_guard(md->Q);
    HPRQP_REQUIRE(p->check_iter > 0, HPRQP_E_ARG, "check_iter must be positive");
    HPRQP_REQUIRE(p->print_frequency == -1 || p->print_frequency > 0, HPRQP_E_ARG,
                  "print_frequency must be -1 or positive");
    HPRQP_REQUIRE(p->sigma == -1 || p->sigma > 0, HPRQP_E_ARG, "sigma must be positive or -1");
    HPRQP_REQUIRE(p->precision == HPRQP_PREC_FP64 || p->precision == HPRQP_PREC_MIXED ||
                      p->precision == HPRQP_PREC_FP32,
                  HPRQP_E_ARG, "unknown precision mode");
    const double setup_start = now_s();
    Reducer red(ctx);
    Scaled S;
    build_scaled(ctx, md, p, red, S);
    check_Q_diagonal(red, S);
    const bool mixed = p->precision != HPRQP_PREC_FP64;
    apply_lanes(ctx->stream, md, p, S, mixed);
    const bool lasso = p->problem_type == HPRQP_PROBLEM_LASSO;
    std::unique_ptr<KernelTimer> ktimer;
    if (p->time_kernels) ktimer.reset(new KernelTimer(ctx->stream));
    Solver<double> D(ctx, S, red, p);
    init_solver(D, md);
    D.kt = ktimer.get();
    // sigma_0 (src/algorithm.jl:839-851)
    double sigma0;
    if (p->sigma == -1) {
        if (S.norm_c > 1e-16 && S.norm_b > 1e-16 && S.norm_b < 1e16 && S.norm_c < 1e16)
            sigma0 = S.norm_b / S.norm_c;
        else
            sigma0 = 1.0;
    } else {
        sigma0 = p->sigma;
    }
    D.sigma = sigma0;
    HPRQP_HIP(hipStreamSynchronize(ctx->stream));
    res->setup_time = now_s() - setup_start;

    const double t_start = now_s();
    // estimate_eigenvalues (src/algorithm.jl:1407-1440)
    estimate_lambdas(D, md, p->eig_factor, &res->power_iter_A, &res->power_iter_Q);
    res->power_time = now_s() - t_start;
    // noC decision uses the unscaled bounds (src/algorithm.jl:1969-1974)
    int64_t nfree = 0;
    for (int32_t i = 0; i < md->n; ++i)
        if (md->l[i] == -INFINITY && md->u[i] == INFINITY) nfree++;
    const bool noC = (double)nfree > 0.8 * (double)md->n;
    D.noC = noC;

    LoopState L;
    L.r.best_sigma = sigma0;
    Solver<double>::Resid R;
    int status;
    int64_t polish_from = -1;
    // FP32 on the standard path with a non-diagonal sparse Q: the offset form
    const bool offset = p->precision == HPRQP_PREC_FP32 && S.qkind == QOP_CSR && !S.q_empty && !noC &&
                        !S.Q_is_diag && lasso == false;
    bool from_D = !mixed;
    if (!mixed) {
        process_initial_points(D, x0, y0);
        if (S.Q_is_diag) D.update_Q_factors();
        status = run_loop(D, L, 0, t_start, md->obj_constant, lasso, hist, hist_cap, false, R, false);
    } else if (offset) {
        process_initial_points(D, x0, y0);
        OffsetSolver O(D);
        O.alloc();
        O.kt = ktimer.get();
        O.reanchor();
        // the offsets carry no fixed FP32 floor: the FP64 polish takes over
        // only if the KKT residual stalls, not near the tolerance
        status = run_loop(O, L, 0, t_start, md->obj_constant, lasso, hist, hist_cap, true, R, false, false);
        if (status == -1) {
            O.promote();
            polish_from = L.iter;
            status = run_loop(D, L, L.iter, t_start, md->obj_constant, lasso, hist, hist_cap, false, R, false);
        }
        from_D = true;
    } else {
        // reduced-precision run, then (only if it stalls) the FP64 polish
        // continuing the same iteration from its state
        auto lowp = [&](auto &F) {
            init_solver(F, md);
            F.kt = ktimer.get();
            F.sigma = sigma0;
            F.lambda_max_A = D.lambda_max_A;
            F.lambda_max_Q = D.lambda_max_Q;
            F.noC = noC;
            process_initial_points(F, x0, y0);
            if (S.Q_is_diag) F.update_Q_factors();
            typename std::remove_reference<decltype(F)>::type::Resid RF;
            status = run_loop(F, L, 0, t_start, md->obj_constant, lasso, hist, hist_cap, true, RF, false);
            if (status == -1) {
                promote_state(F, D);
                polish_from = L.iter;
                status = run_loop(D, L, L.iter, t_start, md->obj_constant, lasso, hist, hist_cap, false, R, false);
            } else {
                R.err_Rp = RF.err_Rp;
                R.err_Rd = RF.err_Rd;
                R.kkt = RF.kkt;
                R.primal_obj = RF.primal_obj;
                R.rel_gap = RF.rel_gap;
                R.dual_obj = RF.dual_obj;
                collect(F, x_out, y_out, z_out, w_out);
                res->sigma = F.sigma;
            }
        };
        // LASSO takes the FP32 data with FP64 iterates in both reduced
        // precisions: its sigma climbs to 1e4-1e5, where the FP32 rounding of
        // the prox argument (sigma * eps * |Qw + c|) exceeds the tolerance
        // and an FP32 iteration drifts away from the solution
        if (p->precision == HPRQP_PREC_MIXED || S.qkind == QOP_LASSO) {
            Solver<double, float> F(ctx, S, red, p);
            lowp(F);
        } else {
            Solver<float, float> F(ctx, S, red, p);
            lowp(F);
        }
    }
    HPRQP_HIP(hipStreamSynchronize(ctx->stream));
    res->time = now_s() - t_start;
    if (from_D || polish_from >= 0) {
        collect(D, x_out, y_out, z_out, w_out);
        res->sigma = D.sigma;
    }
    res->status = status;
    res->iter = L.iter;
    res->iter_4 = L.iter_4 == 0 ? L.iter : L.iter_4;
    res->iter_6 = L.iter_6 == 0 ? L.iter : L.iter_6;
    res->time_4 = L.time_4 == 0.0 ? res->time : L.time_4;
    res->time_6 = L.time_6 == 0.0 ? res->time : L.time_6;
    res->primal_obj = R.primal_obj;
    res->dual_obj = R.dual_obj;
    res->residuals = R.kkt;
    res->gap = R.rel_gap;
    res->err_Rp = R.err_Rp;
    res->err_Rd = R.err_Rd;
    res->lambda_max_A = D.lambda_max_A;
    res->lambda_max_Q = D.lambda_max_Q;
    res->sigma0 = sigma0;
    res->restarts = (int32_t)L.r.times;
    res->Q_is_diag = S.Q_is_diag;
    res->noC = noC;
    res->n_history = L.nh;
    res->polish_iter = polish_from >= 0 ? L.iter - polish_from : 0;
    // value-table shapes reported as -1000 - rows per block
    auto shape = [](const DevCsr &a) {
        return a.dia_nd > 0 ? kLanesDiagonal : a.dict_rb > 0 ? -1000 - a.dict_rb : a.lanes;
    };
    res->lanes_Q = shape(S.Q);
    res->lanes_A = shape(S.A);
    res->lanes_AT = shape(S.AT);
    if (S.qkind == QOP_LASSO) {
        res->lanes_Q = S.qop->la.lanes;  // the LASSO operator's A (its A^T in lanes_AT)
        res->lanes_AT = S.qop->lat.lanes;
    }
    if (S.qkind == QOP_QAP) res->qap_family = qap_family_of(S.qop, p->precision != HPRQP_PREC_FP64);
    if (ktimer) {
        ktimer->flush();
        res->kernel_calls = ktimer->count;
        res->kernel_ms = ktimer->count ? ktimer->total_ms / (double)ktimer->count : 0.0;
    }
}

// ---------------------------------------------------------------------
// measurement and autotuning
// ---------------------------------------------------------------------
struct EventTimer {
    hipEvent_t a, b;
    hipStream_t s;
    explicit EventTimer(hipStream_t st) : s(st) {
        HPRQP_HIP(hipEventCreateWithFlags(&a, hipEventDisableSystemFence));
        HPRQP_HIP(hipEventCreateWithFlags(&b, hipEventDisableSystemFence));
    }
    ~EventTimer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    template <class F>
    double avg_ms(int reps, F &&f) {
        f();  // warm
        HPRQP_HIP(hipEventRecord(a, s));
        for (int r = 0; r < reps; ++r) f();
        HPRQP_HIP(hipEventRecord(b, s));
        HPRQP_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        HPRQP_HIP(hipEventElapsedTime(&ms, a, b));
        return (double)ms / reps;
    }
};

// Epilogue the autotuner times the shapes with: like the iteration's fused
// half-steps it reads four row vectors and writes two, so a shape whose
// epilogue runs on one lane of each row group (its vector accesses poorly
// coalesced) pays for that as it does in the solve
template <class VT>
struct EpiTune {
    VT *y, *z;
    const VT *a, *b, *c, *d;
    template <class AT>
    __device__ void operator()(int32_t i, AT acc) const {
        const double v = (double)acc;
        y[i] = (VT)(v + (double)a[i] * (double)b[i]);
        z[i] = (VT)(v * (double)c[i] - (double)d[i]);
    }
};

// Rows-per-group autotuner for one matrix: times the product for every
// lane count (replaces determine_spmv_mode's cuSPARSE-vs-custom choice), with
// the iteration's own types: matrix values MT, vectors and accumulation VT
// (FP64 / FP64 plain, FP32 / FP64 mixed, FP32 / FP32 fp32), and an epilogue
// with the row-vector traffic of the fused half-steps (EpiTune).
template <class MT, class VT>
static int tune_lanes(hipStream_t s, const DevCsr &a, int reps, bool verbose) {
    const int64_t nr = std::max<int64_t>(1, a.nrow);
    DevVec<VT> x(std::max<int64_t>(1, a.ncol)), y(nr), z(nr), rv(4 * nr);
    {
        VT *xp = x.p, *rp = rv.p;
        each(s, a.ncol, [=] __device__(int64_t i) { xp[i] = (VT)1; });
        each(s, 4 * nr, [=] __device__(int64_t i) { rp[i] = (VT)0.5; });
    }
    const EpiTune<VT> epi{y.p, z.p, rv.p, rv.p + nr, rv.p + 2 * nr, rv.p + 3 * nr};
    EventTimer tm(s);
    const double mean = a.nrow > 0 ? (double)a.nnz / (double)a.nrow : 0.0;
    std::vector<int> cand;
    for (int rb : {256, 128, 64, 32, 16})  // CSR-stream when a block's nonzeros fit in LDS
        if (mean * rb <= 0.75 * kStreamCap) cand.push_back(-rb);
    for (int L = 1; L <= 64; L *= 2) cand.push_back(L);
    if (mean >= 64.0) cand.push_back(kBlock);  // one block per row
    // long rows over many columns: x staged in LDS chunk by chunk
    if (mean >= 256.0 && a.ncol >= 4 * (int64_t)kChunkW) cand.push_back(kLanesChunked);
    auto time_of = [&](int L) {
        return tm.avg_ms(reps, [&] {
            if (L == kLanesChunked)
                launch_chunked<VT>(s, a, csr_vals<MT>(a), (const VT *)x.p, epi);
            else
                launch_csr<VT>(s, L, a.nrow, a.rp.p, a.ci.p, csr_vals<MT>(a), (const VT *)x.p, epi);
        });
    };
    // the heuristic shape stays unless another is clearly faster (10 %), so
    // shapes within timing noise of each other do not flip from run to run
    int best = a.lanes;
    double best_ms = time_of(best);
    if (verbose)
        std::printf("row-product autotune %d x %d (%lld nnz): shape %d %.4f ms (heuristic)\n", a.nrow, a.ncol,
                    (long long)a.nnz, best, best_ms);
    for (int L : cand) {
        if (L == a.lanes) continue;
        const double ms = time_of(L);
        if (verbose) std::printf("row-product autotune   shape %d %.4f ms\n", L, ms);
        if (ms < best_ms * 0.90) {
            best_ms = ms;
            best = L;
        }
    }
    return best;
}

// The value-table form against the matrix's tuned shape (FP64): CSR-stream
// blocks of 256, 128, 64 and 32 rows; taken when 5 % faster.  Returns the rows
// per block, or -1.
static int tune_dict(hipStream_t s, const DevCsr &a, int reps, bool verbose) {
    const int64_t nr = std::max<int64_t>(1, a.nrow);
    DevVec<double> x(std::max<int64_t>(1, a.ncol)), y(nr), z(nr), rv(4 * nr);
    {
        double *xp = x.p, *rp = rv.p;
        each(s, a.ncol, [=] __device__(int64_t i) { xp[i] = 1.0; });
        each(s, 4 * nr, [=] __device__(int64_t i) { rp[i] = 0.5; });
    }
    const EpiTune<double> epi{y.p, z.p, rv.p, rv.p + nr, rv.p + 2 * nr, rv.p + 3 * nr};
    EventTimer tm(s);
    const double plain = tm.avg_ms(reps, [&] { csr_apply<double, double>(s, a, (const double *)x.p, epi); });
    const double mean = a.nrow > 0 ? (double)a.nnz / (double)a.nrow : 0.0;
    int best = -1;
    double best_ms = plain * 0.95;
    for (int rb : {256, 128, 64, 32}) {
        if (mean * rb > 0.75 * kStreamCap) continue;
        const double ms = tm.avg_ms(reps, [&] { launch_dict<double>(s, a, rb, (const double *)x.p, epi); });
        if (verbose) std::printf("row-product autotune   value table, %d rows per block %.4f ms (shape %d %.4f ms)\n",
                                 rb, ms, a.lanes, plain);
        if (ms < best_ms) {
            best_ms = ms;
            best = rb;
        }
    }
    return best;
}

// The diagonals of a pattern, entries (i, j) with j = floor(i * ncol / nrow)
// + off: their count (offsets ascending in off), or -1 when there are more
// than kDiaMax or the diagonal arrays would hold more than 1.5 entries per
// nonzero (padding).  transposed: the pattern of h's transpose.
static int dia_detect(const HostCsr &h, bool transposed, int32_t *off) {
    const int32_t nrow = transposed ? h.ncol : h.nrow, ncol = transposed ? h.nrow : h.ncol;
    if (nrow <= 0 || ncol <= 0 || h.rp.empty() || h.rp[h.nrow] <= 0) return -1;
    int nd = 0;
    for (int32_t r = 0; r < h.nrow; ++r)
        for (int32_t k = h.rp[r]; k < h.rp[r + 1]; ++k) {
            const int32_t i = transposed ? h.ci[k] : r, j = transposed ? r : h.ci[k];
            const int32_t o = j - dia_base(i, nrow, ncol);
            int d = 0;
            while (d < nd && off[d] != o) ++d;
            if (d == nd) {
                if (nd == kDiaMax) return -1;
                off[nd++] = o;
            }
        }
    std::sort(off, off + nd);
    if ((double)nd * (double)nrow > 1.5 * (double)h.rp[h.nrow]) return -1;
    return nd;
}

// Fills a's diagonal arrays from its CSR values (FP64, and FP32 when f32: the
// FP32 values are the FP64 ones rounded, as the CSR form's v32)
static void build_dia(hipStream_t s, DevCsr &a, int nd, const int32_t *offs, bool f32) {
    const int32_t nrow = a.nrow, ncol = a.ncol;
    const int64_t len = (int64_t)nd * nrow;
    DiaOff off;
    for (int d = 0; d < kDiaMax; ++d) off.o[d] = d < nd ? offs[d] : 0;
    for (int d = 0; d < kDiaMax; ++d) a.dia_off[d] = off.o[d];
    if (a.dia_v.n != len) {
        a.dia_v.alloc(len);
        HPRQP_HIP(hipMemsetAsync(a.dia_v.p, 0, sizeof(double) * len, s));
        const int32_t *rp = a.rp.p, *ci = a.ci.p;
        const double *v = a.v.p;
        double *dv = a.dia_v.p;
        each(s, nrow, [=] __device__(int64_t i) {
            const int32_t base = dia_base((int32_t)i, nrow, ncol);
            for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
                const int32_t o = ci[k] - base;
                int d = 0;
                while (d < nd - 1 && off.o[d] != o) ++d;
                dv[(int64_t)d * nrow + i] = v[k];
            }
        });
    }
    if (f32 && a.dia_v32.n != len) {
        a.dia_v32.alloc(len);
        const double *dv = a.dia_v.p;
        float *d32 = a.dia_v32.p;
        each(s, len, [=] __device__(int64_t k) { d32[k] = (float)dv[k]; });
    }
    HPRQP_HIP(hipStreamSynchronize(s));
}

// A matrix's diagonal form against the shape chosen for its CSR form, in the
// iteration's precision (VT vectors, MT values); taken when 5 % faster
template <class VT, class MT>
static bool tune_dia(hipStream_t s, DevCsr &a, int nd, int reps, bool verbose) {
    const int64_t nr = std::max<int64_t>(1, a.nrow);
    DevVec<VT> x(std::max<int64_t>(1, a.ncol)), y(nr), z(nr), rv(4 * nr);
    {
        VT *xp = x.p, *rp = rv.p;
        each(s, a.ncol, [=] __device__(int64_t i) { xp[i] = (VT)1; });
        each(s, 4 * nr, [=] __device__(int64_t i) { rp[i] = (VT)0.5; });
    }
    const EpiTune<VT> epi{y.p, z.p, rv.p, rv.p + nr, rv.p + 2 * nr, rv.p + 3 * nr};
    EventTimer tm(s);
    a.dia_nd = 0;
    const double csr = tm.avg_ms(reps, [&] { csr_apply<VT, MT>(s, a, (const VT *)x.p, epi); });
    a.dia_nd = nd;
    const double dia = tm.avg_ms(reps, [&] { csr_apply<VT, MT>(s, a, (const VT *)x.p, epi); });
    a.dia_nd = 0;
    if (verbose)
        std::printf("row-product autotune %d x %d: diagonal form (%d diagonals) %.4f ms against %.4f ms\n", a.nrow,
                    a.ncol, nd, dia, csr);
    return dia < 0.95 * csr;
}

// Q, A and A^T in the diagonal form: forced (spmv_lanes = kLanesDiagonal:
// Q must have one; A and A^T take theirs where they have one), or where a
// matrix's pattern has at most kDiaMax diagonals and the autotuner times
// that form faster.  The reduced precisions keep FP64 diagonals too (the
// polish and the residuals run in FP64).
static void apply_dia(hipStream_t s, const hprqp_model *md, const hprqp_params *p, Scaled &S, bool forced) {
    const bool q_csr = S.qkind == QOP_CSR && !S.q_empty;
    HPRQP_REQUIRE(!forced || q_csr, HPRQP_E_ARG, "spmv_lanes = -2000 needs a sparse matrix Q");
    DevCsr *mats[3] = {&S.Q, &S.A, &S.AT};
    const bool f32 = p->precision != HPRQP_PREC_FP64;
    for (int k = 0; k < 3; ++k) {
        DevCsr &a = *mats[k];
        if ((k == 0 && !q_csr) || (k > 0 && md->m == 0) || a.nnz == 0) continue;
        if (md->dia_nd[k] == 0)
            md->dia_nd[k] = k == 0 ? dia_detect(md->Qsym, false, md->dia_off[k])
                                   : dia_detect(md->A, k == 2, md->dia_off[k]);
        const int nd = md->dia_nd[k];
        if (nd <= 0) {
            HPRQP_REQUIRE(!(forced && k == 0), HPRQP_E_ARG,
                          "spmv_lanes = -2000: Q has more than 8 diagonals or more padding than 1.5 entries per "
                          "nonzero");
            continue;
        }
        if (forced) {
            build_dia(s, a, nd, md->dia_off[k], f32);
            a.dia_nd = nd;
            continue;
        }
        if (a.nnz < (1 << 18)) continue;
        int &td = md->tuned_dia[p->precision][k];
        if (td == 0) {
            build_dia(s, a, nd, md->dia_off[k], f32);
            bool take;
            if (p->precision == HPRQP_PREC_MIXED)
                take = tune_dia<double, float>(s, a, nd, 20, p->verbose);
            else if (p->precision == HPRQP_PREC_FP32)
                take = tune_dia<float, float>(s, a, nd, 20, p->verbose);
            else
                take = tune_dia<double, double>(s, a, nd, 20, p->verbose);
            td = take ? 1 : -1;
        }
        if (td > 0) {
            build_dia(s, a, nd, md->dia_off[k], f32);
            a.dia_nd = nd;
        } else {
            a.dia_v.release();
            a.dia_v32.release();
        }
    }
}

// the QAP operator's pass family: the caller's, or the one timed fastest at
// the operator's first solve in that precision (the reduced precisions run
// the split-BF16 passes and the FP64 ones in the polish and residuals)
static void apply_qap_family(hipStream_t s, const hprqp_params *p, Scaled &S) {
    const int f = p->qap_family;
    HPRQP_REQUIRE(f == 0 || f == 2 || f == 3 || f == 4, HPRQP_E_ARG, "qap_family must be 0, 2, 3 or 4");
    if (S.qkind != QOP_QAP) return;
    hprqp_qop *q = S.qop;
    const bool reduced = p->precision != HPRQP_PREC_FP64;
    if (f != 0) {
        q->qfam64 = q->qfam16 = f;
        return;
    }
    if (q->qtuned64 == 0) q->qtuned64 = qap_tune(s, q, false, p->verbose);
    q->qfam64 = q->qtuned64;
    if (reduced) {
        if (q->qtuned16 == 0) q->qtuned16 = qap_tune(s, q, true, p->verbose);
        q->qfam16 = q->qtuned16;
    }
}

static void apply_lanes(hipStream_t s, const hprqp_model *md, const hprqp_params *p, Scaled &S, bool mixed) {
    HPRQP_REQUIRE(p->precision == HPRQP_PREC_FP64 || p->precision == HPRQP_PREC_MIXED ||
                      p->precision == HPRQP_PREC_FP32,
                  HPRQP_E_ARG, "unknown precision mode");
    (void)mixed;
    apply_qap_family(s, p, S);
    if (p->spmv_lanes == kLanesDiagonal) {
        // Q in the diagonal form, and A and A^T where their patterns have
        // one; the others keep their default shapes
        apply_dia(s, md, p, S, true);
        return;
    }
    if (p->spmv_lanes <= -1000) {
        // the value-table form forced (rows per block -spmv_lanes - 1000)
        const int rb = -p->spmv_lanes - 1000;
        HPRQP_REQUIRE(rb == 32 || rb == 64 || rb == 128 || rb == 256, HPRQP_E_ARG,
                      "spmv_lanes below -1000 selects the value-table form: -1032, -1064, -1128 or -1256");
        for (DevCsr *a : {&S.Q, &S.A, &S.AT})
            if (a->ndict > 0) a->dict_rb = rb;
        return;
    }
    if (p->spmv_lanes != 0) {
        const int L = p->spmv_lanes;
        const bool vec = L > 0 && (L & (L - 1)) == 0 && (L <= 64 || L == kBlock);
        const bool stream = L == -64 || L == -128 || L == -256;
        HPRQP_REQUIRE(vec || stream || L == kLanesChunked, HPRQP_E_ARG,
                      "spmv_lanes must be 0, a power of two <= 64, 256 (a block per row), -64/-128/-256 "
                      "(rows per block) or 1048576 (column-chunked)");
        S.Q.lanes = S.A.lanes = S.AT.lanes = p->spmv_lanes;
        if (S.qkind == QOP_LASSO) S.qop->la.lanes = S.qop->lat.lanes = p->spmv_lanes;
        return;
    }
    // Q, A, A^T of the scaled problem, and the LASSO operator's data matrix
    // and its transpose
    DevCsr *mats[5] = {&S.Q, &S.A, &S.AT, nullptr, nullptr};
    if (S.qkind == QOP_LASSO) {
        mats[3] = &S.qop->la;
        mats[4] = &S.qop->lat;
    }
    for (int k = 0; k < 5; ++k) {
        if (!mats[k]) continue;
        DevCsr &a = *mats[k];
        if (a.nnz < (1 << 18)) continue;  // heuristic choice is already close
        int &tuned = md->tuned_lanes[p->precision][k];
        if (tuned == 0) {
            if (p->precision == HPRQP_PREC_MIXED) {
                ensure_v32(a);
                tuned = tune_lanes<float, double>(s, a, 20, p->verbose);
            } else if (p->precision == HPRQP_PREC_FP32) {
                ensure_v32(a);
                tuned = tune_lanes<float, float>(s, a, 20, p->verbose);
            } else {
                tuned = tune_lanes<double, double>(s, a, 20, p->verbose);
            }
        }
        a.lanes = tuned;
    }
    // the FP64 iteration may take Q, A, A^T in value-table form
    if (p->precision == HPRQP_PREC_FP64) {
        DevCsr *dm[3] = {&S.Q, &S.A, &S.AT};
        for (int k = 0; k < 3; ++k) {
            DevCsr &a = *dm[k];
            if (a.ndict == 0 || a.nnz < (1 << 18)) continue;
            int &td = md->tuned_dict[k];
            if (td == 0) td = tune_dict(s, a, 20, p->verbose);
            a.dict_rb = td > 0 ? td : 0;
        }
    }
    apply_dia(s, md, p, S, false);
}

template <class VT, class MT>
static void profile_impl(Solver<VT, MT> &W, int reps, double *ms, int *nout) {
    EventTimer tm(W.s);
    const bool standard = W.S.qkind != QOP_LASSO && !W.S.q_empty && !W.noC;
    for (int k = 0; k < 6; ++k) ms[k] = 0.0;
    if (standard) {
        const double h1 = 0.5, h2 = 0.5;
        ms[0] = tm.avg_ms(reps, [&] { W.template step_zxw1<false, false>(h1, h2); });
        if (W.m > 0) {
            ms[1] = tm.avg_ms(reps, [&] { W.step_tempv(); });
            ms[2] = tm.avg_ms(reps, [&] { W.template step_y<false>(W.tempv.p, h1, h2); });
        }
        ms[3] = tm.avg_ms(reps, [&] { W.template step_w2<false, false>(h1, h2); });
    }
    int64_t inner = 0;
    ms[4] = tm.avg_ms(reps, [&] { W.main_update(inner++, false); });
    typename Solver<VT, MT>::Resid r;
    ms[5] = tm.avg_ms(std::max(1, reps / 10), [&] { W.compute_residuals(1, 0.0, false, r); });
    if (nout) *nout = 6;
}

void profile_step(hprqp_ctx *ctx, const hprqp_model *md, const hprqp_params *p, int reps, double *ms_out,
                  int *n_out) {
    OperatorLanesGuard lanes_guard(md->Q);
    Reducer red(ctx);
    Scaled S;
    build_scaled(ctx, md, p, red, S);
    check_Q_diagonal(red, S);
    const bool mixed = p->precision != HPRQP_PREC_FP64;
    apply_lanes(ctx->stream, md, p, S, mixed);
    int64_t nfree = 0;
    for (int32_t i = 0; i < md->n; ++i)
        if (md->l[i] == -INFINITY && md->u[i] == INFINITY) nfree++;
    const bool noC = (double)nfree > 0.8 * (double)md->n;
    auto setup = [&](auto &W) {
        init_solver(W, md);
        W.sigma = 1.0;
        W.lambda_max_A = 1.0;
        W.lambda_max_Q = 1.0;
        W.noC = noC;
        if (S.Q_is_diag) W.update_Q_factors();
    };
    if (p->precision == HPRQP_PREC_MIXED) {
        Solver<double, float> F(ctx, S, red, p);
        setup(F);
        profile_impl(F, reps, ms_out, n_out);
    } else if (p->precision == HPRQP_PREC_FP32) {
        Solver<float, float> F(ctx, S, red, p);
        setup(F);
        profile_impl(F, reps, ms_out, n_out);
    } else {
        Solver<double> D(ctx, S, red, p);
        setup(D);
        profile_impl(D, reps, ms_out, n_out);
    }
    HPRQP_HIP(hipStreamSynchronize(ctx->stream));
}

// ---------------------------------------------------------------------
// ABI helpers
// ---------------------------------------------------------------------
void csr_spmv_fp64(hprqp_ctx *ctx, const DevCsr &a, const double *x, double *y) {
    csr_apply<double, double>(ctx->stream, a, x, EpiStore<double>{y});
}

// the temporaries of an operator's FP64 apply: LASSO A x (rows), QAP A X
// and S X (2 n^2), low rank U^T x (k); CSR and user operators none
int64_t qop_scratch_doubles(const hprqp_qop *q) {
    switch (q->kind) {
        case QOP_LASSO: return q->la.nrow;
        case QOP_QAP: return 2 * (int64_t)q->nq * q->nq;
        case QOP_LOWRANK: return q->lr_k;
        default: return 0;
    }
}

// a product into caller-given scratch must not build the chunked shape's
// segment and partial-sum buffers: that shape gives way to a block per row
static void csr_apply_noalloc(hipStream_t s, const DevCsr &a, const double *x, double *y) {
    if (a.lanes == kLanesChunked && a.nchunk == 0)
        launch_csr<double>(s, kBlock, a.nrow, a.rp.p, a.ci.p, a.v.p, x, EpiStore<double>{y});
    else
        csr_apply<double, double>(s, a, x, EpiStore<double>{y});
}

void qop_apply_fp64(hprqp_ctx *ctx, hprqp_qop *q, const double *x, double *Qx, double *scratch) {
    hipStream_t s = ctx->stream;
    switch (q->kind) {
        case QOP_CSR:
            if (q->q.nnz == 0) {
                // a fill kernel rather than a memset node: graph-capture safe
                each(s, q->n, [=] __device__(int64_t i) { Qx[i] = 0.0; });
            } else if (scratch) {
                csr_apply_noalloc(s, q->q, x, Qx);
            } else {
                csr_apply<double, double>(s, q->q, x, EpiStore<double>{Qx});
            }
            break;
        case QOP_LASSO: {
            double *t = scratch ? scratch : q->ltemp.p;
            if (scratch) {
                csr_apply_noalloc(s, q->la, x, t);
                csr_apply_noalloc(s, q->lat, (const double *)t, Qx);
            } else {
                csr_apply<double, double>(s, q->la, x, EpiStore<double>{t});
                csr_apply<double, double>(s, q->lat, (const double *)t, EpiStore<double>{Qx});
            }
            break;
        }
        case QOP_QAP:
            qap_apply_f64(s, q, x, Qx, scratch);
            break;
        case QOP_LOWRANK:
            lowrank_apply<double, double, double>(s, q->n, q->lr_k, q->lr_U.p, q->lr_d.p, x,
                                                  scratch ? scratch : q->lr_t.p, Qx);
            break;
        default: {
            int rc = q->vt.apply(q->user, x, Qx, (void *)s);
            HPRQP_REQUIRE(rc == 0, HPRQP_E_ARG, "user Q operator apply failed");
        }
    }
}

void qop_apply_mixed(hprqp_ctx *ctx, hprqp_qop *q, const double *x, double *Qx) {
    hipStream_t s = ctx->stream;
    switch (q->kind) {
        case QOP_CSR:
            if (q->q.nnz == 0) {
                HPRQP_HIP(hipMemsetAsync(Qx, 0, sizeof(double) * q->n, s));
            } else {
                ensure_v32(q->q);
                csr_apply<double, float>(s, q->q, x, EpiStore<double>{Qx});
            }
            break;
        case QOP_LASSO:
            ensure_v32(q->la);
            ensure_v32(q->lat);
            csr_apply<double, float>(s, q->la, x, EpiStore<double>{q->ltemp.p});
            csr_apply<double, float>(s, q->lat, (const double *)q->ltemp.p, EpiStore<double>{Qx});
            break;
        case QOP_QAP:
            qap_apply_bf16x3<double, double>(s, q, x, Qx);
            break;
        case QOP_LOWRANK:
            lowrank_apply<float, double, double>(s, q->n, q->lr_k, q->lr_U32.p, q->lr_d32.p, x, q->lr_t.p, Qx);
            break;
        default: {
            int rc = q->vt.apply(q->user, x, Qx, (void *)s);
            HPRQP_REQUIRE(rc == 0, HPRQP_E_ARG, "user Q operator apply failed");
        }
    }
}

}  // namespace hprqp

// ---------------------------------------------------------------------
// step-level workspace: the state of solve() before its loop, with each
// device operation of the loop as its own call (include/hprqp.h)
// ---------------------------------------------------------------------
struct hprqp_ws {
    hprqp_ws_info info{};
    virtual ~hprqp_ws() = default;
    virtual hprqp_ctx *ctx() const = 0;
    virtual void step(int kind, const hprqp_step_params *sp) = 0;
    virtual void kkt(int64_t iter, hprqp_kkt_out *out) = 0;
    virtual double mnorm(const hprqp_step_params *sp) = 0;
    virtual void sigma_dots(double *out) = 0;
    virtual double sigma_diag(double a, double b) = 0;
    virtual void q_factors(double sigma) = 0;
    virtual void restart() = 0;
    virtual void collect(double *x, double *y, double *z, double *w) = 0;
    virtual void download(int which, double *host, int64_t len) = 0;
    virtual void upload(int which, const double *host, int64_t len) = 0;
};

namespace hprqp {

// the half-steps by entry point; kWsMain runs main_update along the path
enum WsStep {
    kWsMain = 0,
    kWsZxw1,
    kWsTempv,
    kWsY,
    kWsW2,
    kWsNocZxw,
    kWsNocTempv,
    kWsLasso,
    kWsLpZx,
    kWsLpY
};

template <class VT, class MT>
struct WsImpl final : hprqp_ws {
    hprqp_ctx *c_;
    const hprqp_model *md;
    hprqp_params p;
    Reducer red;
    Scaled S;
    std::unique_ptr<Solver<VT, MT>> W;
    int la_lanes = 0, lat_lanes = 0;  // the LASSO operator's shapes for this workspace

    WsImpl(hprqp_ctx *c, const hprqp_model *m, const hprqp_params *pp) : c_(c), md(m), p(*pp), red(c) {
        p.save_cb = nullptr;  // a step-level host saves for itself
        p.auto_save = 0;
    }

    hprqp_ctx *ctx() const override { return c_; }

    // the shared LASSO operator carries its row-product shapes; a workspace
    // sets its own for each call and puts the previous ones back after it
    struct Lanes {
        WsImpl *w;
        OperatorLanesGuard g;
        explicit Lanes(WsImpl *ws) : w(ws), g(ws->md->Q) {
            if (w->S.qkind == QOP_LASSO) {
                w->S.qop->la.lanes = w->la_lanes;
                w->S.qop->lat.lanes = w->lat_lanes;
            }
        }
    };

    // solve() up to its loop (src/algorithm.jl:1904-1976)
    void create(const double *x0, const double *y0) {
        OperatorLanesGuard guard(md->Q);
        const double setup_start = now_s();
        build_scaled(c_, md, &p, red, S);
        check_Q_diagonal(red, S);
        apply_lanes(c_->stream, md, &p, S, p.precision != HPRQP_PREC_FP64);
        if (S.qkind == QOP_LASSO) {
            la_lanes = S.qop->la.lanes;
            lat_lanes = S.qop->lat.lanes;
        }
        W.reset(new Solver<VT, MT>(c_, S, red, &p));
        Solver<VT, MT> &D = *W;
        init_solver(D, md);
        double sigma0 = 1.0;
        if (p.sigma == -1) {
            if (S.norm_c > 1e-16 && S.norm_b > 1e-16 && S.norm_b < 1e16 && S.norm_c < 1e16)
                sigma0 = S.norm_b / S.norm_c;
        } else {
            sigma0 = p.sigma;
        }
        D.sigma = sigma0;
        HPRQP_HIP(hipStreamSynchronize(c_->stream));
        info.setup_time = now_s() - setup_start;
        const double t0 = now_s();
        estimate_lambdas(D, md, p.eig_factor, &info.power_iter_A, &info.power_iter_Q);
        info.power_time = now_s() - t0;
        int64_t nfree = 0;
        for (int32_t i = 0; i < md->n; ++i)
            if (md->l[i] == -INFINITY && md->u[i] == INFINITY) nfree++;
        D.noC = (double)nfree > 0.8 * (double)md->n;
        process_initial_points(D, x0, y0);
        if (S.Q_is_diag) D.update_Q_factors();
        HPRQP_HIP(hipStreamSynchronize(c_->stream));
        info.n = S.n;
        info.m = S.m;
        info.Q_is_diag = S.Q_is_diag;
        info.noC = D.noC;
        info.path = S.qkind == QOP_LASSO ? HPRQP_PATH_LASSO
                    : S.q_empty          ? HPRQP_PATH_LP
                    : D.noC              ? HPRQP_PATH_NOC
                                         : HPRQP_PATH_STANDARD;
        info.precision = p.precision;
        info.lambda_max_A = D.lambda_max_A;
        info.lambda_max_Q = D.lambda_max_Q;
        info.sigma0 = sigma0;
        info.b_scale = S.b_scale;
        info.c_scale = S.c_scale;
        info.norm_b = S.norm_b;
        info.norm_c = S.norm_c;
    }

    void set_scalars(const hprqp_step_params *sp) {
        W->sigma = sp->sigma;
        W->lambda_max_A = sp->lambda_max_A;
        W->lambda_max_Q = sp->lambda_max_Q;
    }

    template <bool F, bool DG>
    void run(int kind, double h1, double h2) {
        Solver<VT, MT> &D = *W;
        switch (kind) {
            case kWsMain: D.template update_impl<F, DG>(h1, h2); break;
            case kWsZxw1: D.template step_zxw1<F, DG>(h1, h2); break;
            case kWsTempv: D.step_tempv(); break;
            case kWsY: D.template step_y<F>(D.tempv.p, h1, h2); break;
            case kWsW2: D.template step_w2<F, DG>(h1, h2); break;
            case kWsNocZxw: D.template step_noc_zxw<F, DG>(h1, h2); break;
            case kWsNocTempv: D.step_noc_tempv(h1, h2); break;
            case kWsLasso: D.template step_lasso<F>(h1, h2); break;
            case kWsLpZx: D.template step_lp_zx<F>(h1, h2); break;
            default: D.template step_y<F>(D.x_hat.p, h1, h2); break;  // kWsLpY
        }
    }

    void step(int kind, const hprqp_step_params *sp) override {
        static const int path_of[] = {-1,
                                      HPRQP_PATH_STANDARD,
                                      HPRQP_PATH_STANDARD,
                                      HPRQP_PATH_STANDARD,
                                      HPRQP_PATH_STANDARD,
                                      HPRQP_PATH_NOC,
                                      HPRQP_PATH_NOC,
                                      HPRQP_PATH_LASSO,
                                      HPRQP_PATH_LP,
                                      HPRQP_PATH_LP};
        HPRQP_REQUIRE(kind >= kWsMain && kind <= kWsLpY, HPRQP_E_ARG, "unknown half-step");
        // the y step is shared by the standard and noC paths
        const bool on_path = kind == kWsMain || path_of[kind] == info.path ||
                             (kind == kWsY && info.path == HPRQP_PATH_NOC);
        HPRQP_REQUIRE(on_path, HPRQP_E_STATE, "half-step is not on this problem's path (hprqp_ws_info.path)");
        HPRQP_REQUIRE(sp->sigma > 0 && sp->h1 >= 0 && sp->h1 <= 1, HPRQP_E_ARG, "step scalars out of range");
        Lanes scope(this);
        // a half-step on its own leaves the QAP one-apply form's products
        // behind the iterates
        if (kind != kWsMain) W->invalidate_products();
        set_scalars(sp);
        const bool full = sp->full != 0, diag = S.Q_is_diag;
        if (full)
            diag ? run<true, true>(kind, sp->h1, sp->h2) : run<true, false>(kind, sp->h1, sp->h2);
        else
            diag ? run<false, true>(kind, sp->h1, sp->h2) : run<false, false>(kind, sp->h1, sp->h2);
    }

    void kkt(int64_t iter, hprqp_kkt_out *o) override {
        Lanes scope(this);
        typename Solver<VT, MT>::Resid r;
        W->compute_residuals(iter, md->obj_constant, p.problem_type == HPRQP_PROBLEM_LASSO, r);
        o->err_Rp = r.err_Rp;
        o->err_Rd = r.err_Rd;
        o->rel_gap = r.rel_gap;
        o->primal_obj = r.primal_obj;
        o->dual_obj = r.dual_obj;
        o->kkt = r.kkt;
    }

    double mnorm(const hprqp_step_params *sp) override {
        Lanes scope(this);
        set_scalars(sp);
        return W->compute_M_norm();
    }

    void sigma_dots(double *out) override {
        Lanes scope(this);
        hprqp::sigma_dots(*W, out);
    }

    double sigma_diag(double a, double b) override {
        HPRQP_REQUIRE(S.Q_is_diag && S.m > 0, HPRQP_E_STATE, "Q is not diagonal or there are no constraints");
        HPRQP_REQUIRE(a > 0 && b > 0, HPRQP_E_ARG, "a and b must be positive");
        return sigma_estimate_diag(*W, a, b);
    }

    void q_factors(double sigma) override {
        HPRQP_REQUIRE(S.Q_is_diag, HPRQP_E_STATE, "Q is not diagonal");
        HPRQP_REQUIRE(sigma > 0, HPRQP_E_ARG, "sigma must be positive");
        W->sigma = sigma;
        W->update_Q_factors();
    }

    void restart() override {
        Lanes scope(this);
        restart_vectors(*W);
    }

    void collect(double *x, double *y, double *z, double *w) override { hprqp::collect(*W, x, y, z, w); }

    // the workspace vector `which` (HPRQP_VEC_*): iterate-precision pointer
    // or FP64 pointer (anchors, diagonal of Q), and its length
    void vec_of(int which, VT *&v, double *&vd, int64_t &len) {
        Solver<VT, MT> &D = *W;
        v = nullptr;
        vd = nullptr;
        len = D.n;
        switch (which) {
            case HPRQP_VEC_X_BAR: v = D.x_bar.p; break;
            case HPRQP_VEC_Y_BAR: v = D.y_bar.p, len = D.m; break;
            case HPRQP_VEC_Z_BAR: v = D.z_bar.p; break;
            case HPRQP_VEC_W_BAR: v = D.w_bar.p; break;
            case HPRQP_VEC_ATDY: v = D.ATdy.p; break;
            case HPRQP_VEC_QATDY: v = D.QATdy.p; break;
            case HPRQP_VEC_DIAG_Q: vd = S.diag_Q.p; break;
            case HPRQP_VEC_X: v = D.x.p; break;
            case HPRQP_VEC_W: v = D.w.p; break;
            case HPRQP_VEC_Y: v = D.y.p, len = D.m; break;
            case HPRQP_VEC_ATY: v = D.ATy.p; break;
            case HPRQP_VEC_QW: v = D.Qw.p; break;
            case HPRQP_VEC_X_HAT: v = D.x_hat.p; break;
            case HPRQP_VEC_DX: v = D.dx.p; break;
            case HPRQP_VEC_DW: v = D.dw.p; break;
            case HPRQP_VEC_DY: v = D.dy.p, len = D.m; break;
            case HPRQP_VEC_S: v = D.s_.p, len = D.m; break;
            case HPRQP_VEC_ATY_BAR: v = D.ATy_bar.p; break;
            case HPRQP_VEC_TEMPV: v = D.tempv.p; break;
            case HPRQP_VEC_QW_BAR: v = D.Qw_bar.p; break;
            case HPRQP_VEC_LAST_X: vd = D.last_x.p; break;
            case HPRQP_VEC_LAST_Y: vd = D.last_y.p, len = D.m; break;
            case HPRQP_VEC_LAST_W: vd = D.last_w.p; break;
            case HPRQP_VEC_LAST_ATY: vd = D.last_ATy.p; break;
            case HPRQP_VEC_LAST_QW: vd = D.last_Qw.p; break;
            case HPRQP_VEC_FACT1: v = D.fact1.p; break;
            case HPRQP_VEC_FACT2: v = D.fact2.p; break;
            case HPRQP_VEC_FACT: v = D.fact.p; break;
            default: throw Error(HPRQP_E_ARG, "unknown vector");
        }
    }

    void download(int which, double *host, int64_t len) override {
        Solver<VT, MT> &D = *W;
        D.flush();
        VT *src = nullptr;
        double *srcd = nullptr;
        int64_t want = 0;
        vec_of(which, src, srcd, want);
        HPRQP_REQUIRE(host != nullptr && len == want, HPRQP_E_ARG, "host buffer length differs from the vector's");
        if (want == 0) return;
        if (!srcd) {
            DevVec<double> t(want);
            double *tp = t.p;
            const VT *sp = src;
            each(c_->stream, want, [=] __device__(int64_t i) { tp[i] = (double)sp[i]; });
            HPRQP_HIP(hipMemcpyAsync(host, t.p, sizeof(double) * want, hipMemcpyDeviceToHost, c_->stream));
            HPRQP_HIP(hipStreamSynchronize(c_->stream));
        } else {
            HPRQP_HIP(hipMemcpyAsync(host, srcd, sizeof(double) * want, hipMemcpyDeviceToHost, c_->stream));
            HPRQP_HIP(hipStreamSynchronize(c_->stream));
        }
    }

    void upload(int which, const double *host, int64_t len) override {
        Solver<VT, MT> &D = *W;
        D.invalidate_products();
        HPRQP_REQUIRE(which != HPRQP_VEC_DIAG_Q, HPRQP_E_ARG, "the diagonal of Q is problem data, not state");
        VT *dst = nullptr;
        double *dstd = nullptr;
        int64_t want = 0;
        vec_of(which, dst, dstd, want);
        HPRQP_REQUIRE(host != nullptr && len == want, HPRQP_E_ARG, "host buffer length differs from the vector's");
        if (want == 0) return;
        DevVec<double> t(want);
        HPRQP_HIP(hipMemcpyAsync(t.p, host, sizeof(double) * want, hipMemcpyHostToDevice, c_->stream));
        if (dstd) {
            HPRQP_HIP(hipMemcpyAsync(dstd, t.p, sizeof(double) * want, hipMemcpyDeviceToDevice, c_->stream));
        } else {
            const double *tp = t.p;
            VT *dp = dst;
            each(c_->stream, want, [=] __device__(int64_t i) { dp[i] = (VT)tp[i]; });
        }
        HPRQP_HIP(hipStreamSynchronize(c_->stream));
    }
};

hprqp_ws *ws_create(hprqp_ctx *ctx, const hprqp_model *md, const hprqp_params *p, const double *x0,
                    const double *y0) {
    HPRQP_REQUIRE(p->check_iter > 0, HPRQP_E_ARG, "check_iter must be positive");
    HPRQP_REQUIRE(p->sigma == -1 || p->sigma > 0, HPRQP_E_ARG, "sigma must be positive or -1");
    std::unique_ptr<hprqp_ws> w;
    switch (p->precision) {
        case HPRQP_PREC_FP64: {
            auto *i = new WsImpl<double, double>(ctx, md, p);
            w.reset(i);
            i->create(x0, y0);
            break;
        }
        case HPRQP_PREC_MIXED: {
            auto *i = new WsImpl<double, float>(ctx, md, p);
            w.reset(i);
            i->create(x0, y0);
            break;
        }
        case HPRQP_PREC_FP32: {
            auto *i = new WsImpl<float, float>(ctx, md, p);
            w.reset(i);
            i->create(x0, y0);
            break;
        }
        default: throw Error(HPRQP_E_ARG, "unknown precision mode");
    }
    return w.release();
}

void ws_info(const hprqp_ws *w, hprqp_ws_info *info) { *info = w->info; }
hprqp_ctx *ws_ctx(const hprqp_ws *w) { return w->ctx(); }
void ws_step(hprqp_ws *w, int kind, const hprqp_step_params *sp) { w->step(kind, sp); }
void ws_kkt(hprqp_ws *w, int64_t iter, hprqp_kkt_out *out) { w->kkt(iter, out); }
double ws_mnorm(hprqp_ws *w, const hprqp_step_params *sp) { return w->mnorm(sp); }
void ws_sigma_dots(hprqp_ws *w, double *out) { w->sigma_dots(out); }
double ws_sigma_diag(hprqp_ws *w, double a, double b) { return w->sigma_diag(a, b); }
void ws_q_factors(hprqp_ws *w, double sigma) { w->q_factors(sigma); }
void ws_restart(hprqp_ws *w) { w->restart(); }
void ws_collect(hprqp_ws *w, double *x, double *y, double *z, double *wv) { w->collect(x, y, z, wv); }
void ws_download(hprqp_ws *w, int which, double *host, int64_t len) { w->download(which, host, len); }
void ws_upload(hprqp_ws *w, int which, const double *host, int64_t len) { w->upload(which, host, len); }
void ws_destroy(hprqp_ws *w) { delete w; }

}  // namespace hprqp
