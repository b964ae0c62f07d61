// hprqp_batch.hip -- many HPR-QP instances of uneven sizes solved together.
//
// Each instance i:  min 1/2 x'(diag(d_i) + U_i U_i')x + c_i'x
//                   s.t. AL_i <= A_i x <= AU_i,  l_i <= x <= u_i
// with U_i dense n_i x k_i (column-major) and A_i sparse (CSR, with its
// stored transpose).  Q_i is an operator, so -- as for every operator Q in
// the reference (src/utils.jl:364-406) -- no scaling is applied.
//
// Device side: one 256-thread workgroup owns one instance for a whole check
// interval (the iterations between two residual evaluations of the
// reference's cadence, src/algorithm.jl:1977-2038): it runs the standard
// z/x/w1 - tempv - y - w2 steps (src/kernels.jl:2301-2832) with block-level
// reductions, the M-norm(s) the interval needs, the FP64 residuals at the
// interval's end and the dot products a sigma update would need.  All
// instances share the cadence, so one launch covers one interval for the
// whole batch; retired instances return at once.
// Host side: per-instance check_break / check_restart / update_sigma /
// do_restart decisions from a 32-double record per instance
// (src/algorithm.jl:323-576), exactly as the single-instance loop.
//
// A second kind of batch (hprqp_batch_create_csr) holds sparse Q_i (CSR,
// symmetric pattern): each solve first scales every instance as the single
// solve does (k_batch_scale: Ruiz x 10 and Pock-Chambolle,
// src/utils.jl:350-804, then (Q + Q')/2), the kernels take QK = 1 and form
// Q v by CSR row products, and the residuals carry the row and column norms
// (compute_residuals!, src/algorithm.jl:37-133).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "hprqp_internal.hpp"
#include "hprqp_kernels.hpp"

namespace hprqp {

// threads per instance workgroup (a multiple of 64)
constexpr int kBB = 256;
constexpr int kBW = kBB / 64;

static double bnow() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// per-instance control (host -> device) and record (device -> host)
struct BCtl {
    double sigma, lamA, lamQ, h_inner;  // h_inner: Halpern counter at interval start
    int32_t active, K, restart, check_iter;
    int64_t iter0;
    int32_t want_last_gap, want_cur_gap, init, pad;
};
enum {
    R_CX = 0, R_XQX, R_ZX, R_YS, R_RD, R_ATY, R_QX, R_RP, R_AX, R_BND,  // residual parts
    R_LASTGAP, R_CURGAP,                                                // M-norms
    R_DXDX, R_DQWATDY, R_DQWDQW, R_DWDQW, R_ATDYQATDY, R_DYDY,          // sigma dots
    R_POW_A, R_POW_Q, R_POW_ITA, R_POW_ITQ,                             // power method
    R_DONE,  // iteration at which the instance met the tolerance inside an interval, -1 if not
    R_NB,    // ||b||_2, ||c||_2 of the scaled data (sigma_0)
    R_NC,
    R_COUNT = 26
};

template <class MT>
struct BArgs {
    int32_t count;
    const int32_t *order;  // blockIdx -> instance (largest first)
    const int32_t *n, *m, *k;
    const int64_t *noff, *moff, *uoff, *arp_off, *atrp_off;
    const MT *U, *d;            // iteration precision
    const double *U64, *d64;    // FP64 (residuals, power method)
    const int32_t *a_rp, *a_ci, *at_rp, *at_ci;  // packed CSR, row pointers local to the instance
    const MT *a_v, *at_v;
    const double *a_v64, *at_v64;
    // CSR Q (QK = 1): packed, row pointers global (instance b's at
    // qrp_off[b]), column indices local; the scaling's column and row norms
    // (null for the operator kind, which is not scaled)
    const int32_t *q_rp, *q_ci;
    const int64_t *qrp_off;
    const MT *q_v;
    const double *q_v64, *cn, *rn;
    const double *c, *l, *u, *AL, *AU;
    const double *c0, *nbi, *nci;  // objective constant, |b|_inf, |c|_inf (residual norms)
    double tol;
    int64_t print_freq;  // -1: the adaptive print cadence
    const MT *cM, *lM, *uM, *ALM, *AUM;
    double *x, *w, *w_bar, *x_hat, *x_bar, *z_bar, *dx, *dw, *Qw, *Qw_bar, *tempv, *ATy, *ATy_bar, *ATdy, *QATdy,
        *dQw;
    double *y, *y_bar, *s, *dy;
    double *last_x, *last_w, *last_y, *last_ATy;
    double *pq, *pz, *pq2, *pz2;  // eigenvalue scratch (n-packed and m-packed)
    // A_i U_i (m_i x k_i row-major, FP64 sums of the iteration's A and U
    // values), packed at sum_j m_j k_j
    double *AUk;
    const int64_t *auoff;
    BCtl *ctl;
    double *rec;
};

// ---------------- block primitives (256 threads, 4 waves) ----------------
__device__ __forceinline__ double bsum(double v, double *sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < kBW; ++q) r += sh[q];
    return r;
}
__device__ __forceinline__ double bmax(double v, double *sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = dmax(v, __shfl_xor(v, off));
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int q = 1; q < kBW; ++q) r = dmax(r, sh[q]);
    return r;
}

// t[j] = sum_i U[i + j n] v[i], j < k  (t in LDS).  Columns go in groups of
// kG: each thread keeps kG partial sums over its rows, so one round of
// shuffles and one barrier pair serve kG columns and the loads of a group
// are independent.  Columns past k in the last group read column k-1 again
// (cache hits) and are dropped, so no load sits behind a branch.
constexpr int kG = 8;
template <class UT, class VT>
__device__ void lr_tvec(int32_t n, int32_t k, const UT *__restrict__ U, const VT *__restrict__ v, double *t,
                        double *sh) {
    __shared__ double part[kBB / 64][kG];
    (void)sh;
    for (int32_t j0 = 0; j0 < k; j0 += kG) {
        double acc[kG];
        int32_t co[kG];  // column offsets (an instance's U has < 2^31 entries)
#pragma unroll
        for (int g = 0; g < kG; ++g) {
            acc[g] = 0.0;
            co[g] = min(j0 + g, k - 1) * n;
        }
        for (int32_t i = threadIdx.x; i < n; i += kBB) {
            const double vi = (double)v[i];
#pragma unroll
            for (int g = 0; g < kG; ++g) acc[g] += (double)U[co[g] + i] * vi;
        }
#pragma unroll
        for (int g = 0; g < kG; ++g) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc[g] += __shfl_xor(acc[g], off);
        }
        if ((threadIdx.x & 63) == 0) {
#pragma unroll
            for (int g = 0; g < kG; ++g) part[threadIdx.x >> 6][g] = acc[g];
        }
        __syncthreads();
        const int gmax = min(kG, k - j0);
        if (threadIdx.x < gmax) {
            const int g = threadIdx.x;
            double r = 0.0;
#pragma unroll
            for (int q = 0; q < kBW; ++q) r += part[q][g];
            t[j0 + g] = r;
        }
        __syncthreads();
    }
    __syncthreads();
}
// U_i. t with four independent accumulators
template <class UT>
__device__ __forceinline__ double lr_row(int32_t n, int32_t k, const UT *__restrict__ U, int32_t i, const double *t) {
    const UT *p = U + i;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int32_t j = 0;
    for (; j + 4 <= k; j += 4) {
        a0 += (double)p[(int64_t)j * n] * t[j];
        a1 += (double)p[(int64_t)(j + 1) * n] * t[j + 1];
        a2 += (double)p[(int64_t)(j + 2) * n] * t[j + 2];
        a3 += (double)p[(int64_t)(j + 3) * n] * t[j + 3];
    }
    for (; j < k; ++j) a0 += (double)p[(int64_t)j * n] * t[j];
    return (a0 + a1) + (a2 + a3);
}
// Qv for all i into out: out_i = d_i v_i + U_i. t  (t = U'v already in LDS)
template <class UT, class VT>
__device__ void lr_apply(int32_t n, int32_t k, const UT *U, const UT *d, const VT *v, double *out, double *t,
                         double *sh) {
    lr_tvec(n, k, U, v, t, sh);
    for (int32_t i = threadIdx.x; i < n; i += kBB) out[i] = (double)d[i] * (double)v[i] + lr_row(n, k, U, i, t);
    __syncthreads();
}
// y_r = sum A_r. v  for the instance's m rows: one wave per row, or, with
// fewer rows than waves (a budget row of n nonzeros), each row by the whole
// block with four loads per thread in flight -- one wave walking a long row
// waits on its loads n/64 times in a row
template <class AT_, class VT>
__device__ void csr_rows_wave(int32_t m, const int32_t *rp, const int32_t *ci, const AT_ *v, const VT *x,
                              double *out) {
    if (m < kBW) {
        __shared__ double shr[kBW];
        for (int32_t r = 0; r < m; ++r) {
            const int32_t s = rp[r], e = rp[r + 1];
            double a[4] = {0.0, 0.0, 0.0, 0.0};
            for (int32_t q0 = s + (int32_t)threadIdx.x; q0 < e; q0 += 4 * kBB) {
                AT_ vv[4];
                int32_t cc[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {  // past the row's end: its first nonzero, weight zero
                    const int32_t q = q0 + u * kBB, at = q < e ? q : s;
                    vv[u] = v[at];
                    cc[u] = ci[at];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double pr = (double)vv[u] * (double)x[cc[u]];
                    a[u] += q0 + u * kBB < e ? pr : 0.0;
                }
            }
            const double tot = bsum((a[0] + a[1]) + (a[2] + a[3]), shr);
            if (threadIdx.x == 0) out[r] = tot;
        }
        __syncthreads();
        return;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int32_t r = wave; r < m; r += kBB / 64) {
        double acc = 0.0;
        for (int32_t q = rp[r] + lane; q < rp[r + 1]; q += 64) acc += (double)v[q] * (double)x[ci[q]];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (lane == 0) out[r] = acc;
    }
    __syncthreads();
}
// transpose product, one thread per column j (rows of the stored transpose)
template <class AT_, class VT>
__device__ __forceinline__ double csr_row(const int32_t *rp, const int32_t *ci, const AT_ *v, const VT *x, int32_t j) {
    double acc = 0.0;
    for (int32_t q = rp[j]; q < rp[j + 1]; ++q) acc += (double)v[q] * (double)x[ci[q]];
    return acc;
}

// An instance's arrays.  The table lives in LDS (filled once per block) and
// the step code reads each pointer from it where it is used: kept in
// registers, the fifty pointers overflowed the scalar registers into vector
// registers and scratch and capped the interval kernel at three waves per
// SIMD.
template <class MT>
struct IPtr {
    const MT *U, *d;
    const double *U64, *d64;
    const int32_t *arp, *aci, *atrp, *atci;
    const MT *av, *atv;
    const double *av64, *atv64;
    const int32_t *qrp, *qci;
    const MT *qv;
    const double *qv64, *cn, *rn;
    const double *c, *l, *u, *AL, *AU;
    const MT *cM, *lM, *uM, *ALM, *AUM;
    double *x, *w, *w_bar, *x_hat, *x_bar, *z_bar, *dx, *dw, *Qw, *Qw_bar, *tempv, *ATy, *ATy_bar, *ATdy, *QATdy,
        *dQw;
    double *y, *y_bar, *s, *dy, *last_x, *last_w, *last_y, *last_ATy, *pq, *pz, *pq2, *pz2;
    const double *AUk;  // A U (m x k row-major)
};
template <class MT>
struct Inst {
    int32_t n, m, k;
    const IPtr<MT> *P;
};

// fills the block's table (thread 0) and returns the instance; every thread
// of the block must call it
template <class MT>
__device__ Inst<MT> inst_of(const BArgs<MT> &a, int b, IPtr<MT> *tab) {
    Inst<MT> I;
    I.n = a.n[b];
    I.m = a.m[b];
    I.k = a.k[b];
    I.P = tab;
    if (threadIdx.x == 0) {
        IPtr<MT> &T = *tab;
        const int64_t no = a.noff[b], mo = a.moff[b], uo = a.uoff[b];
        T.U = a.U + uo;
        T.d = a.d + no;
        T.U64 = a.U64 + uo;
        T.d64 = a.d64 + no;
        T.arp = a.a_rp + a.arp_off[b];
        T.atrp = a.at_rp + a.atrp_off[b];
        T.aci = a.a_ci;
        T.atci = a.at_ci;
        T.av = a.a_v;
        T.atv = a.at_v;
        T.av64 = a.a_v64;
        T.atv64 = a.at_v64;
        T.qrp = a.q_rp ? a.q_rp + a.qrp_off[b] : nullptr;
        T.qci = a.q_ci;
        T.qv = a.q_v;
        T.qv64 = a.q_v64;
        T.cn = a.cn ? a.cn + no : nullptr;
        T.rn = a.rn ? a.rn + mo : nullptr;
        T.c = a.c + no;
        T.l = a.l + no;
        T.u = a.u + no;
        T.AL = a.AL + mo;
        T.AU = a.AU + mo;
        T.cM = a.cM + no;
        T.lM = a.lM + no;
        T.uM = a.uM + no;
        T.ALM = a.ALM + mo;
        T.AUM = a.AUM + mo;
#define P_(f) T.f = a.f + no
        P_(x); P_(w); P_(w_bar); P_(x_hat); P_(x_bar); P_(z_bar); P_(dx); P_(dw); P_(Qw); P_(Qw_bar); P_(tempv);
        P_(ATy); P_(ATy_bar); P_(ATdy); P_(QATdy); P_(dQw); P_(last_x); P_(last_w); P_(last_ATy); P_(pq); P_(pq2);
#undef P_
        T.y = a.y + mo;
        T.y_bar = a.y_bar + mo;
        T.s = a.s + mo;
        T.dy = a.dy + mo;
        T.last_y = a.last_y + mo;
        T.pz = a.pz + mo;
        T.pz2 = a.pz2 + mo;
        T.AUk = a.AUk ? a.AUk + a.auoff[b] : nullptr;
    }
    __syncthreads();
    return I;
}

// The projections of the iteration's n-vectors on the columns of U, kept in
// k-space for a block's whole check interval: tw = U' w, taty = U' ATy and
// those of the Halpern anchors, tlw = U' last_w, tlaty = U' last_ATy; tx =
// U' x_hat of the current step.  They start exact at every interval (the
// anchors change only at restarts, which happen between intervals) and
// follow the iteration's own linear updates inside it.
struct KSpace {
    double tw[64], taty[64], tlw[64], tlaty[64], tx[64];
};

template <class MT>
__device__ void bkspace_init(Inst<MT> &I, KSpace &ks, double *sh) {
    const int32_t n = I.n, k = I.k;
    lr_tvec(n, k, I.P->U, I.P->w, ks.tw, sh);
    lr_tvec(n, k, I.P->U, I.P->ATy, ks.taty, sh);
    lr_tvec(n, k, I.P->U, I.P->last_w, ks.tlw, sh);
    lr_tvec(n, k, I.P->U, I.P->last_ATy, ks.tlaty, sh);
}

// out_r = sum A_r. x for `rows` rows with L lanes per row (L divides 64):
// the block takes kBB / L rows at a time; each lane loads eight of its row's
// nonzeros (stride L, past the row's end the row's first entry with weight
// zero), then their eight gathers, so each lane has eight gathers in flight;
// the row's lanes sum by xor shuffles
template <int L, class AT_, class XF>
__device__ void csr_rows_lanes_f(int32_t rows, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                 const AT_ *__restrict__ v, const XF &xf, double *__restrict__ out) {
    constexpr int G = kBB / L, U = 8;
    const int lane = (int)threadIdx.x % L, grp = (int)threadIdx.x / L;
    for (int32_t r0 = 0; r0 < rows; r0 += G) {
        const int32_t r = r0 + grp;
        double acc = 0.0;
        if (r < rows) {
            const int32_t s = rp[r], e = rp[r + 1];
            for (int32_t q0 = s + lane; q0 < e; q0 += U * L) {
                int32_t cc[U];
                double vv[U], xx[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int32_t q = q0 + u * L, at = q < e ? q : s;
                    cc[u] = ci[at];
                    vv[u] = q < e ? (double)v[at] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) xx[u] = xf(cc[u]);
#pragma unroll
                for (int u = 0; u < U; ++u) acc += vv[u] * xx[u];
            }
        }
        if constexpr (L > 1) {
#pragma unroll
            for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        }
        if (lane == 0 && r < rows) out[r] = acc;
    }
    __syncthreads();
}
template <int L, class AT_, class VT>
__device__ void csr_rows_lanes(int32_t rows, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                               const AT_ *__restrict__ v, const VT *__restrict__ x, double *__restrict__ out) {
    csr_rows_lanes_f<L>(rows, rp, ci, v, [=](int32_t c) { return (double)x[c]; }, out);
}
// the lanes per row from the matrix's mean row length (block-uniform): one
// lane for rows of a few nonzeros, four up to a few hundred (a row then
// takes few rounds of eight gathers per lane), sixteen beyond
template <class AT_, class VT>
__device__ void csr_rows_auto(int32_t rows, const int32_t *rp, const int32_t *ci, const AT_ *v, const VT *x,
                              double *out) {
    if (rows <= 0) return;
    const int32_t mean = (rp[rows] - rp[0]) / rows;
    if (mean <= 6)
        csr_rows_lanes<1>(rows, rp, ci, v, x, out);
    else if (mean <= 256)
        csr_rows_lanes<4>(rows, rp, ci, v, x, out);
    else
        csr_rows_lanes<16>(rows, rp, ci, v, x, out);
}

// Q's row products gather from an LDS copy of the vector when it fits
// (kQxs doubles): Q has ~100+ nonzeros per row, and a block's gathers from
// L1 bound the product (every 8-byte gather is its own cache-line access)
constexpr int32_t kQxs = 4096;
template <class AT_, class VT>
__device__ void qcsr_rows(int32_t n, const int32_t *rp, const int32_t *ci, const AT_ *v, const VT *x,
                          double *out) {
    if (n <= 0) return;
    if (n > kQxs) {
        csr_rows_auto(n, rp, ci, v, x, out);
        return;
    }
    __shared__ double xs[kQxs];
    for (int32_t i = threadIdx.x; i < n; i += kBB) xs[i] = (double)x[i];
    __syncthreads();
    const int32_t mean = (rp[n] - rp[0]) / n;
    auto xf = [&](int32_t c) { return xs[c]; };
    if (mean <= 6)
        csr_rows_lanes_f<1>(n, rp, ci, v, xf, out);
    else if (mean <= 256)
        csr_rows_lanes_f<4>(n, rp, ci, v, xf, out);
    else
        csr_rows_lanes_f<16>(n, rp, ci, v, xf, out);
}

// out = Q v for a CSR instance (QK = 1), in the iteration's precision (qv)
// or in FP64 (qv64)
template <class MT, class AT_, class VT>
__device__ void qcsr_apply(const Inst<MT> &I, const AT_ *qv, const VT *v, double *out) {
    qcsr_rows(I.n, I.P->qrp, I.P->qci, qv, v, out);
}

// Q v of either kind: out = Q v; t, sh LDS scratch of the low-rank product
template <int QK, class MT, class VT>
__device__ void bq_apply(const Inst<MT> &I, const VT *v, double *out, double *t, double *sh, bool fp64 = false) {
    if constexpr (QK == 0) {
        if (fp64)
            lr_apply(I.n, I.k, I.P->U64, I.P->d64, v, out, t, sh);
        else
            lr_apply(I.n, I.k, I.P->U, I.P->d, v, out, t, sh);
    } else {
        if (fp64)
            qcsr_apply(I, I.P->qv64, v, out);
        else
            qcsr_apply(I, I.P->qv, v, out);
    }
}

// One HPR step of the standard path with a sparse Q (QK = 1): the
// reference's four half-steps (src/kernels.jl:2301-2832), each a row
// product (Q w, Q w_bar, A tempv, A^T y_bar) by csr_rows_auto into a work
// vector followed by its elementwise update
template <class MT>
__device__ void bstep_csr(Inst<MT> &I, double sigma, double lamA, double lamQ, double h1, double h2, bool full) {
    const int32_t n = I.n, m = I.m;
    const double f2s = 1.0 / (1.0 + sigma * lamQ), f1s = 1.0 - f2s;
    // z/x/w1 on Q w
    qcsr_rows(n, I.P->qrp, I.P->qci, I.P->qv, I.P->w, I.P->Qw);
    for (int32_t i = threadIdx.x; i < n; i += kBB) {
        const double qw = I.P->Qw[i];
        const double wi = I.P->w[i], xi = I.P->x[i];
        const double z_raw = xi + sigma * (-qw + I.P->ATy[i] - (double)I.P->cM[i]);
        const double xb = proj(z_raw, (double)I.P->lM[i], (double)I.P->uM[i]);
        const double xh = 2.0 * xb - xi;
        if (full) {
            I.P->dx[i] = xb - xi;
            I.P->x_bar[i] = xb;
            I.P->z_bar[i] = (xb - z_raw) / sigma;
        }
        I.P->x[i] = fma(h2, xh, h1 * I.P->last_x[i]);
        I.P->x_hat[i] = xh;
        I.P->w_bar[i] = fma(f1s, wi, f2s * xh);
    }
    __syncthreads();
    if (m > 0) {
        // tempv = x_hat + sigma (Q w - Q w_bar)
        qcsr_rows(n, I.P->qrp, I.P->qci, I.P->qv, I.P->w_bar, I.P->Qw_bar);
        for (int32_t i = threadIdx.x; i < n; i += kBB)
            I.P->tempv[i] = I.P->x_hat[i] + sigma * (I.P->Qw[i] - I.P->Qw_bar[i]);
        __syncthreads();
        // y step on A tempv
        csr_rows_auto(m, I.P->arp, I.P->aci, I.P->av, I.P->tempv, I.P->pz);
        const double fa = lamA * sigma, fb = 1.0 / fa;
        for (int32_t r = threadIdx.x; r < m; r += kBB) {
            const double yi = I.P->y[r];
            const double s_raw = I.P->pz[r] - fa * yi;
            const double sp = proj(s_raw, (double)I.P->ALM[r], (double)I.P->AUM[r]);
            const double yb = fb * (sp - s_raw);
            if (full) {
                I.P->s[r] = sp;
                I.P->dy[r] = yb - yi;
            }
            I.P->y_bar[r] = yb;
            I.P->y[r] = fma(h2, 2.0 * yb - yi, h1 * I.P->last_y[r]);
        }
        __syncthreads();
        // A^T y_bar into tempv (no longer needed this step)
        csr_rows_auto(n, I.P->atrp, I.P->atci, I.P->atv, I.P->y_bar, I.P->tempv);
    }
    // w2 step
    const double fs = sigma / (1.0 + sigma * lamQ);
    for (int32_t i = threadIdx.x; i < n; i += kBB) {
        const double atyb = m > 0 ? I.P->tempv[i] : 0.0;
        const double wi = I.P->w[i], wbi = I.P->w_bar[i], atyi = I.P->ATy[i];
        const double wbn = wbi + fs * (atyb - atyi);
        I.P->w[i] = h1 * I.P->last_w[i] + h2 * (2.0 * wbn - wi);
        I.P->ATy[i] = h1 * I.P->last_ATy[i] + h2 * (2.0 * atyb - atyi);
        if (full) {
            I.P->w_bar[i] = wbn;
            I.P->dw[i] = wbn - wi;
            I.P->ATdy[i] = atyb - atyi;
            I.P->ATy_bar[i] = atyb;
        }
    }
    __syncthreads();
}

// One HPR step of the standard path (Q = diag(d) + U U' an operator, noC
// false; src/kernels.jl:2301-2832) with two passes over U instead of four:
//  * z/x/w1 takes Q w = d.*w + U tw from tw in k-space and leaves x_hat;
//    one pass over U then gives tx = U' x_hat;
//  * w_bar = f1 w + f2 x_hat gives U' w_bar = f1 tw + f2 tx, so
//    tempv = x_hat + sigma (Q w - Q w_bar)
//          = x_hat + sigma f2 (d.*(w - x_hat) + U (tw - tx)),
//    and the y step needs only A tempv = A v + sigma f2 (A U)(tw - tx) with
//    v = x_hat + sigma f2 d.*(w - x_hat): no pass over U;
//  * the w2 step's updates of w and ATy are linear in A^T y_bar, whose
//    projection is U' A^T y_bar = (A U)' y_bar, so tw and taty follow in
//    k-space.
template <class MT>
__device__ void bstep(Inst<MT> &I, KSpace &ks, double sigma, double lamA, double lamQ, double h1, double h2,
                      bool full, double *sh) {
    const int32_t n = I.n, m = I.m, k = I.k;
    const double f2s = 1.0 / (1.0 + sigma * lamQ), f1s = 1.0 - f2s, sf2 = sigma * f2s;
    // z/x/w1: Qw = d.*w + U tw
    for (int32_t i = threadIdx.x; i < n; i += kBB) {
        const double wi = I.P->w[i], xi = I.P->x[i], di = (double)I.P->d[i];
        const double qw = di * wi + lr_row(n, k, I.P->U, i, ks.tw);
        const double z_raw = xi + sigma * (-qw + I.P->ATy[i] - (double)I.P->cM[i]);
        const double xb = proj(z_raw, (double)I.P->lM[i], (double)I.P->uM[i]);
        const double xh = 2.0 * xb - xi;
        if (full) {
            I.P->dx[i] = xb - xi;
            I.P->x_bar[i] = xb;
            I.P->z_bar[i] = (xb - z_raw) / sigma;
        }
        I.P->x[i] = fma(h2, xh, h1 * I.P->last_x[i]);
        I.P->x_hat[i] = xh;
        I.P->w_bar[i] = fma(f1s, wi, f2s * xh);
        if (m > 0) I.P->tempv[i] = fma(sf2, di * (wi - xh), xh);  // v
    }
    __syncthreads();
    lr_tvec(n, k, I.P->U, I.P->x_hat, ks.tx, sh);
    if (m > 0) {
        // y step: A tempv = A v + sigma f2 (A U)(tw - tx)
        csr_rows_wave(m, I.P->arp, I.P->aci, I.P->av, I.P->tempv, I.P->pz);
        const double fa = lamA * sigma, fb = 1.0 / fa;
        for (int32_t r = threadIdx.x; r < m; r += kBB) {
            const double *au = I.P->AUk + (int64_t)r * k;
            double corr = 0.0;
            for (int32_t j = 0; j < k; ++j) corr += au[j] * (ks.tw[j] - ks.tx[j]);
            const double yi = I.P->y[r];
            const double s_raw = fma(sf2, corr, I.P->pz[r]) - fa * yi;
            const double sp = proj(s_raw, (double)I.P->ALM[r], (double)I.P->AUM[r]);
            const double yb = fb * (sp - s_raw);
            if (full) {
                I.P->s[r] = sp;
                I.P->dy[r] = yb - yi;
            }
            I.P->y_bar[r] = yb;
            I.P->y[r] = fma(h2, 2.0 * yb - yi, h1 * I.P->last_y[r]);
        }
        __syncthreads();
    }
    // w2 step: A^T y_bar
    const double fs = sigma / (1.0 + sigma * lamQ);
    for (int32_t i = threadIdx.x; i < n; i += kBB) {
        const double atyb = m > 0 ? csr_row(I.P->atrp, I.P->atci, I.P->atv, I.P->y_bar, i) : 0.0;
        const double wi = I.P->w[i], wbi = I.P->w_bar[i], atyi = I.P->ATy[i];
        const double wbn = wbi + fs * (atyb - atyi);
        I.P->w[i] = h1 * I.P->last_w[i] + h2 * (2.0 * wbn - wi);
        I.P->ATy[i] = h1 * I.P->last_ATy[i] + h2 * (2.0 * atyb - atyi);
        if (full) {
            I.P->w_bar[i] = wbn;
            I.P->dw[i] = wbn - wi;
            I.P->ATdy[i] = atyb - atyi;
            I.P->ATy_bar[i] = atyb;
        }
    }
    // the same updates in k-space: ta = (A U)' y_bar = U' A^T y_bar
    for (int32_t j = threadIdx.x; j < k; j += kBB) {
        double ta = 0.0;
        for (int32_t r = 0; r < m; ++r) ta += I.P->AUk[(int64_t)r * k + j] * I.P->y_bar[r];
        const double twj = ks.tw[j], tatyj = ks.taty[j];
        const double twbn = fma(f1s, twj, f2s * ks.tx[j]) + fs * (ta - tatyj);
        ks.tw[j] = h1 * ks.tlw[j] + h2 * (2.0 * twbn - twj);
        ks.taty[j] = h1 * ks.tlaty[j] + h2 * (2.0 * ta - tatyj);
    }
    __syncthreads();
}

// compute_M_norm! (src/algorithm.jl:689-739), non-diagonal Q
template <int QK, class MT>
__device__ double bmnorm(Inst<MT> &I, double sigma, double lamA, double lamQ, double *t, double *sh) {
    const int32_t n = I.n, m = I.m;
    bq_apply<QK>(I, I.P->dw, I.P->dQw, t, sh);
    if (m > 0) {
        for (int32_t i = threadIdx.x; i < n; i += kBB) I.P->ATdy[i] = csr_row(I.P->atrp, I.P->atci, I.P->atv, I.P->dy, i);
        __syncthreads();
        bq_apply<QK>(I, I.P->ATdy, I.P->QATdy, t, sh);
    }
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s6 = 0, s7 = 0;
    for (int32_t i = threadIdx.x; i < n; i += kBB) {
        const double a = I.P->dx[i], b = I.P->dQw[i], c = m > 0 ? I.P->ATdy[i] : 0.0, d = m > 0 ? I.P->QATdy[i] : 0.0,
                     e = I.P->dw[i];
        s0 += a * a;
        s1 += b * a;
        s2 += b * c;
        s3 += c * a;
        s4 += c * d;
        s6 += e * b;
    }
    for (int32_t r = threadIdx.x; r < m; r += kBB) s7 += I.P->dy[r] * I.P->dy[r];
    s0 = bsum(s0, sh);
    s1 = bsum(s1, sh);
    s2 = bsum(s2, sh);
    s3 = bsum(s3, sh);
    s4 = bsum(s4, sh);
    s6 = bsum(s6, sh);
    s7 = bsum(s7, sh);
    double M1 = 0.0, M3 = 0.0, M2 = (1.0 / sigma) * s0 - 2.0 * s1;
    if (m > 0) {
        M1 = sigma * lamA * s7 - 2.0 * sigma * s2;
        M2 += 2.0 * s3;
        M3 = (sigma * sigma) / (1.0 + sigma * lamQ) * s4;
        M1 += sigma * lamQ * s6;
    } else {
        M1 = sigma * lamQ * s6;
    }
    M2 += dmax(M1, 0.0);
    return sqrt(dmax(M2, 0.0) + dmax(M3, 0.0));
}

// residual parts in FP64 against the FP64 data (src/algorithm.jl:37-133)
template <int QK, class MT>
__device__ void bresid(Inst<MT> &I, bool iter0, double *rec, double *t, double *sh) {
    const int32_t n = I.n, m = I.m;
    // Qx (into dQw as scratch), A^T y_bar (into ATdy), A x_bar (into pz)
    bq_apply<QK>(I, I.P->x_bar, I.P->dQw, t, sh, true);
    if (m > 0) {
        for (int32_t i = threadIdx.x; i < n; i += kBB) I.P->ATdy[i] = csr_row(I.P->atrp, I.P->atci, I.P->atv64, I.P->y_bar, i);
        csr_rows_wave(m, I.P->arp, I.P->aci, I.P->av64, I.P->x_bar, I.P->pz);
    }
    double cx = 0, xqx = 0, zx = 0, rd = 0, aty = 0, qx = 0, bnd = 0, ys = 0, rp = 0, ax = 0;
    for (int32_t i = threadIdx.x; i < n; i += kBB) {
        const double x = I.P->x_bar[i], q = I.P->dQw[i], z = I.P->z_bar[i], ci = I.P->c[i], a = m > 0 ? I.P->ATdy[i] : 0.0;
        // scaled instances: the original-space residuals (column norms)
        const double sf = QK == 1 ? I.P->cn[i] : 1.0;
        cx += ci * x;
        xqx += x * q;
        zx += z * x;
        rd = dmax(rd, fabs((q + ci - a - z) * sf));
        aty = dmax(aty, fabs(a * sf));
        qx = dmax(qx, fabs(q * sf));
        if (iter0) bnd = dmax(bnd, (dmax(I.P->l[i] - x, 0.0) + dmax(x - I.P->u[i], 0.0)) / sf);
    }
    for (int32_t r = threadIdx.x; r < m; r += kBB) {
        ys += I.P->y_bar[r] * I.P->s[r];
        const double v = I.P->pz[r];
        const double sf = QK == 1 ? I.P->rn[r] : 1.0;
        rp = dmax(rp, fabs((proj(v, I.P->AL[r], I.P->AU[r]) - v) * sf));
        ax = dmax(ax, fabs(v * sf));
    }
    cx = bsum(cx, sh);
    xqx = bsum(xqx, sh);
    zx = bsum(zx, sh);
    ys = bsum(ys, sh);
    rd = bmax(rd, sh);
    aty = bmax(aty, sh);
    qx = bmax(qx, sh);
    rp = bmax(rp, sh);
    ax = bmax(ax, sh);
    bnd = bmax(bnd, sh);
    if (threadIdx.x == 0) {
        rec[R_CX] = cx;
        rec[R_XQX] = xqx;
        rec[R_ZX] = zx;
        rec[R_YS] = ys;
        rec[R_RD] = rd;
        rec[R_ATY] = aty;
        rec[R_QX] = qx;
        rec[R_RP] = rp;
        rec[R_AX] = ax;
        rec[R_BND] = bnd;
    }
    __syncthreads();
}

// the relative primal residual alone (compute_residuals!'s err_Rp,
// src/algorithm.jl:37-133), FP64: A x_bar over the FP64 A, no pass over Q.
// The KKT residual is the max of err_Rp, err_Rd and the gap, so err_Rp >= tol
// settles a print step's termination test without the FP64 Q x
template <int QK, class MT>
__device__ double bprimal_err(Inst<MT> &I, double nbi, double *sh) {
    const int32_t m = I.m;
    csr_rows_wave(m, I.P->arp, I.P->aci, I.P->av64, I.P->x_bar, I.P->pz);
    double rp = 0, ax = 0;
    for (int32_t r = threadIdx.x; r < m; r += kBB) {
        const double v = I.P->pz[r];
        const double sf = QK == 1 ? I.P->rn[r] : 1.0;
        rp = dmax(rp, fabs((proj(v, I.P->AL[r], I.P->AU[r]) - v) * sf));
        ax = dmax(ax, fabs(v * sf));
    }
    rp = bmax(rp, sh);
    ax = bmax(ax, sh);
    return rp / (1.0 + fmax(nbi, ax));
}

// the dot products of update_sigma! (src/algorithm.jl:332-364)
template <int QK, class MT>
__device__ void bsigma_dots(Inst<MT> &I, double *rec, double *t, double *sh) {
    const int32_t n = I.n, m = I.m;
    for (int32_t i = threadIdx.x; i < n; i += kBB) {
        I.P->dx[i] = I.P->x_bar[i] - I.P->last_x[i];
        I.P->dw[i] = I.P->w_bar[i] - I.P->last_w[i];
        if (m > 0) I.P->ATdy[i] = I.P->ATy_bar[i] - I.P->last_ATy[i];
    }
    for (int32_t r = threadIdx.x; r < m; r += kBB) I.P->dy[r] = I.P->y_bar[r] - I.P->last_y[r];
    __syncthreads();
    bq_apply<QK>(I, I.P->dw, I.P->dQw, t, sh);
    if (m > 0) bq_apply<QK>(I, I.P->ATdy, I.P->QATdy, t, sh);
    double a = 0, b = 0, c = 0, d = 0, e = 0, f = 0;
    for (int32_t i = threadIdx.x; i < n; i += kBB) {
        const double dxi = I.P->dx[i], dq = I.P->dQw[i], at = m > 0 ? I.P->ATdy[i] : 0.0, qat = m > 0 ? I.P->QATdy[i] : 0.0;
        a += dxi * dxi;
        b += dq * at;
        c += dq * dq;
        d += I.P->dw[i] * dq;
        e += at * qat;
    }
    for (int32_t r = threadIdx.x; r < m; r += kBB) f += I.P->dy[r] * I.P->dy[r];
    a = bsum(a, sh);
    b = bsum(b, sh);
    c = bsum(c, sh);
    d = bsum(d, sh);
    e = bsum(e, sh);
    f = bsum(f, sh);
    if (threadIdx.x == 0) {
        rec[R_DXDX] = a;
        rec[R_DQWATDY] = b;
        rec[R_DQWDQW] = c;
        rec[R_DWDQW] = d;
        rec[R_ATDYQATDY] = e;
        rec[R_DYDY] = f;
    }
    __syncthreads();
}

// do_restart! (src/algorithm.jl:522-555)
template <class MT>
__device__ void brestart(Inst<MT> &I) {
    const int32_t n = I.n, m = I.m;
    for (int32_t r = threadIdx.x; r < m; r += kBB) {
        I.P->y[r] = I.P->y_bar[r];
        I.P->last_y[r] = I.P->y_bar[r];
    }
    __syncthreads();
    for (int32_t i = threadIdx.x; i < n; i += kBB) {
        I.P->x[i] = I.P->x_bar[i];
        I.P->w[i] = I.P->w_bar[i];
        I.P->last_x[i] = I.P->x_bar[i];
        I.P->last_w[i] = I.P->w_bar[i];
        if (m > 0) {
            const double a = csr_row(I.P->atrp, I.P->atci, I.P->atv, I.P->y_bar, i);
            I.P->ATy_bar[i] = a;
            I.P->last_ATy[i] = a;
            I.P->ATy[i] = a;
        }
    }
    __syncthreads();
}

static __device__ int64_t bprint_step(int64_t k) {
    if (k <= 0) return 10;
    int64_t p = 1;
    while (p <= k / 10) p *= 10;
    return p / 10 > 10 ? p / 10 : 10;
}

// relative KKT residual from a residual record, the host's formula
// (compute_residuals!, src/algorithm.jl:37-133, operator Q: no scaling)
__device__ __forceinline__ double bkkt(const double *r, double c0, double nbi, double nci, int32_t m) {
    const double pobj = r[R_CX] + 0.5 * r[R_XQX] + c0;
    double dobj = -0.5 * r[R_XQX] + r[R_ZX] + c0;
    if (m > 0) dobj += r[R_YS];
    const double gap = fabs(pobj - dobj) / (1.0 + fmax(fabs(pobj), fabs(dobj)));
    const double errd = r[R_RD] / (1.0 + fmax(fmax(nci, r[R_ATY]), r[R_QX]));
    const double errp = m > 0 ? r[R_RP] / (1.0 + fmax(nbi, r[R_AX])) : 0.0;
    return fmax(fmax(errp, errd), gap);
}

// one check interval for every active instance
template <class MT, int QK>
__global__ __launch_bounds__(kBB) void k_batch_interval(BArgs<MT> a) {
    __shared__ double t[64];
    __shared__ double sh[kBW];
    __shared__ KSpace ks;
    const int b = a.order[blockIdx.x];
    const BCtl c = a.ctl[b];
    if (!c.active) return;
    __shared__ IPtr<MT> tab;
    Inst<MT> I = inst_of(a, b, &tab);
    double *rec = a.rec + (int64_t)b * R_COUNT;
    if (c.restart) brestart(I);
    if constexpr (QK == 0) bkspace_init(I, ks, sh);
    if (threadIdx.x == 0) rec[R_DONE] = -1.0;
    double inner = c.h_inner;
    for (int s = 0; s < c.K; ++s) {
        const int64_t it = c.iter0 + s, next = it + 1;
        const bool print_yes = next % (a.print_freq > 0 ? a.print_freq : bprint_step(next)) == 0;
        const bool full = (s == c.K - 1) || (s == 0 && c.restart) || (next % c.check_iter == 0) || print_yes;
        const double h1 = 1.0 / (inner + 2.0);
        if constexpr (QK == 0)
            bstep(I, ks, c.sigma, c.lamA, c.lamQ, h1, 1.0 - h1, full, sh);
        else
            bstep_csr(I, c.sigma, c.lamA, c.lamQ, h1, 1.0 - h1, full);
        if (s == 0 && c.want_last_gap) {
            const double g = bmnorm<QK>(I, c.sigma, c.lamA, c.lamQ, t, sh);
            if (threadIdx.x == 0) rec[R_LASTGAP] = g;
        }
        if (next % c.check_iter == 0) {
            const double g = bmnorm<QK>(I, c.sigma, c.lamA, c.lamQ, t, sh);
            if (threadIdx.x == 0) rec[R_CURGAP] = g;
        }
        inner += 1.0;
        // a residual iteration before the interval's end (a print step between
        // two check iterations): only the termination test can act there, so
        // it runs here and a converged instance retires without the host; the
        // FP64 residuals run only once the primal part is below the tolerance
        // (they write scratch vectors the step does not read, so skipping them
        // leaves the trajectory as it is)
        if (print_yes && s < c.K - 1 && (I.m == 0 || bprimal_err<QK>(I, a.nbi[b], sh) < a.tol)) {
            bresid<QK>(I, false, rec, t, sh);
            __syncthreads();
            if (bkkt(rec, a.c0[b], a.nbi[b], a.nci[b], I.m) < a.tol) {
                if (threadIdx.x == 0) rec[R_DONE] = (double)next;
                return;
            }
        }
    }
    bresid<QK>(I, false, rec, t, sh);
    bsigma_dots<QK>(I, rec, t, sh);
}

// A_i U_i of every instance (row-major m_i x k_i), from the A and U values
// the iteration reads, summed in FP64
template <class MT>
__global__ __launch_bounds__(kBB) void k_batch_au(BArgs<MT> a) {
    const int b = blockIdx.x;
    const int32_t n = a.n[b], m = a.m[b], k = a.k[b];
    const int32_t *rp = a.a_rp + a.arp_off[b];
    const MT *U = a.U + a.uoff[b];
    double *AU = a.AUk + a.auoff[b];
    for (int64_t t = threadIdx.x; t < (int64_t)m * k; t += kBB) {
        const int32_t r = (int32_t)(t / k), j = (int32_t)(t % k);
        double acc = 0.0;
        for (int32_t q = rp[r]; q < rp[r + 1]; ++q) acc += (double)a.a_v[q] * (double)U[a.a_ci[q] + (int64_t)j * n];
        AU[t] = acc;
    }
}

// iteration-0 residuals (with the bound-violation term) and warm start
template <class MT, int QK>
__global__ __launch_bounds__(kBB) void k_batch_init(BArgs<MT> a, int has_x0, int has_y0) {
    __shared__ double t[64];
    __shared__ double sh[kBW];
    const int b = a.order[blockIdx.x];
    const BCtl c = a.ctl[b];
    if (!c.active) return;
    __shared__ IPtr<MT> tab;
    Inst<MT> I = inst_of(a, b, &tab);
    double *rec = a.rec + (int64_t)b * R_COUNT;
    const int32_t n = I.n, m = I.m;
    // process_initial_points! (src/algorithm.jl:1783-1836): x0/y0 were
    // copied into x_bar / y_bar by the host, in the original space (scaled
    // instances: x col_norm, y row_norm; b_scale = c_scale = 1)
    if (has_x0) {
        for (int32_t i = threadIdx.x; i < n; i += kBB) {
            const double v = QK == 1 ? I.P->x_bar[i] * I.P->cn[i] : I.P->x_bar[i];
            I.P->x_bar[i] = v;
            I.P->x[i] = I.P->w[i] = I.P->w_bar[i] = I.P->last_x[i] = I.P->last_w[i] = v;
        }
    }
    __syncthreads();
    if (has_y0) {
        for (int32_t r = threadIdx.x; r < m; r += kBB) {
            const double v = QK == 1 ? I.P->y_bar[r] * I.P->rn[r] : I.P->y_bar[r];
            I.P->y_bar[r] = I.P->y[r] = I.P->last_y[r] = v;
        }
        __syncthreads();
        for (int32_t i = threadIdx.x; i < n; i += kBB) {
            const double v = m > 0 ? csr_row(I.P->atrp, I.P->atci, I.P->atv, I.P->y_bar, i) : 0.0;
            I.P->ATy_bar[i] = I.P->ATy[i] = I.P->last_ATy[i] = v;
        }
        __syncthreads();
        bq_apply<QK>(I, I.P->x_bar, I.P->dQw, t, sh);
        for (int32_t i = threadIdx.x; i < n; i += kBB) {
            const double z_raw = I.P->x_bar[i] + c.sigma * (-I.P->dQw[i] + I.P->ATy_bar[i] - (double)I.P->cM[i]);
            I.P->z_bar[i] = (I.P->x_bar[i] - z_raw) / c.sigma;
        }
        if (m > 0) {
            csr_rows_wave(m, I.P->arp, I.P->aci, I.P->av, I.P->x_bar, I.P->pz);
            for (int32_t r = threadIdx.x; r < m; r += kBB)
                I.P->s[r] = dmin(dmax(I.P->pz[r] - c.lamA * c.sigma * I.P->y[r], (double)I.P->ALM[r]), (double)I.P->AUM[r]);
        }
        __syncthreads();
    }
    bresid<QK>(I, true, rec, t, sh);
}

// ---------------- scaling of CSR instances -------------------------------
// scaling! (src/utils.jl:350-804) of every instance by its own block, on
// FP64 working copies of the data: nruiz Ruiz passes (statistic sqrt(max|.|))
// then, with pc, one Pock-Chambolle pass (sqrt(sum|.|)); the column
// statistic runs over A's column and Q's row, the row statistic over A's
// row; then Q = (Q + Q')/2 over the symmetric pattern, and the 2-norms of the
// scaled b and c for sigma_0.  The formulas are build_scaled's
// (hprqp_solver.hip), so a batched instance is scaled exactly as a single one.
struct BScale {
    const int32_t *n, *m;
    const int64_t *noff, *moff, *qrp_off, *arp_off, *atrp_off;
    const int32_t *q_rp, *q_ci, *q_tpos, *a_rp, *a_ci, *at_rp, *at_ci;
    double *qv, *av, *atv, *c, *l, *u, *AL, *AU, *cn, *rn, *tc, *tr, *qtmp, *rec;
    int nruiz, pc;
};

__global__ __launch_bounds__(kBB) void k_batch_scale(BScale g) {
    __shared__ double sh[kBW];
    const int b = blockIdx.x;
    const int32_t n = g.n[b], m = g.m[b];
    const int64_t no = g.noff[b], mo = g.moff[b];
    const int32_t *qrp = g.q_rp + g.qrp_off[b], *arp = g.a_rp + g.arp_off[b], *atrp = g.at_rp + g.atrp_off[b];
    double *cn = g.cn + no, *rn = g.rn + mo, *tc = g.tc + no, *tr = g.tr + mo;
    for (int32_t i = threadIdx.x; i < n; i += kBB) cn[i] = 1.0;
    for (int32_t r = threadIdx.x; r < m; r += kBB) rn[r] = 1.0;
    const int total = g.nruiz + (g.pc ? 1 : 0);
    for (int pass = 0; pass < total; ++pass) {
        const bool ruiz = pass < g.nruiz;
        for (int32_t j = threadIdx.x; j < n; j += kBB) {
            double sa = 0.0, sq = 0.0;
            for (int32_t k = atrp[j]; k < atrp[j + 1]; ++k) sa = ruiz ? dmax(sa, fabs(g.atv[k])) : sa + fabs(g.atv[k]);
            for (int32_t k = qrp[j]; k < qrp[j + 1]; ++k) sq = ruiz ? dmax(sq, fabs(g.qv[k])) : sq + fabs(g.qv[k]);
            const double t = ruiz ? dmax(sq, sa) : sq + sa;
            tc[j] = t > 0.0 ? sqrt(t) : 1.0;
        }
        for (int32_t r = threadIdx.x; r < m; r += kBB) {
            double sr = 0.0;
            for (int32_t k = arp[r]; k < arp[r + 1]; ++k) sr = ruiz ? dmax(sr, fabs(g.av[k])) : sr + fabs(g.av[k]);
            tr[r] = sr > 0.0 ? sqrt(sr) : 1.0;
        }
        __syncthreads();
        for (int32_t j = threadIdx.x; j < n; j += kBB) {
            const double t = tc[j];
            cn[j] *= t;
            g.c[no + j] /= t;
            g.l[no + j] *= t;
            g.u[no + j] *= t;
        }
        for (int32_t r = threadIdx.x; r < m; r += kBB) {
            const double t = tr[r];
            rn[r] *= t;
            g.AL[mo + r] /= t;
            g.AU[mo + r] /= t;
        }
        for (int32_t i = threadIdx.x; i < n; i += kBB) {
            const double inv = 1.0 / tc[i];
            for (int32_t k = qrp[i]; k < qrp[i + 1]; ++k) g.qv[k] = (g.qv[k] * inv) / tc[g.q_ci[k]];
        }
        for (int32_t r = threadIdx.x; r < m; r += kBB) {
            const double inv = 1.0 / tr[r];
            for (int32_t k = arp[r]; k < arp[r + 1]; ++k) g.av[k] = (g.av[k] * inv) / tc[g.a_ci[k]];
        }
        for (int32_t j = threadIdx.x; j < n; j += kBB) {
            const double inv = 1.0 / tc[j];
            for (int32_t k = atrp[j]; k < atrp[j + 1]; ++k) g.atv[k] = (g.atv[k] * inv) / tr[g.at_ci[k]];
        }
        __syncthreads();
    }
    // Q = (Q + Q')/2 (src/utils.jl:772-777)
    const int32_t q0 = qrp[0], q1 = qrp[n];
    for (int32_t k = q0 + (int32_t)threadIdx.x; k < q1; k += kBB) g.qtmp[k] = (g.qv[k] + g.qv[g.q_tpos[k]]) / 2.0;
    __syncthreads();
    for (int32_t k = q0 + (int32_t)threadIdx.x; k < q1; k += kBB) g.qv[k] = g.qtmp[k];
    // norms of the scaled b (max(|AL|, |AU|) of the finite bounds) and c
    double nb = 0.0, nc = 0.0;
    for (int32_t r = threadIdx.x; r < m; r += kBB) {
        const double lo = g.AL[mo + r], hi = g.AU[mo + r];
        const double t = dmax(fabs(lo == -INFINITY ? 0.0 : lo), fabs(hi == INFINITY ? 0.0 : hi));
        nb += t * t;
    }
    for (int32_t j = threadIdx.x; j < n; j += kBB) nc += g.c[no + j] * g.c[no + j];
    nb = bsum(nb, sh);
    nc = bsum(nc, sh);
    if (threadIdx.x == 0) {
        g.rec[(int64_t)b * R_COUNT + R_NB] = sqrt(nb);
        g.rec[(int64_t)b * R_COUNT + R_NC] = sqrt(nc);
    }
}

// ---------------- lambda_max by Lanczos ----------------------------------
// The reference estimates lambda_max(A A^T) and lambda_max(Q) with the power
// method (src/utils.jl:1335-1461: start randn + 1e-8, stop when
// |Op q - lambda q| / (|Op q| + lambda) < 1e-4, at most 5000 steps).  Batched
// instances are often factor models whose top eigenvalues are close, where
// the power method needs thousands of steps; here each block runs Lanczos
// (no reorthogonalisation) from the same start vector and applies the same
// test to the top Ritz pair (theta, y): |Op y - theta y| = beta_m |s_m|, s
// the eigenvector of the tridiagonal T_m.  theta is found by multisection
// with Sturm counts (all threads), s_m by the three-term recurrence.
constexpr int kLzMax = 512;  // Lanczos steps kept (LDS)

__device__ __forceinline__ int sturm_below(const double *al, const double *be, int m, double x) {
    double d = al[0] - x;
    int c = d < 0.0;
    for (int i = 1; i < m; ++i) {
        if (d == 0.0) d = -1e-300;
        d = al[i] - x - be[i - 1] * be[i - 1] / d;
        c += d < 0.0;
    }
    return c;
}

// largest eigenvalue of the m x m tridiagonal (al, be), lo a lower bound
__device__ double tri_max(const double *al, const double *be, int m, double lo, int *imin) {
    double hi = -INFINITY;
    for (int i = 0; i < m; ++i) {
        lo = fmax(lo, al[i]);
        hi = fmax(hi, al[i] + (i > 0 ? fabs(be[i - 1]) : 0.0) + (i + 1 < m ? fabs(be[i]) : 0.0));
    }
    for (int round = 0; round < 10 && hi - lo > 1e-15 * fmax(fabs(hi), fabs(lo)); ++round) {
        const double x = lo + (hi - lo) * (double)(threadIdx.x + 1) / (double)(kBB + 1);
        const bool above = sturm_below(al, be, m, x) == m;
        __syncthreads();
        if (threadIdx.x == 0) *imin = kBB;
        __syncthreads();
        if (above) atomicMin(imin, (int)threadIdx.x);
        __syncthreads();
        const int t = *imin;
        const double w = (hi - lo) / (double)(kBB + 1);
        const double nlo = lo + w * (double)t;
        const double nhi = t < kBB ? lo + w * (double)(t + 1) : hi;
        lo = nlo;
        hi = nhi;
    }
    return 0.5 * (lo + hi);
}

// |last component| of the unit eigenvector of T_m for eigenvalue theta
__device__ double tri_last(const double *al, const double *be, int m, double theta) {
    double xp = 0.0, xc = 1.0, ss = 1.0;
    for (int i = 0; i + 1 < m; ++i) {
        if (be[i] == 0.0) return 0.0;
        const double xn = -((al[i] - theta) * xc + (i > 0 ? be[i - 1] * xp : 0.0)) / be[i];
        xp = xc;
        xc = xn;
        ss += xn * xn;
        if (fabs(xc) > 1e100) {
            xp *= 1e-100;
            xc *= 1e-100;
            ss *= 1e-200;
        }
    }
    return fabs(xc) / sqrt(ss);
}

__device__ __forceinline__ bool lz_check(int m) {
    if (m <= 24) return m % 4 == 0;
    if (m <= 48) return m % 8 == 0;
    if (m <= 96) return m % 16 == 0;
    return m % 32 == 0;
}

// Lanczos on a dim-dimensional operator; v0 holds the start vector (it is
// overwritten), v1 and v2 are scratch of the same length
template <class Op>
__device__ double blanczos(int32_t dim, Op op, double *v0, double *v1, double *v2, double tol, int max_steps,
                           int *steps, double *sh) {
    __shared__ double al[kLzMax], be[kLzMax];
    __shared__ double bc[2];
    __shared__ int imin;
    max_steps = min(max_steps, kLzMax);
    double nz = 0.0;
    for (int32_t i = threadIdx.x; i < dim; i += kBB) nz += v0[i] * v0[i];
    nz = sqrt(bsum(nz, sh));
    for (int32_t i = threadIdx.x; i < dim; i += kBB) v0[i] /= nz;
    __syncthreads();
    double *qp = v1, *qc = v0, *w = v2;
    double beta_prev = 0.0, theta = -INFINITY;
    int j = 0;
    for (; j < max_steps; ++j) {
        op(qc, w);
        double a = 0.0;
        for (int32_t i = threadIdx.x; i < dim; i += kBB) a += qc[i] * w[i];
        const double alpha = bsum(a, sh);
        double ww = 0.0;
        for (int32_t i = threadIdx.x; i < dim; i += kBB) {
            const double v = w[i] - alpha * qc[i] - (j > 0 ? beta_prev * qp[i] : 0.0);
            w[i] = v;
            ww += v * v;
        }
        const double beta = sqrt(bsum(ww, sh));
        if (threadIdx.x == 0) {
            al[j] = alpha;
            be[j] = beta;
        }
        __syncthreads();
        const int m = j + 1;
        const bool tiny = !(beta > 1e-13 * (fabs(alpha) + beta_prev));
        if (lz_check(m) || m == max_steps || tiny) {
            theta = tri_max(al, be, m, theta, &imin);
            if (threadIdx.x == 0) {
                const double res = tiny ? 0.0 : beta * tri_last(al, be, m, theta);
                const double den = sqrt(theta * theta + res * res) + fabs(theta);
                bc[0] = (res == 0.0 || (den > 0.0 && res / den < tol)) ? 1.0 : 0.0;
            }
            __syncthreads();
            if (bc[0] != 0.0) break;
        }
        for (int32_t i = threadIdx.x; i < dim; i += kBB) w[i] /= beta;
        __syncthreads();
        double *tmp = qp;
        qp = qc;
        qc = w;
        w = tmp;
        beta_prev = beta;
    }
    *steps = j < max_steps ? j + 1 : max_steps;
    return theta;
}

// per-instance lambda_max of A A^T and Q in one launch; start vectors were
// written by the host (m-part in pz, n-part in dQw)
template <class MT, int QK>
__global__ __launch_bounds__(kBB) void k_batch_power(BArgs<MT> a, double tol, int max_it) {
    __shared__ double t[64];
    __shared__ double sh[kBW];
    const int b = a.order[blockIdx.x];
    if (!a.ctl[b].active) return;
    __shared__ IPtr<MT> tab;
    Inst<MT> I = inst_of(a, b, &tab);
    double *rec = a.rec + (int64_t)b * R_COUNT;
    const int32_t n = I.n, m = I.m, k = I.k;
    double lam = 0.0;
    int steps = 0;
    if (m > 0) {
        // A A^T v = A (A^T v), A^T v into pq (n)
        auto opA = [&](const double *v, double *out) {
            for (int32_t i = threadIdx.x; i < n; i += kBB) I.P->pq[i] = csr_row(I.P->atrp, I.P->atci, I.P->atv, v, i);
            __syncthreads();
            csr_rows_wave(m, I.P->arp, I.P->aci, I.P->av, I.P->pq, out);
        };
        lam = blanczos(m, opA, I.P->pz, I.P->s, I.P->pz2, tol, max_it, &steps, sh);
    }
    if (threadIdx.x == 0) {
        rec[R_POW_A] = lam;
        rec[R_POW_ITA] = (double)steps;
    }
    auto opQ = [&](const double *v, double *out) { bq_apply<QK>(I, v, out, t, sh); };
    const double lq = blanczos(n, opQ, I.P->dQw, I.P->pq, I.P->pq2, tol, max_it, &steps, sh);
    if (threadIdx.x == 0) {
        rec[R_POW_Q] = lq;
        rec[R_POW_ITQ] = (double)steps;
    }
}

// ----------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------
static uint64_t bmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// same counter-based N(0,1) stream as the single-instance power method
static void brandn(uint64_t seed, int64_t count, double *out) {
    const double two_pi = 6.283185307179586476925286766559;
    const uint64_t ks = bmix64(seed + 0x9E3779B97F4A7C15ULL);
    for (int64_t i = 0; i < count; ++i) {
        const uint64_t k0 = 2 * (uint64_t)i, k1 = k0 + 1;
        const double a = (double)(bmix64(ks ^ (k0 * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL)) >> 11) *
                         (1.0 / 9007199254740992.0);
        const double b = (double)(bmix64(ks ^ (k1 * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL)) >> 11) *
                         (1.0 / 9007199254740992.0);
        out[i] = std::sqrt(-2.0 * std::log(1.0 - a)) * std::cos(two_pi * b);
    }
}

static double bgolden(double ap, double bp, double cp, double dp) {
    const double phi = (std::sqrt(5.0) - 1.0) / 2.0;
    auto f = [&](double x) { return ap * x + bp / x + cp * x * x / (1 + dp * x); };
    double a = 1e-12, b = 1e12;
    double c = b - phi * (b - a), d = a + phi * (b - a);
    double fc = f(c), fd = f(d);
    for (int i = 1; i <= 200; ++i) {
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
        if ((b - a) < 1e-13) break;
    }
    return 0.5 * (a + b);
}

struct BatchWork {
    DevVec<double> vn[21], vm[7];
    DevVec<double> au;  // A_i U_i of every instance
    DevVec<BCtl> ctl;
    DevVec<double> rec;
    // CSR kind: the scaled data of this solve (FP64, and FP32 for the mixed
    // form), the scaling's column and row norms, scratch of the (Q + Q')/2
    DevVec<double> sq, sa, sat, sc, sl, su, sAL, sAU, cn, rn, qtmp;
    DevVec<float> sq32, sa32, sat32, sc32, sl32, su32, sAL32, sAU32;
};

template <class F>
static void beach(hipStream_t s, int64_t n, const F &f) {
    if (n <= 0) return;
    k_each<F><<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(n, f);
    HPRQP_HIP(hipGetLastError());
}

// the CSR kind's per-solve scaling: working copies of the data, scaled on
// the device by k_batch_scale (column / row statistics' scratch in tc, tr)
static void batch_scale_data(hipStream_t s, hprqp_batch *B, BatchWork &W, const hprqp_params *p, bool mixed,
                             double *tc, double *tr) {
    const int64_t N = B->Ntot, M = B->Mtot;
    auto copy = [&](DevVec<double> &dst, const DevVec<double> &src, int64_t len) {
        dst.alloc(std::max<int64_t>(len, 1));
        if (len) HPRQP_HIP(hipMemcpyAsync(dst.p, src.p, sizeof(double) * len, hipMemcpyDeviceToDevice, s));
    };
    copy(W.sq, B->q_v, B->nnzQ);
    copy(W.sa, B->a_v, B->nnzA);
    copy(W.sat, B->at_v, B->nnzA);
    copy(W.sc, B->c, N);
    copy(W.sl, B->l, N);
    copy(W.su, B->u, N);
    copy(W.sAL, B->AL, M);
    copy(W.sAU, B->AU, M);
    W.cn.alloc(std::max<int64_t>(N, 1));
    W.rn.alloc(std::max<int64_t>(M, 1));
    W.qtmp.alloc(std::max<int64_t>(B->nnzQ, 1));
    BScale g{};
    g.n = B->dn.p;
    g.m = B->dm.p;
    g.noff = B->dnoff.p;
    g.moff = B->dmoff.p;
    g.qrp_off = B->dqrp_off.p;
    g.arp_off = B->darp_off.p;
    g.atrp_off = B->datrp_off.p;
    g.q_rp = B->q_rp.p;
    g.q_ci = B->q_ci.p;
    g.q_tpos = B->q_tpos.p;
    g.a_rp = B->a_rp.p;
    g.a_ci = B->a_ci.p;
    g.at_rp = B->at_rp.p;
    g.at_ci = B->at_ci.p;
    g.qv = W.sq.p;
    g.av = W.sa.p;
    g.atv = W.sat.p;
    g.c = W.sc.p;
    g.l = W.sl.p;
    g.u = W.su.p;
    g.AL = W.sAL.p;
    g.AU = W.sAU.p;
    g.cn = W.cn.p;
    g.rn = W.rn.p;
    g.tc = tc;
    g.tr = tr;
    g.qtmp = W.qtmp.p;
    g.rec = W.rec.p;
    g.nruiz = p->use_Ruiz_scaling ? 10 : 0;
    g.pc = p->use_Pock_Chambolle_scaling ? 1 : 0;
    k_batch_scale<<<B->count, kBB, 0, s>>>(g);
    HPRQP_HIP(hipGetLastError());
    if (mixed) {
        auto to32 = [&](DevVec<float> &dst, const DevVec<double> &src, int64_t len) {
            dst.alloc(std::max<int64_t>(len, 1));
            float *d = dst.p;
            const double *a = src.p;
            beach(s, len, [=] __device__(int64_t i) { d[i] = (float)a[i]; });
        };
        to32(W.sq32, W.sq, B->nnzQ);
        to32(W.sa32, W.sa, B->nnzA);
        to32(W.sat32, W.sat, B->nnzA);
        to32(W.sc32, W.sc, N);
        to32(W.sl32, W.sl, N);
        to32(W.su32, W.su, N);
        to32(W.sAL32, W.sAL, M);
        to32(W.sAU32, W.sAU, M);
    }
}

template <class MT>
static BArgs<MT> make_args(hprqp_batch *B, BatchWork &W) {
    BArgs<MT> a{};
    a.count = B->count;
    a.order = B->dorder.p;
    a.n = B->dn.p;
    a.m = B->dm.p;
    a.k = B->dk.p;
    a.noff = B->dnoff.p;
    a.moff = B->dmoff.p;
    a.uoff = B->duoff.p;
    a.arp_off = B->darp_off.p;
    a.atrp_off = B->datrp_off.p;
    a.U64 = B->U.p;
    a.d64 = B->d.p;
    a.a_rp = B->a_rp.p;
    a.a_ci = B->a_ci.p;
    a.at_rp = B->at_rp.p;
    a.at_ci = B->at_ci.p;
    a.a_v64 = B->a_v.p;
    a.at_v64 = B->at_v.p;
    a.c = B->c.p;
    a.l = B->l.p;
    a.u = B->u.p;
    a.AL = B->AL.p;
    a.AU = B->AU.p;
    a.c0 = B->dc0.p;
    a.nbi = B->dnbi.p;
    a.nci = B->dnci.p;
    if (B->qkind == 1) {
        // the scaled data of this solve
        a.a_v64 = W.sa.p;
        a.at_v64 = W.sat.p;
        a.c = W.sc.p;
        a.l = W.sl.p;
        a.u = W.su.p;
        a.AL = W.sAL.p;
        a.AU = W.sAU.p;
        a.q_rp = B->q_rp.p;
        a.q_ci = B->q_ci.p;
        a.qrp_off = B->dqrp_off.p;
        a.q_v64 = W.sq.p;
        a.cn = W.cn.p;
        a.rn = W.rn.p;
        if constexpr (std::is_same<MT, double>::value) {
            a.q_v = W.sq.p;
            a.a_v = W.sa.p;
            a.at_v = W.sat.p;
            a.cM = W.sc.p;
            a.lM = W.sl.p;
            a.uM = W.su.p;
            a.ALM = W.sAL.p;
            a.AUM = W.sAU.p;
        } else {
            a.q_v = W.sq32.p;
            a.a_v = W.sa32.p;
            a.at_v = W.sat32.p;
            a.cM = W.sc32.p;
            a.lM = W.sl32.p;
            a.uM = W.su32.p;
            a.ALM = W.sAL32.p;
            a.AUM = W.sAU32.p;
        }
    }
    if (B->qkind == 1) {
        // filled after the scaling below
    } else if constexpr (std::is_same<MT, double>::value) {
        a.U = B->U.p;
        a.d = B->d.p;
        a.a_v = B->a_v.p;
        a.at_v = B->at_v.p;
        a.cM = B->c.p;
        a.lM = B->l.p;
        a.uM = B->u.p;
        a.ALM = B->AL.p;
        a.AUM = B->AU.p;
    } else {
        a.U = B->U32.p;
        a.d = B->d32.p;
        a.a_v = B->a_v32.p;
        a.at_v = B->at_v32.p;
        a.cM = B->c32.p;
        a.lM = B->l32.p;
        a.uM = B->u32.p;
        a.ALM = B->AL32.p;
        a.AUM = B->AU32.p;
    }
    double **np[] = {&a.x, &a.w, &a.w_bar, &a.x_hat, &a.x_bar, &a.z_bar, &a.dx, &a.dw, &a.Qw, &a.Qw_bar,
                     &a.tempv, &a.ATy, &a.ATy_bar, &a.ATdy, &a.QATdy, &a.dQw, &a.last_x, &a.last_w, &a.last_ATy,
                     &a.pq, &a.pq2};
    for (int i = 0; i < 21; ++i) *np[i] = W.vn[i].p;
    double **mp[] = {&a.y, &a.y_bar, &a.s, &a.dy, &a.last_y, &a.pz, &a.pz2};
    for (int i = 0; i < 7; ++i) *mp[i] = W.vm[i].p;
    a.AUk = W.au.p;
    a.auoff = B->dauoff.p;
    a.ctl = W.ctl.p;
    a.rec = W.rec.p;
    return a;
}

struct BInstState {
    int64_t restart_flag = 0;
    bool first_restart = true;
    double last_gap = INFINITY, current_gap = INFINITY, save_gap = INFINITY;
    int64_t inner = 0, times = 0;
    double weighted_norm = INFINITY, best_gap = INFINITY, best_sigma = 1.0;
    double sigma = 1.0, lamA = 0.0, lamQ = 0.0;
    int status = -2;
    int64_t iter = 0, power_it = 0;
    double kkt = INFINITY, pobj = 0, dobj = 0, gap = 0, errp = 0, errd = 0;
};

template <class MT>
static void batch_solve_impl(hprqp_ctx *ctx, hprqp_batch *B, const hprqp_params *p, const double *x0,
                             const double *y0, const double *x0_dev, const double *y0_dev, double *x_out, double *y_out, int32_t *status_out, int64_t *iter_out,
                             double *obj_out, double *kkt_out, hprqp_batch_result *res) {
    hipStream_t s = ctx->stream;
    const int32_t cnt = B->count;
    const int64_t N = B->Ntot, M = B->Mtot;
    BatchWork W;
    for (auto &v : W.vn) {
        v.alloc(std::max<int64_t>(N, 1));
        HPRQP_HIP(hipMemsetAsync(v.p, 0, sizeof(double) * v.n, s));
    }
    for (auto &v : W.vm) {
        v.alloc(std::max<int64_t>(M, 1));
        HPRQP_HIP(hipMemsetAsync(v.p, 0, sizeof(double) * v.n, s));
    }
    W.au.alloc(std::max<int64_t>(B->AUtot, 1));
    W.ctl.alloc(cnt);
    W.rec.alloc((int64_t)cnt * R_COUNT);
    HPRQP_HIP(hipMemsetAsync(W.rec.p, 0, sizeof(double) * W.rec.n, s));
    const bool csrq = B->qkind == 1;
    // CSR kind: scale every instance (the statistics' scratch: pq, pz, which
    // the power method overwrites later)
    if (csrq) batch_scale_data(s, B, W, p, !std::is_same<MT, double>::value, W.vn[19].p, W.vm[5].p);
    BArgs<MT> a = make_args<MT>(B, W);
    if (!csrq) {
        k_batch_au<MT><<<cnt, kBB, 0, s>>>(a);
        HPRQP_HIP(hipGetLastError());
    }
    a.tol = p->stoptol;
    a.print_freq = p->print_frequency;
    // pinned staging for the per-instance control and records
    BCtl *hctl = nullptr;
    double *hrec = nullptr;
    HPRQP_HIP(hipHostMalloc((void **)&hctl, sizeof(BCtl) * cnt, hipHostMallocDefault));
    HPRQP_HIP(hipHostMalloc((void **)&hrec, sizeof(double) * R_COUNT * cnt, hipHostMallocDefault));
    struct PinnedFree {
        void *a, *b;
        ~PinnedFree() {
            (void)hipHostFree(a);
            (void)hipHostFree(b);
        }
    } pf{hctl, hrec};
    std::vector<BInstState> st(cnt);
    std::memset(hctl, 0, sizeof(BCtl) * cnt);
    for (int b = 0; b < cnt; ++b) {
        hctl[b].active = 1;
        hctl[b].check_iter = (int32_t)p->check_iter;
    }
    auto put_ctl = [&] {
        HPRQP_HIP(hipMemcpyAsync(W.ctl.p, hctl, sizeof(BCtl) * cnt, hipMemcpyHostToDevice, s));
    };
    auto get_rec = [&] {
        HPRQP_HIP(hipMemcpyAsync(hrec, W.rec.p, sizeof(double) * R_COUNT * cnt, hipMemcpyDeviceToHost, s));
        HPRQP_HIP(hipStreamSynchronize(s));
    };
    const double t_start = bnow();
    // power method start vectors: z (m) in pz, z (n) in dQw
    {
        std::vector<double> zm(std::max<int64_t>(M, 1)), zn(std::max<int64_t>(N, 1));
        for (int b = 0; b < cnt; ++b) {
            brandn(1, B->m[b], zm.data() + B->moff[b]);
            for (int32_t r = 0; r < B->m[b]; ++r) zm[B->moff[b] + r] += 1e-8;
            brandn(1, B->n[b], zn.data() + B->noff[b]);
            for (int32_t i = 0; i < B->n[b]; ++i) zn[B->noff[b] + i] += 1e-8;
        }
        HPRQP_HIP(hipMemcpyAsync(a.pz, zm.data(), sizeof(double) * M, hipMemcpyHostToDevice, s));
        HPRQP_HIP(hipMemcpyAsync(a.dQw, zn.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
        put_ctl();
        if (csrq)
            k_batch_power<MT, 1><<<cnt, kBB, 0, s>>>(a, 1e-4, 5000);
        else
            k_batch_power<MT, 0><<<cnt, kBB, 0, s>>>(a, 1e-4, 5000);
        HPRQP_HIP(hipGetLastError());
        get_rec();
    }
    for (int b = 0; b < cnt; ++b) {
        BInstState &S = st[b];
        const double *r = hrec + (int64_t)b * R_COUNT;
        S.lamA = B->m[b] > 0 ? r[R_POW_A] * p->eig_factor : 0.0;
        S.lamQ = r[R_POW_Q] * p->eig_factor;
        S.power_it = (int64_t)(r[R_POW_ITA] + r[R_POW_ITQ]);
        // sigma_0 (src/algorithm.jl:839-851) from the norms of the data the
        // iteration sees (the operator kind is not scaled)
        const double nb = csrq ? r[R_NB] : B->norm_b2[b], nc = csrq ? r[R_NC] : B->norm_c2[b];
        S.sigma = (p->sigma == -1) ? ((nc > 1e-16 && nb > 1e-16 && nb < 1e16 && nc < 1e16) ? nb / nc : 1.0)
                                   : p->sigma;
        S.best_sigma = S.sigma;
        hctl[b].sigma = S.sigma;
        hctl[b].lamA = S.lamA;
        hctl[b].lamQ = S.lamQ;
    }
    res->power_time = bnow() - t_start;
    // warm start: x0 / y0 go to x_bar / y_bar, the init kernel spreads them
    if (x0_dev)
        HPRQP_HIP(hipMemcpyAsync(a.x_bar, x0_dev, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    else if (x0)
        HPRQP_HIP(hipMemcpyAsync(a.x_bar, x0, sizeof(double) * N, hipMemcpyHostToDevice, s));
    if (y0_dev && M)
        HPRQP_HIP(hipMemcpyAsync(a.y_bar, y0_dev, sizeof(double) * M, hipMemcpyDeviceToDevice, s));
    else if (y0 && M)
        HPRQP_HIP(hipMemcpyAsync(a.y_bar, y0, sizeof(double) * M, hipMemcpyHostToDevice, s));
    put_ctl();
    {
        const int hx = x0 != nullptr || x0_dev != nullptr, hy = y0 != nullptr || y0_dev != nullptr;
        if (csrq)
            k_batch_init<MT, 1><<<cnt, kBB, 0, s>>>(a, hx, hy);
        else
            k_batch_init<MT, 0><<<cnt, kBB, 0, s>>>(a, hx, hy);
    }
    HPRQP_HIP(hipGetLastError());
    get_rec();
    // one check interval = ctl upload + interval kernel + record download,
    // captured once in a HIP graph and replayed (ctl/rec live in pinned
    // memory at fixed addresses; the interval length travels in ctl)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    HPRQP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    HPRQP_HIP(hipMemcpyAsync(W.ctl.p, hctl, sizeof(BCtl) * cnt, hipMemcpyHostToDevice, s));
    if (csrq)
        k_batch_interval<MT, 1><<<cnt, kBB, 0, s>>>(a);
    else
        k_batch_interval<MT, 0><<<cnt, kBB, 0, s>>>(a);
    HPRQP_HIP(hipMemcpyAsync(hrec, W.rec.p, sizeof(double) * R_COUNT * cnt, hipMemcpyDeviceToHost, s));
    HPRQP_HIP(hipStreamEndCapture(s, &graph));
    HPRQP_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    struct GraphFree {
        hipGraph_t g;
        hipGraphExec_t e;
        ~GraphFree() {
            if (e) (void)hipGraphExecDestroy(e);
            if (g) (void)hipGraphDestroy(g);
        }
    } gf{graph, gexec};

    // each replay carries an event pair: the summed spans are the time of
    // the interval kernels (with the two small copies of the replay)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    HPRQP_HIP(hipEventCreateWithFlags(&ev0, hipEventDisableSystemFence));
    HPRQP_HIP(hipEventCreateWithFlags(&ev1, hipEventDisableSystemFence));
    struct EventFree {
        hipEvent_t a, b;
        ~EventFree() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } evf{ev0, ev1};
    double interval_ms = 0.0;

    int64_t it = 0, intervals = 0, solved = 0;
    int active = cnt;
    const double tol = p->stoptol;
    while (active > 0) {
        // decisions at residual iteration `it` for every active instance
        for (int b = 0; b < cnt; ++b) {
            BInstState &S = st[b];
            if (!hctl[b].active) continue;
            const double *r = hrec + (int64_t)b * R_COUNT;
            // converged at a residual iteration inside the last interval
            const bool done_early = it > 0 && r[R_DONE] >= 0.0;
            const int64_t rit = done_early ? (int64_t)r[R_DONE] : it;
            // residuals (operator Q: no scaling, b_scale = c_scale = 1)
            S.pobj = r[R_CX] + 0.5 * r[R_XQX] + B->obj_constant[b];
            S.dobj = -0.5 * r[R_XQX] + r[R_ZX] + B->obj_constant[b];
            if (B->m[b] > 0) S.dobj += r[R_YS];
            S.gap = std::fabs(S.pobj - S.dobj) / (1.0 + std::max(std::fabs(S.pobj), std::fabs(S.dobj)));
            S.errd = r[R_RD] / (1.0 + std::max(std::max(B->norm_c_inf[b], r[R_ATY]), r[R_QX]));
            S.errp = B->m[b] > 0 ? r[R_RP] / (1.0 + std::max(B->norm_b_inf[b], r[R_AX])) : 0.0;
            if (it == 0) S.errp = std::max(S.errp, r[R_BND]);
            S.kkt = std::max(std::max(S.errp, S.errd), S.gap);
            S.iter = rit;
            int stat = -2;
            if (S.kkt < tol || done_early)
                stat = HPRQP_STATUS_OPTIMAL;
            else if (it >= p->max_iter)
                stat = HPRQP_STATUS_MAX_ITER;
            else if (bnow() - t_start > p->time_limit)
                stat = HPRQP_STATUS_TIME_LIMIT;
            if (stat != -2) {
                S.status = stat;
                hctl[b].active = 0;
                --active;
                if (stat == HPRQP_STATUS_OPTIMAL) ++solved;
                continue;
            }
            // check_restart (src/algorithm.jl:441-492)
            S.restart_flag = 0;
            if (S.first_restart) {
                if (it == p->check_iter) {
                    S.first_restart = false;
                    S.restart_flag = 1;
                    S.weighted_norm = S.current_gap;
                    S.best_gap = S.current_gap;
                    S.best_sigma = S.sigma;
                }
            } else if (it % p->check_iter == 0) {
                if (S.current_gap < 0) S.current_gap = 1e-6;
                if (S.current_gap <= 0.2 * S.last_gap) S.restart_flag = 1;
                if (S.current_gap <= 0.8 * S.last_gap && S.current_gap > S.save_gap) S.restart_flag = 2;
                const double fact = (S.current_gap / S.weighted_norm > 1e-1) ? 0.5 : 0.2;
                if ((double)S.inner >= fact * (double)it) S.restart_flag = 3;
                if (S.best_gap > S.current_gap) {
                    S.best_gap = S.current_gap;
                    S.best_sigma = S.sigma;
                }
                S.save_gap = S.current_gap;
            }
            // update_sigma! (src/algorithm.jl:323-423), non-diagonal operator Q
            if (!p->sigma_fixed && S.restart_flag >= 1) {
                const int32_t m = B->m[b];
                double aa = 0.0, bb = std::max(r[R_DXDX], 1e-12), cc = 0.0, dd = 0.0;
                if (m > 0) aa = S.lamA * r[R_DYDY] - 2 * r[R_DQWATDY];
                aa += S.lamQ * r[R_DWDQW];
                if (m > 0) {
                    cc = r[R_ATDYQATDY];
                    dd = S.lamQ;
                }
                aa = std::max(aa, 1e-12);
                const double est = m > 0 ? bgolden(aa, bb, cc, dd) : std::sqrt(bb / aa);
                const double fact = std::exp(-0.05 * (S.current_gap / S.best_gap));
                const double temp_1 = std::max(std::min(S.errd, S.errp), std::min(S.gap, S.current_gap));
                const double cand = std::exp(fact * std::log(est) + (1 - fact) * std::log(S.best_sigma));
                auto clampj = [](double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); };
                double kappa = 1.0;
                if (temp_1 <= 9e-10)
                    kappa = temp_1 > 5e-10 ? clampj(std::sqrt(S.errd / S.errp), 1e-2, 100.0)
                                           : clampj(S.errd / S.errp, 1e-2, 100.0);
                S.sigma = kappa * cand;
            }
            // do_restart! bookkeeping (the vectors are reset on the device)
            hctl[b].restart = S.restart_flag > 0;
            if (S.restart_flag > 0) {
                S.last_gap = S.current_gap;
                S.save_gap = INFINITY;
                S.times++;
                S.inner = 0;
            }
            hctl[b].sigma = S.sigma;
            hctl[b].h_inner = (double)S.inner;
            hctl[b].want_last_gap = S.restart_flag > 0;
            hctl[b].iter0 = it;
        }
        if (active == 0) break;
        // interval to the next check iteration (the print steps in between
        // are handled on the device: only termination acts there)
        int64_t K = 1;
        while (!(((it + K) % p->check_iter) == 0 || it + K >= p->max_iter)) ++K;
        for (int b = 0; b < cnt; ++b)
            if (hctl[b].active) hctl[b].K = (int32_t)K;
        HPRQP_HIP(hipEventRecord(ev0, s));
        HPRQP_HIP(hipGraphLaunch(gexec, s));
        HPRQP_HIP(hipEventRecord(ev1, s));
        HPRQP_HIP(hipStreamSynchronize(s));
        {
            float ms = 0.f;
            HPRQP_HIP(hipEventElapsedTime(&ms, ev0, ev1));
            interval_ms += ms;
        }
        ++intervals;
        for (int b = 0; b < cnt; ++b) {
            if (!hctl[b].active) continue;
            BInstState &S = st[b];
            const double *r = hrec + (int64_t)b * R_COUNT;
            if (hctl[b].want_last_gap) S.last_gap = r[R_LASTGAP];
            if ((it + K) % p->check_iter == 0) S.current_gap = r[R_CURGAP];
            S.inner += K;
            hctl[b].restart = 0;
            hctl[b].want_last_gap = 0;
        }
        it += K;
    }
    HPRQP_HIP(hipStreamSynchronize(s));
    res->time = bnow() - t_start;
    res->solved = solved;
    res->intervals = intervals;
    res->interval_time = 1e-3 * interval_ms;
    res->max_iter_reached = it;
    int64_t tot_it = 0;
    for (int b = 0; b < cnt; ++b) {
        const BInstState &S = st[b];
        tot_it += S.iter;
        if (status_out) status_out[b] = S.status;
        if (iter_out) iter_out[b] = S.iter;
        if (obj_out) obj_out[b] = S.pobj;
        if (kkt_out) kkt_out[b] = S.kkt;
    }
    res->total_iter = tot_it;
    if (csrq) {
        // results in the original space (src/algorithm.jl:630-640; b_scale =
        // c_scale = 1): x = x_bar / col_norm, y = y_bar / row_norm
        double *xb = a.x_bar, *yb = a.y_bar;
        const double *cn = W.cn.p, *rn = W.rn.p;
        beach(s, N, [=] __device__(int64_t i) { xb[i] /= cn[i]; });
        beach(s, M, [=] __device__(int64_t i) { yb[i] /= rn[i]; });
        HPRQP_HIP(hipStreamSynchronize(s));
    }
    if (x_out) HPRQP_HIP(hipMemcpy(x_out, a.x_bar, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (y_out && M) HPRQP_HIP(hipMemcpy(y_out, a.y_bar, sizeof(double) * M, hipMemcpyDeviceToHost));
}

void batch_solve(hprqp_ctx *ctx, hprqp_batch *B, const hprqp_params *p, const double *x0, const double *y0,
                 const double *x0_dev, const double *y0_dev, double *x_out, double *y_out, int32_t *status_out,
                 int64_t *iter_out, double *obj_out, double *kkt_out, hprqp_batch_result *res) {
    std::memset(res, 0, sizeof(*res));
    HPRQP_REQUIRE(p->check_iter > 0, HPRQP_E_ARG, "check_iter must be positive");
    if (p->precision == HPRQP_PREC_FP64)
        batch_solve_impl<double>(ctx, B, p, x0, y0, x0_dev, y0_dev, x_out, y_out, status_out, iter_out, obj_out,
                                 kkt_out, res);
    else
        batch_solve_impl<float>(ctx, B, p, x0, y0, x0_dev, y0_dev, x_out, y_out, status_out, iter_out, obj_out,
                                kkt_out, res);
}

}  // namespace hprqp
