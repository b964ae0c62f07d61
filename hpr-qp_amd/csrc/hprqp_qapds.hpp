// hprqp_qapds.hpp -- the QAP relaxation's iteration with one operator apply
// per step.
//
// The relaxation's constraints are X 1 = 1 (rows 0..n-1 of A) and
// X^T 1 = 1 (rows n..2n-1), X = reshape(x, n, n) column-major
// (build_from_ABST, src/utils.jl:1133-1180).  So A^T y, as a matrix, is
// G_ij = a_i + b_j with a = y[0:n], b = y[n:2n], and its image under
// Q(X) = 2(A X B - S X - X T) (src/Q_operators/QAP_operator.jl:52-76) is
// rank-structured:
//   Q(G)_ij = 2 (P_i cB_j + rA_i R_j - Sa_i - rS_i b_j - a_i cT_j - Tb_j)
// with P = A a, R = B^T b, Sa = S a, Tb = T^T b for each y and the
// constants rA = A 1, cB = B^T 1, rS = S 1, cT = T^T 1 (diagonal S, T:
// S a = s .* a and so on): four n x n matrix-vector products instead of an
// apply.
//
// The HPR step (src/kernels.jl:422-537, 674-705, 760-862, 878-991) needs
// Q w and Q w_bar.  Everything in its second w update is linear in
// A^T y_bar, so with Qg = Q A^T y_bar and the products of the Halpern
// anchors held alongside,
//   Q w_bar' = Q w_bar + fs (Qg - Q ATy)
//   Q w      = h1 Q last_w   + h2 (2 Q w_bar' - Q w)
//   Q ATy    = h1 Q last_ATy + h2 (2 Qg - Q ATy)
// the products follow the iterates exactly in exact arithmetic, and only
// one apply per iteration remains.  It is taken of d = w - x_hat, since
// w_bar = f1 w + f2 x_hat gives Q w - Q w_bar = f2 Q d: tempv =
// x_hat + sigma f2 Q d and Q w_bar = Q w - f2 Q d.  The apply's rounding then
// scales with |w - x_hat|, which shrinks as the iteration converges, so the
// split-BF16 x3 apply of the reduced-precision forms perturbs tempv no more
// than the reference's two applies of nearby vectors do.  The
// w update of one iteration and the z/x/w_bar update of the next then run
// as one elementwise kernel (both are elementwise once Q w is known).  At a
// restart w = w_bar' and ATy = A^T y_bar, whose products the full-step w
// update has just formed, so no apply is needed there either; the products
// are formed by applies only when the iterates are set from outside the
// iteration (start, warm start, polish hand-over, a step-level caller's
// half-steps).
#pragma once
#include "hprqp_kernels.hpp"

namespace hprqp {

// out[c][i] = sum_k M_c[i + k n] v_c[k] for the column-major n x n matrices
// M_0..M_{cnt-1}; a 256-thread block owns 64 outputs of one matrix, its four
// waves take a quarter of k each and are summed in wave order
// (deterministic).  v_c null: the all-ones vector.
template <class VT>
struct QapGemv {
    int32_t n, cnt;
    const double *M[4];
    const VT *v[4];
    double *out[4];
};

template <class VT>
__global__ __launch_bounds__(256) void k_qap_gemv(QapGemv<VT> g) {
    __shared__ double red[4][64];
    const int n = g.n, c = blockIdx.y;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 64 + lane;
    const int kq = (n + 3) / 4, k0 = w * kq, k1 = min(n, k0 + kq);
    const double *__restrict__ M = g.M[c];
    const VT *__restrict__ v = g.v[c];
    const int ii = i < n ? i : 0;
    double acc = 0.0;
    int k = k0;
    for (; k + 4 <= k1; k += 4) {
        double m4[4], v4[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            m4[t] = M[ii + (int64_t)(k + t) * n];
            v4[t] = v ? (double)v[k + t] : 1.0;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) acc = fma(m4[t], v4[t], acc);
    }
    for (; k < k1; ++k) acc = fma(M[ii + (int64_t)k * n], v ? (double)v[k] : 1.0, acc);
    red[w][lane] = acc;
    __syncthreads();
    if (w == 0 && i < n) g.out[c][i] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// the products of y_bar (or of ones, for the constants): P = A a, R = B^T b,
// and Sa = S a, Tb = T^T b for a full S / T; Bt, Tt are the column-major
// transposes of B and T
template <class VT>
static void qap_ds_gemv(hipStream_t s, int32_t n, const double *A, const double *Bt, const double *S,
                        const double *Tt, const VT *yb, double *P, double *R, double *Sa, double *Tb) {
    QapGemv<VT> g{};
    g.n = n;
    g.M[0] = A;
    g.v[0] = yb;
    g.out[0] = P;
    g.M[1] = Bt;
    g.v[1] = yb ? yb + n : nullptr;
    g.out[1] = R;
    g.cnt = 2;
    if (S) {
        g.M[g.cnt] = S;
        g.v[g.cnt] = yb;
        g.out[g.cnt++] = Sa;
    }
    if (Tt) {
        g.M[g.cnt] = Tt;
        g.v[g.cnt] = yb ? yb + n : nullptr;
        g.out[g.cnt++] = Tb;
    }
    launch_k(k_qap_gemv<VT>, dim3((unsigned)((n + 63) / 64), (unsigned)g.cnt), dim3(256), 0, s, g);
    HPRQP_HIP(hipGetLastError());
}

// The fused step.  W2 (0 none, 1 partial, 2 full): the second w update of
// the pending iteration (EpiW2's formulas, with A^T y_bar formed from y_bar
// and the products carried); ZX (0 none, 1 partial, 2 full): the z/x/w_bar
// update of the next iteration (EpiZxw1's formulas) on the new w, ATy, Q w.
template <class VT, class DT>
struct QapDsStep {
    int32_t n;  // nq; the iterates have n * n entries
    // y_bar, its products, the constants; Sa / Tb null: diagonal S / T
    const VT *yb;
    const double *P, *R, *Sa, *Tb, *rA, *cB, *rS, *cT, *sd, *td;
    // the operator's A, B^T, S, T^T (column-major; S, T^T null when
    // diagonal) for the tile kernel's own products of y_bar
    const double *gA, *gBt, *gS, *gTt;
    // second w update (scalars of the pending iteration)
    const double *last_w, *last_ATy, *last_Qw, *last_QATy;
    VT *w, *w_bar, *ATy, *Qw, *QATy;
    const VT *Qw_bar;
    VT *dw, *ATdy, *ATy_bar, *Qwb_new, *Qg;  // full step: the restart's products
    double fs, wh1, wh2;
    // z/x/w_bar update
    const DT *c, *l, *u;
    double lS, uS;
    const double *last_x;
    VT *x, *x_hat, *dx, *x_bar, *z_bar, *d;  // d = w - x_hat, the apply's input
    double sigma, f1s, f2s, h1, h2;
};

template <class VT, class DT, int W2, int ZX>
__global__ __launch_bounds__(kBlock) void k_qap_ds_step(QapDsStep<VT, DT> e) {
    const int64_t nn = (int64_t)e.n * e.n;
    const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (o >= nn) return;
    double wi, atyi, qwi;
    if constexpr (W2 != 0) {
        const int32_t i = (int32_t)(o % e.n), j = (int32_t)(o / e.n);
        const double a = (double)e.yb[i], b = (double)e.yb[e.n + j];
        const double g = a + b;  // (A^T y_bar)_o: the two nonzeros of column o, in row order
        const double sa = e.Sa ? e.Sa[i] : e.sd[i] * a, rs = e.Sa ? e.rS[i] : e.sd[i];
        const double tb = e.Tb ? e.Tb[j] : e.td[j] * b, ct = e.Tb ? e.cT[j] : e.td[j];
        const double qg = 2.0 * (e.P[i] * e.cB[j] + e.rA[i] * e.R[j] - sa - rs * b - a * ct - tb);
        const double w0 = (double)e.w[o], wb = (double)e.w_bar[o], t0 = (double)e.ATy[o];
        const double qw0 = (double)e.Qw[o], qt0 = (double)e.QATy[o];
        const double wbn = wb + e.fs * (g - t0);
        const double wn = e.wh1 * e.last_w[o] + e.wh2 * (2.0 * wbn - w0);
        const double atyn = e.wh1 * e.last_ATy[o] + e.wh2 * (2.0 * g - t0);
        const double qwbn = (double)e.Qw_bar[o] + e.fs * (qg - qt0);
        const double qwn = e.wh1 * e.last_Qw[o] + e.wh2 * (2.0 * qwbn - qw0);
        const double qtn = e.wh1 * e.last_QATy[o] + e.wh2 * (2.0 * qg - qt0);
        e.w[o] = (VT)wn;
        e.ATy[o] = (VT)atyn;
        e.Qw[o] = (VT)qwn;
        e.QATy[o] = (VT)qtn;
        if constexpr (W2 == 2) {
            e.w_bar[o] = (VT)wbn;
            e.dw[o] = (VT)(wbn - w0);
            e.ATdy[o] = (VT)(g - t0);
            e.ATy_bar[o] = (VT)g;
            e.Qwb_new[o] = (VT)qwbn;
            e.Qg[o] = (VT)qg;
        }
        wi = (double)(VT)wn;
        atyi = (double)(VT)atyn;
        qwi = (double)(VT)qwn;
    } else {
        wi = (double)e.w[o];
        atyi = (double)e.ATy[o];
        qwi = (double)e.Qw[o];
    }
    if constexpr (ZX != 0) {
        const double xi = (double)e.x[o];
        const double z_raw = xi + e.sigma * (-qwi + atyi - (double)e.c[o]);
        const double xb = proj(z_raw, e.l ? (double)e.l[o] : e.lS, e.u ? (double)e.u[o] : e.uS);
        const double xh = 2.0 * xb - xi;
        const double xn = fma(e.h2, xh, e.h1 * e.last_x[o]);
        const double wbz = fma(e.f1s, wi, e.f2s * xh);
        if constexpr (ZX == 2) {
            e.dx[o] = (VT)(xb - xi);
            e.x_bar[o] = (VT)xb;
            e.z_bar[o] = (VT)((xb - z_raw) / e.sigma);
        }
        e.x[o] = (VT)xn;
        e.x_hat[o] = (VT)xh;
        e.w_bar[o] = (VT)wbz;
        e.d[o] = (VT)(wi - (double)(VT)xh);
    }
}

template <class VT, class DT, int W2, int ZX>
static void qap_ds_step(hipStream_t s, const QapDsStep<VT, DT> &e) {
    const int64_t nn = (int64_t)e.n * e.n;
    launch_k(k_qap_ds_step<VT, DT, W2, ZX>, dim3((unsigned)((nn + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, e);
    HPRQP_HIP(hipGetLastError());
}

// ---------------- nq < 256: three launches per step ---------------------------
// With the small-n passes (one 16 x 16 tile per block) the y update joins
// the fused kernel too: pass 2 hands tempv to EpiTempvTiles, which returns it
// for the tile's row and column sums instead of storing it, and the fused
// kernel forms A tempv from those sums (row i: the sums of the tiles along
// row i; row n + j: down column j, in tile order), then y_bar, the next y
// (into the other buffer of a pair, since every block reads the current
// one), A^T y_bar and the products of y_bar for its own tile.  A step is
// then: fused kernel, pass 1, pass 2.

// from the product Q d: Q w_bar = Q w - f2 Q d (stored) and tempv =
// x_hat + sigma f2 Q d, stored (TILES false) or returned for the tile sums
template <class VT, bool TILES>
struct EpiTempvDiff {
    static constexpr bool kTileSums = TILES;
    const VT *x_hat, *Qw;
    VT *Qw_bar, *tempv;
    double *tsum;
    double sigma, f2;
    template <class AT>
    __device__ double operator()(int32_t i, AT qd_) const {
        const double qd = f2 * (double)qd_;
        Qw_bar[i] = (VT)((double)Qw[i] - qd);
        const double t = (double)x_hat[i] + sigma * qd;
        if constexpr (!TILES) tempv[i] = (VT)t;
        return t;
    }
};

// the y update's data (EpiY's formulas, src/kernels.jl:760-862)
template <class VT, class DT>
struct QapDsY {
    const double *tsum;  // 2 nt n tile sums of tempv
    const DT *AL, *AU;   // AU null: every row an equality
    const double *last_y;
    const VT *y;         // current y (read by every block)
    VT *y_next, *y_bar, *s, *dy;  // written by block (0, 0)
    double fact1, fact2, h1, h2;
};

// grid nt x nt of 256-thread blocks; block (bx, by) owns rows 16 bx.. and
// columns 16 by.. of the n x n iterates, thread t entry (16 bx + t % 16,
// 16 by + t / 16).  YW (0 none, 1 partial, 2 full): the y update and the
// second w update of the pending step; ZX as in k_qap_ds_step.
template <class VT, class DT, int YW, int ZX>
__global__ __launch_bounds__(256) void k_qap_ds_tile(QapDsStep<VT, DT> e, QapDsY<VT, DT> q) {
    __shared__ double ys[512];          // y_bar (2n <= 510)
    __shared__ double red[4][16][17];   // products' partials: [product][k segment][entry]
    __shared__ double pv[4][16];        // P, R, Sa, Tb of the tile's rows / columns
    const int n = e.n, t = threadIdx.x, li = t & 15, lj = t >> 4;
    const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16, i = i0 + li, j = j0 + lj;
    const bool in = i < n && j < n;
    const int64_t o = in ? (int64_t)i + (int64_t)j * n : 0;
    // everything that does not depend on y_bar is loaded first, so its
    // latency overlaps the tile sums and the products
    double w0 = (double)e.w[o], t0 = (double)e.ATy[o], qw0 = (double)e.Qw[o];
    double wb = 0.0, qt0 = 0.0, qwb = 0.0, lw = 0.0, lt = 0.0, lqw = 0.0, lqt = 0.0;
    if constexpr (YW != 0) {
        wb = (double)e.w_bar[o];
        qt0 = (double)e.QATy[o];
        qwb = (double)e.Qw_bar[o];
        lw = e.last_w[o];
        lt = e.last_ATy[o];
        lqw = e.last_Qw[o];
        lqt = e.last_QATy[o];
    }
    double xi = 0.0, ci = 0.0, lx = 0.0, lo_b = e.lS, hi_b = e.uS;
    if constexpr (ZX != 0) {
        xi = (double)e.x[o];
        ci = (double)e.c[o];
        lx = e.last_x[o];
        if (e.l) lo_b = (double)e.l[o];
        if (e.u) hi_b = (double)e.u[o];
    }
    double wi = w0, atyi = t0, qwi = qw0;
    if constexpr (YW != 0) {
        // the products' operands: thread (li, lj) takes k in segment lj of
        // row i0 + li of A and S and of column j0 + li of B and T (as rows of
        // their transposes); kseg <= 16 for n < 256
        const int kseg = (n + 15) / 16, ka = lj * kseg;
        const int ri = min(i0 + li, n - 1), rj = min(j0 + li, n - 1);
        const bool fs_ = e.Sa != nullptr, ft_ = e.Tb != nullptr;
        double ma[16], mb[16], ms[16], mt[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            ma[u] = mb[u] = ms[u] = mt[u] = 0.0;
            if (u < kseg) {
                const int k = min(ka + u, n - 1);
                ma[u] = e.gA[ri + (int64_t)k * n];
                mb[u] = e.gBt[rj + (int64_t)k * n];
                if (fs_) ms[u] = e.gS[ri + (int64_t)k * n];
                if (ft_) mt[u] = e.gTt[rj + (int64_t)k * n];
            }
        }
        const int m = 2 * n, nt = (int)gridDim.x;
        const bool lead = blockIdx.x == 0 && blockIdx.y == 0;
        for (int r = t; r < m; r += 256) {
            const double *ps = r < n ? q.tsum + r : q.tsum + (int64_t)nt * n + (r - n);
            double ax = 0.0;
            for (int b = 0; b < nt; ++b) ax += ps[(int64_t)b * n];
            const double yi = (double)q.y[r];
            const double s_raw = ax - q.fact1 * yi;
            const double lo = (double)q.AL[r];
            const double sp = proj(s_raw, lo, q.AU ? (double)q.AU[r] : lo);
            const double yb = q.fact2 * (sp - s_raw);
            ys[r] = (double)(VT)yb;  // y_bar as the iteration stores it
            if (lead) {
                const double yh = 2.0 * yb - yi;
                q.y_next[r] = (VT)fma(q.h2, yh, q.h1 * q.last_y[r]);
                q.y_bar[r] = (VT)yb;
                if constexpr (YW == 2) {
                    q.s[r] = (VT)sp;
                    q.dy[r] = (VT)(yb - yi);
                }
            }
        }
        __syncthreads();
        // P_i = (A a)_i, R_j = (B^T b)_j and, for a full S / T, (S a)_i,
        // (T^T b)_j, with a = y_bar[0:n], b = y_bar[n:2n]
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int k = ka + u;
            const bool kin = u < kseg && k < n;
            const double a = kin ? ys[k] : 0.0, b = kin ? ys[n + k] : 0.0;
            acc[0] = fma(ma[u], a, acc[0]);
            acc[1] = fma(mb[u], b, acc[1]);
            acc[2] = fma(ms[u], a, acc[2]);
            acc[3] = fma(mt[u], b, acc[3]);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) red[c][lj][li] = acc[c];
        __syncthreads();
        if (t < 64) {
            const int c = t >> 4, r = t & 15;
            double v = 0.0;
            for (int sgm = 0; sgm < 16; ++sgm) v += red[c][sgm][r];
            pv[c][r] = v;
        }
        __syncthreads();
        if (in) {
            const double a = ys[i], b = ys[n + j];
            const double g = a + b;
            const double sa = fs_ ? pv[2][li] : e.sd[i] * a, rs = fs_ ? e.rS[i] : e.sd[i];
            const double tb = ft_ ? pv[3][lj] : e.td[j] * b, ct = ft_ ? e.cT[j] : e.td[j];
            const double qg = 2.0 * (pv[0][li] * e.cB[j] + e.rA[i] * pv[1][lj] - sa - rs * b - a * ct - tb);
            const double wbn = wb + e.fs * (g - t0);
            const double wn = e.wh1 * lw + e.wh2 * (2.0 * wbn - w0);
            const double atyn = e.wh1 * lt + e.wh2 * (2.0 * g - t0);
            const double qwbn = qwb + e.fs * (qg - qt0);
            const double qwn = e.wh1 * lqw + e.wh2 * (2.0 * qwbn - qw0);
            const double qtn = e.wh1 * lqt + e.wh2 * (2.0 * qg - qt0);
            e.w[o] = (VT)wn;
            e.ATy[o] = (VT)atyn;
            e.Qw[o] = (VT)qwn;
            e.QATy[o] = (VT)qtn;
            if constexpr (YW == 2) {
                e.w_bar[o] = (VT)wbn;
                e.dw[o] = (VT)(wbn - w0);
                e.ATdy[o] = (VT)(g - t0);
                e.ATy_bar[o] = (VT)g;
                e.Qwb_new[o] = (VT)qwbn;
                e.Qg[o] = (VT)qg;
            }
            wi = (double)(VT)wn;
            atyi = (double)(VT)atyn;
            qwi = (double)(VT)qwn;
        }
    }
    if constexpr (ZX != 0) {
        if (!in) return;
        const double z_raw = xi + e.sigma * (-qwi + atyi - ci);
        const double xb = proj(z_raw, lo_b, hi_b);
        const double xh = 2.0 * xb - xi;
        const double xn = fma(e.h2, xh, e.h1 * lx);
        const double wbz = fma(e.f1s, wi, e.f2s * xh);
        if constexpr (ZX == 2) {
            e.dx[o] = (VT)(xb - xi);
            e.x_bar[o] = (VT)xb;
            e.z_bar[o] = (VT)((xb - z_raw) / e.sigma);
        }
        e.x[o] = (VT)xn;
        e.x_hat[o] = (VT)xh;
        e.w_bar[o] = (VT)wbz;
        e.d[o] = (VT)(wi - (double)(VT)xh);
    }
}

template <class VT, class DT, int YW, int ZX>
static void qap_ds_tile(hipStream_t s, const QapDsStep<VT, DT> &e, const QapDsY<VT, DT> &q) {
    const unsigned nt = (unsigned)((e.n + 15) / 16);
    launch_k(k_qap_ds_tile<VT, DT, YW, ZX>, dim3(nt, nt), dim3(256), 0, s, e, q);
    HPRQP_HIP(hipGetLastError());
}

}  // namespace hprqp
