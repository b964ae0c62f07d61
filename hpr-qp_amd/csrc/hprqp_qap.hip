// hprqp_qap.hip -- host side of the QAP operator: operand preparation
// (BF16 splitting, diagonal detection) and the plain applies.  The kernels
// are in hprqp_qap.hpp, where the solver also instantiates them with its
// fused half-step epilogues.
#include <cstdio>
#include <vector>

#include "hprqp_kernels.hpp"
#include "hprqp_qap.hpp"

namespace hprqp {

void qap_prepare(hipStream_t s, hprqp_qop *q, const double *A, const double *B, const double *S, const double *T) {
    const int n = q->nq;
    // pieces padded to a multiple of 32 (the K step), of 128 (the largest
    // block edge) where the blocked passes run
    const int np = n >= kQapBlockedMin ? (n + 127) / 128 * 128 : (n + 31) / 32 * 32;
    q->qnp = np;
    const int64_t pp = (int64_t)np * np;
    auto is_diag = [&](const double *M) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                if (i != j && M[i + (int64_t)j * n] != 0.0) return false;
        return true;
    };
    q->qs_diag = is_diag(S);
    q->qt_diag = is_diag(T);
    std::vector<double> dg(n);
    if (q->qs_diag) {
        for (int i = 0; i < n; ++i) dg[i] = S[i + (int64_t)i * n];
        q->qsd.alloc(n);
        HPRQP_HIP(hipMemcpy(q->qsd.p, dg.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    if (q->qt_diag) {
        for (int i = 0; i < n; ++i) dg[i] = T[i + (int64_t)i * n];
        q->qtd.alloc(n);
        HPRQP_HIP(hipMemcpy(q->qtd.p, dg.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    // split-BF16 pieces, zero-padded to np x np.  row_major: piece[i*np + k]
    // = M[i, k] (left operands A, S); otherwise piece[k + j*np] = M[k, j]
    // (right operands B, T).  A matrix is exact when piece 0 reproduces it.
    std::vector<unsigned short> p0(pp), p1(pp);
    auto split = [&](const double *M, bool row_major, hprqp::DevVec<uint16_t> &d0, hprqp::DevVec<uint16_t> &d1) {
        std::fill(p0.begin(), p0.end(), 0);
        std::fill(p1.begin(), p1.end(), 0);
        bool exact = true;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double v = M[i + (int64_t)j * n];
                const float f = (float)v;
                const unsigned short h = bf16_rn_host(f);
                const uint32_t hb = (uint32_t)h << 16;
                float hf;
                std::memcpy(&hf, &hb, 4);
                const unsigned short l = bf16_rn_host(f - hf);
                if ((double)hf != v) exact = false;
                const int64_t o = row_major ? (int64_t)i * np + j : i + (int64_t)j * np;
                p0[o] = h;
                p1[o] = l;
            }
        d0.alloc(pp);
        d1.alloc(pp);
        HPRQP_HIP(hipMemcpy(d0.p, p0.data(), sizeof(uint16_t) * pp, hipMemcpyHostToDevice));
        HPRQP_HIP(hipMemcpy(d1.p, p1.data(), sizeof(uint16_t) * pp, hipMemcpyHostToDevice));
        return exact;
    };
    q->qa_exact = split(A, true, q->qa0, q->qa1);
    q->qb_exact = split(B, false, q->qb0, q->qb1);
    if (!q->qs_diag) q->qs_exact = split(S, true, q->qs0, q->qs1);
    if (!q->qt_diag) q->qt_exact = split(T, false, q->qt0, q->qt1);
    // the two n x n temporaries (Y, Z) of each precision
    const int64_t nn = (int64_t)n * n;
    q->qtemp.alloc(2 * nn);
    q->qtemp32.alloc(2 * nn);
    // the large-n split-BF16 passes read their data operands as BF16 pieces;
    // the padding stays zero (the kernels write only the n x n part)
    if (n >= kQapBlockedMin) {
        q->qxp.alloc(3 * pp);
        q->qyp.alloc(3 * pp);
        HPRQP_HIP(hipMemsetAsync(q->qxp.p, 0, sizeof(uint16_t) * 3 * pp, s));
        HPRQP_HIP(hipMemsetAsync(q->qyp.p, 0, sizeof(uint16_t) * 3 * pp, s));
        if (!q->qt_diag) {
            q->qxr.alloc(3 * pp);
            HPRQP_HIP(hipMemsetAsync(q->qxr.p, 0, sizeof(uint16_t) * 3 * pp, s));
        }
        HPRQP_HIP(hipStreamSynchronize(s));
    }
}

void qap_apply_f64(hipStream_t s, const hprqp_qop *q, const double *X, double *QX, double *tmp) {
    qap_launch_f64(s, q, X, EpiStore<double>{QX}, tmp);
}

template <class IT, class OT>
void qap_apply_bf16x3(hipStream_t s, const hprqp_qop *q, const IT *X, OT *QX) {
    qap_launch_bf16x3(s, q, X, EpiStore<OT>{QX});
}

template void qap_apply_bf16x3<double, double>(hipStream_t, const hprqp_qop *, const double *, double *);
template void qap_apply_bf16x3<float, float>(hipStream_t, const hprqp_qop *, const float *, float *);

__global__ void k_qap_tune_fill(int64_t nn, int n, double *x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) x[i] = (1.0 + 0.25 * (double)((i * 7919) % 13)) / n;
}

int qap_family_of(const hprqp_qop *q, bool bf16) {
    return qap_family(qap_args(q), bf16 ? q->qfam16 : q->qfam64, bf16);
}

// Each family that can run on the operator, timed on a plain apply (the
// iteration's epilogues add the same per-element work to every family): 2
// warm-up and 10 timed applies of an n x n input, the fastest kept.  Below
// n = 256, or with a full S or T (blocked only), nothing is timed.
int qap_tune(hipStream_t s, hprqp_qop *q, bool bf16, bool verbose) {
    const QapArgs g = qap_args(q);
    const bool diag = !g.s_full && !g.t_full;
    if (g.n < kQapBlockedMin) return kQapFamSmall;
    if (!diag) return kQapFamBlocked;
    std::vector<int> cand = {kQapFamBlocked, kQapFamLds};
    if (bf16 && (int64_t)(g.np / 64) * (g.np / 128) >= 512) cand.push_back(kQapFamLdsWide);
    const int64_t nn = (int64_t)g.n * g.n;
    DevVec<double> x(nn), y(nn);
    // a doubly stochastic-like iterate: entries near 1/n
    launch_k(k_qap_tune_fill, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, nn, g.n, x.p);
    HPRQP_HIP(hipGetLastError());
    int &fam = bf16 ? q->qfam16 : q->qfam64;
    const int saved = fam;
    hipEvent_t a, b;
    HPRQP_HIP(hipEventCreate(&a));
    HPRQP_HIP(hipEventCreate(&b));
    int best = kQapBySize;
    float best_ms = 0.f;
    for (int c : cand) {
        fam = c;
        auto run = [&] {
            if (bf16)
                qap_apply_bf16x3<double, double>(s, q, (const double *)x.p, y.p);
            else
                qap_apply_f64(s, q, (const double *)x.p, y.p);
        };
        run();
        run();
        HPRQP_HIP(hipEventRecord(a, s));
        for (int r = 0; r < 10; ++r) run();
        HPRQP_HIP(hipEventRecord(b, s));
        HPRQP_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        HPRQP_HIP(hipEventElapsedTime(&ms, a, b));
        ms /= 10.f;
        if (verbose)
            std::printf("QAP kernel autotune    %s family %d: %.4f ms per apply\n", bf16 ? "split-BF16" : "FP64", c,
                        (double)ms);
        if (best == kQapBySize || ms < best_ms) {
            best = c;
            best_ms = ms;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    fam = saved;
    return best;
}

}  // namespace hprqp
