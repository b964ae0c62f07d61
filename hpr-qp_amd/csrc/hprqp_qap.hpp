// hprqp_qap.hpp -- QAP relaxation operator Q(X) = 2(A X B - S X - X T)
// (src/Q_operators/QAP_operator.jl:52-76) on dense column-major nq x nq
// matrices, X = reshape(x, nq, nq).
//
// The operator is the only dense contraction of the solver, so it runs on
// the matrix cores, in two precisions:
//   * FP64: v_mfma_f64_16x16x4f64, FP64 accumulation (plain FP64 solves and
//     every residual evaluation);
//   * split-BF16 x3: v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  Each
//     product a*b becomes three BF16 products: when the coefficient matrix
//     (A, B, S or T) is exactly representable in BF16 (integer flows and
//     distances are), the data operand is split into three BF16 pieces
//     b0 + b1 + b2 (24 significant bits, FP32 accuracy) and the terms are
//     a*b0 + a*b1 + a*b2; otherwise both are split in two and the terms are
//     a0*b0 + a0*b1 + a1*b0 (16 significant bits).
//
// Both run the operator as two tiled passes, one wavefront per 16 x 16
// result tile, the K dimension in slabs of kKS staged through LDS by
// coalesced loads:
//   pass 1: Y = A X and, for a full S, Z = S X              -> temporaries
//   pass 2: R = Y B - X T (full T);
//           out = 2 (R - Z, or - s_i X_ij for diagonal S, - X_ij t_j for
//           diagonal T)
#pragma once
#include <cmath>
#include <cstring>

#include "hprqp_internal.hpp"

namespace hprqp {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) unsigned short u16x8;
typedef __attribute__((ext_vector_type(4))) float floatx4;
typedef __attribute__((ext_vector_type(4))) double doublex4;

// MFMA operand layouts (16 x 16 tiles, lane l, group g = l / 16, measured
// on gfx950 with lane-index probes):
//   bf16 16x16x32: left operand row l%16, k = 8g..8g+7; right operand column
//   l%16, same k; result element e (0..3) at row 4g + e, column l%16.
//   f64 16x16x4: left row l%16, k = g; right column l%16, k = g; result
//   element e at row g + 4e, column l%16.
constexpr int kKS = 128;  // K slab

// ---------------- BF16 splitting ---------------------------------------
__device__ __forceinline__ unsigned short bf16_rn(float f) {
    const uint32_t u = __float_as_uint(f);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

static inline unsigned short bf16_rn_host(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ void split2(const float v[8], bf16x8 &p0, bf16x8 &p1) {
    u16x8 a, b;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const unsigned short h = bf16_rn(v[e]);
        a[e] = h;
        b[e] = bf16_rn(v[e] - bf16_f(h));
    }
    p0 = __builtin_bit_cast(bf16x8, a);
    p1 = __builtin_bit_cast(bf16x8, b);
}
__device__ __forceinline__ void split3(const float v[8], bf16x8 &p0, bf16x8 &p1, bf16x8 &p2) {
    u16x8 a, b, c;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const unsigned short h = bf16_rn(v[e]);
        const float r = v[e] - bf16_f(h);
        const unsigned short m = bf16_rn(r);
        a[e] = h;
        b[e] = m;
        c[e] = bf16_rn(r - bf16_f(m));
    }
    p0 = __builtin_bit_cast(bf16x8, a);
    p1 = __builtin_bit_cast(bf16x8, b);
    p2 = __builtin_bit_cast(bf16x8, c);
}
__device__ __forceinline__ floatx4 mfma_bf16(bf16x8 a, bf16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// acc += C D with the coefficient C (pieces c0, c1) as the left operand
__device__ __forceinline__ floatx4 mma3_cl(floatx4 acc, bf16x8 c0, bf16x8 c1, bool exact, const float d[8]) {
    if (exact) {
        bf16x8 d0, d1, d2;
        split3(d, d0, d1, d2);
        acc = mfma_bf16(c0, d2, acc);
        acc = mfma_bf16(c0, d1, acc);
        acc = mfma_bf16(c0, d0, acc);
    } else {
        bf16x8 d0, d1;
        split2(d, d0, d1);
        acc = mfma_bf16(c1, d0, acc);
        acc = mfma_bf16(c0, d1, acc);
        acc = mfma_bf16(c0, d0, acc);
    }
    return acc;
}
// acc += D C with the coefficient as the right operand
__device__ __forceinline__ floatx4 mma3_cr(floatx4 acc, bf16x8 c0, bf16x8 c1, bool exact, const float d[8]) {
    if (exact) {
        bf16x8 d0, d1, d2;
        split3(d, d0, d1, d2);
        acc = mfma_bf16(d2, c0, acc);
        acc = mfma_bf16(d1, c0, acc);
        acc = mfma_bf16(d0, c0, acc);
    } else {
        bf16x8 d0, d1;
        split2(d, d0, d1);
        acc = mfma_bf16(d0, c1, acc);
        acc = mfma_bf16(d1, c0, acc);
        acc = mfma_bf16(d0, c0, acc);
    }
    return acc;
}

struct QapArgs {
    int32_t n, np;
    const double *A, *B, *S, *T, *sd, *td;
    const unsigned short *A0, *A1, *B0, *B1, *S0, *S1, *T0, *T1;
    int a_exact, b_exact, s_exact, t_exact, s_full, t_full;
};

static inline QapArgs qap_args(const hprqp_qop *q) {
    QapArgs g{};
    g.n = q->nq;
    g.np = q->qnp;
    g.A = q->qa.p;
    g.B = q->qb.p;
    g.S = q->qs.p;
    g.T = q->qt.p;
    g.sd = q->qsd.p;
    g.td = q->qtd.p;
    g.A0 = q->qa0.p;
    g.A1 = q->qa1.p;
    g.B0 = q->qb0.p;
    g.B1 = q->qb1.p;
    g.S0 = q->qs0.p;
    g.S1 = q->qs1.p;
    g.T0 = q->qt0.p;
    g.T1 = q->qt1.p;
    g.a_exact = q->qa_exact;
    g.b_exact = q->qb_exact;
    g.s_exact = q->qs_exact;
    g.t_exact = q->qt_exact;
    g.s_full = !q->qs_diag;
    g.t_full = !q->qt_diag;
    return g;
}

// ---------------- FP64 passes ----------------------------------------------
constexpr int kLd64 = kKS + 1;  // double stride of staged column blocks

// Staging: every lane issues all of its loads of a slab before the first
// LDS store (fully unrolled, so the loads are in flight together), then
// stores.  A slab is 16 rows or columns x kKS; entries outside the matrix
// are zero.  Out-of-range lanes load element 0 and discard it, so that no
// load sits behind a branch (a branch around each load would serialise
// them).
constexpr int kSlabIt = 16 * kKS / 64;  // elements per lane per slab

// rows r0..r0+15 of column-major M over columns k0..k0+kKS-1 into L[k][i]
// (ld 16), negated when neg
template <class T, class D>
__device__ __forceinline__ void stage_rows(const T *M, int n, int r0, int k0, D *L, bool neg, int ldk, bool kmajor) {
    const int lane = threadIdx.x, i = lane & 15, r = r0 + i;
    D v[kSlabIt];
#pragma unroll
    for (int t = 0; t < kSlabIt; ++t) {
        const int k = k0 + t * 4 + (lane >> 4);
        const bool in = r < n && k < n;
        const D a = (D)M[in ? r + (int64_t)k * n : 0];
        v[t] = in ? a : (D)0;
    }
#pragma unroll
    for (int t = 0; t < kSlabIt; ++t) {
        const int kk = t * 4 + (lane >> 4);
        L[kmajor ? kk * 16 + i : i * ldk + kk] = neg ? -v[t] : v[t];
    }
}
// columns c0..c0+15 of column-major M over rows k0..k0+kKS-1 into L[j][k]
template <class T, class D>
__device__ __forceinline__ void stage_cols(const T *M, int n, int c0, int k0, D *L, int ldk) {
    const int lane = threadIdx.x;
    D v[kSlabIt];
#pragma unroll
    for (int t = 0; t < kSlabIt; ++t) {
        const int kk = (t & 1) * 64 + lane, c = c0 + (t >> 1), k = k0 + kk;
        const bool in = c < n && k < n;
        const D a = (D)M[in ? k + (int64_t)c * n : 0];
        v[t] = in ? a : (D)0;
    }
#pragma unroll
    for (int t = 0; t < kSlabIt; ++t) L[(t >> 1) * ldk + (t & 1) * 64 + lane] = v[t];
}

static __global__ __launch_bounds__(64) void k_qap64_p1(QapArgs g, const double *__restrict__ x, double *__restrict__ Y,
                                                  double *__restrict__ Z) {
    __shared__ double As[kKS * 16], Ss[kKS * 16], Xs[16 * kLd64];
    const int n = g.n, r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
    const int li = threadIdx.x & 15, lg = threadIdx.x >> 4;
    doublex4 ay = {0.0, 0.0, 0.0, 0.0}, az = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < n; k0 += kKS) {
        const int kwp = (min(kKS, n - k0) + 3) & ~3;
        stage_rows(g.A, n, r0, k0, As, false, 0, true);
        if (g.s_full) stage_rows(g.S, n, r0, k0, Ss, false, 0, true);
        stage_cols(x, n, c0, k0, Xs, kLd64);
        __syncthreads();
        for (int kq = 0; kq < kwp; kq += 4) {
            const double b = Xs[li * kLd64 + kq + lg];
            ay = __builtin_amdgcn_mfma_f64_16x16x4f64(As[(kq + lg) * 16 + li], b, ay, 0, 0, 0);
            if (g.s_full) az = __builtin_amdgcn_mfma_f64_16x16x4f64(Ss[(kq + lg) * 16 + li], b, az, 0, 0, 0);
        }
        __syncthreads();
    }
    const int j = c0 + li;
    if (j >= n) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = r0 + lg + 4 * e;
        if (r >= n) continue;
        Y[r + (int64_t)j * n] = ay[e];
        if (g.s_full) Z[r + (int64_t)j * n] = az[e];
    }
}

template <class Epi>
__global__ __launch_bounds__(64) void k_qap64_p2(QapArgs g, const double *__restrict__ x,
                                                  const double *__restrict__ Y, const double *__restrict__ Z, Epi epi) {
    __shared__ double Ls[kKS * 16], Xr[kKS * 16], Bs[16 * kLd64], Ts[16 * kLd64];
    const int n = g.n, r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
    const int li = threadIdx.x & 15, lg = threadIdx.x >> 4;
    doublex4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < n; k0 += kKS) {
        const int kwp = (min(kKS, n - k0) + 3) & ~3;
        stage_rows(Y, n, r0, k0, Ls, false, 0, true);
        stage_cols(g.B, n, c0, k0, Bs, kLd64);
        if (g.t_full) {
            stage_rows(x, n, r0, k0, Xr, true, 0, true);
            stage_cols(g.T, n, c0, k0, Ts, kLd64);
        }
        __syncthreads();
        for (int kq = 0; kq < kwp; kq += 4) {
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Ls[(kq + lg) * 16 + li], Bs[li * kLd64 + kq + lg], acc, 0, 0,
                                                       0);
            if (g.t_full)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Xr[(kq + lg) * 16 + li], Ts[li * kLd64 + kq + lg], acc,
                                                           0, 0, 0);
        }
        __syncthreads();
    }
    const int j = c0 + li;
    if (j >= n) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = r0 + lg + 4 * e;
        if (r >= n) continue;
        const int64_t o = r + (int64_t)j * n;
        double v = acc[e];
        if (g.s_full)
            v -= Z[o];
        else
            v -= g.sd[r] * x[o];
        if (!g.t_full) v -= x[o] * g.td[j];
        epi((int32_t)o, 2.0 * v);
    }
}

// ---------------- split-BF16 x3 passes ---------------------------------------
constexpr int kLdB = kKS + 8;  // u16 row stride of staged coefficient pieces
constexpr int kLdF = kKS + 4;  // float row stride of staged data

// rows b0..b0+15 of a padded np x np piece stored P[b * np + k] (the row-major
// left pieces A, S, and the column-major right pieces B, T, whose column b
// is contiguous) over k0..k0+kKS-1 into L[i][k], 16-byte loads
__device__ __forceinline__ void stage_piece(const unsigned short *P, int np, int b0, int k0, unsigned short *L) {
    constexpr int kChunks = 16 * kKS / 8 / 64;  // 16-byte chunks per lane
    const int lane = threadIdx.x;
    u16x8 v[kChunks];
#pragma unroll
    for (int t = 0; t < kChunks; ++t) {
        const int e = t * 64 + lane, i = e / (kKS / 8), k = k0 + (e % (kKS / 8)) * 8;
        const bool in = k < np;
        const u16x8 a = *(const u16x8 *)(P + (in ? (int64_t)(b0 + i) * np + k : 0));
        const u16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
        v[t] = in ? a : z;
    }
#pragma unroll
    for (int t = 0; t < kChunks; ++t) {
        const int e = t * 64 + lane, i = e / (kKS / 8), kk = (e % (kKS / 8)) * 8;
        *(u16x8 *)(L + i * kLdB + kk) = v[t];
    }
}
__device__ __forceinline__ void frag_f(const float *L, int li, int k, float v[8]) {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = L[li * kLdF + k + e];
}
__device__ __forceinline__ bf16x8 frag_b(const unsigned short *L, int li, int k) {
    return __builtin_bit_cast(bf16x8, *(const u16x8 *)(L + li * kLdB + k));
}

template <class IT>
__global__ __launch_bounds__(64) void k_qapbf_p1(QapArgs g, const IT *__restrict__ x, float *__restrict__ Y,
                                                  float *__restrict__ Z) {
    __shared__ __attribute__((aligned(16))) unsigned short A0s[16 * kLdB], A1s[16 * kLdB], S0s[16 * kLdB],
        S1s[16 * kLdB];
    __shared__ float Xs[16 * kLdF];
    const int n = g.n, np = g.np, r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
    const int li = threadIdx.x & 15, lg = threadIdx.x >> 4;
    floatx4 ay = {0.f, 0.f, 0.f, 0.f}, az = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < np; k0 += kKS) {
        stage_piece(g.A0, np, r0, k0, A0s);
        if (!g.a_exact) stage_piece(g.A1, np, r0, k0, A1s);
        if (g.s_full) {
            stage_piece(g.S0, np, r0, k0, S0s);
            if (!g.s_exact) stage_piece(g.S1, np, r0, k0, S1s);
        }
        stage_cols(x, n, c0, k0, Xs, kLdF);
        __syncthreads();
        const int kend = min(kKS, np - k0);
        for (int kq = 0; kq < kend; kq += 32) {
            const int k = kq + 8 * lg;
            float v[8];
            frag_f(Xs, li, k, v);
            const bf16x8 a0 = frag_b(A0s, li, k);
            ay = mma3_cl(ay, a0, g.a_exact ? a0 : frag_b(A1s, li, k), g.a_exact, v);
            if (g.s_full) {
                const bf16x8 s0 = frag_b(S0s, li, k);
                az = mma3_cl(az, s0, g.s_exact ? s0 : frag_b(S1s, li, k), g.s_exact, v);
            }
        }
        __syncthreads();
    }
    const int j = c0 + li;
    if (j >= n) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = r0 + 4 * lg + e;
        if (r >= n) continue;
        Y[r + (int64_t)j * n] = ay[e];
        if (g.s_full) Z[r + (int64_t)j * n] = az[e];
    }
}

template <class IT, class Epi>
__global__ __launch_bounds__(64) void k_qapbf_p2(QapArgs g, const IT *__restrict__ x, const float *__restrict__ Y,
                                                  const float *__restrict__ Z, Epi epi) {
    __shared__ __attribute__((aligned(16))) unsigned short B0s[16 * kLdB], B1s[16 * kLdB], T0s[16 * kLdB],
        T1s[16 * kLdB];
    __shared__ float Ys[16 * kLdF], Xr[16 * kLdF];
    const int n = g.n, np = g.np, r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
    const int li = threadIdx.x & 15, lg = threadIdx.x >> 4;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < np; k0 += kKS) {
        stage_rows(Y, n, r0, k0, Ys, false, kLdF, false);
        stage_piece(g.B0, np, c0, k0, B0s);
        if (!g.b_exact) stage_piece(g.B1, np, c0, k0, B1s);
        if (g.t_full) {
            stage_rows(x, n, r0, k0, Xr, true, kLdF, false);
            stage_piece(g.T0, np, c0, k0, T0s);
            if (!g.t_exact) stage_piece(g.T1, np, c0, k0, T1s);
        }
        __syncthreads();
        const int kend = min(kKS, np - k0);
        for (int kq = 0; kq < kend; kq += 32) {
            const int k = kq + 8 * lg;
            float v[8];
            frag_f(Ys, li, k, v);
            const bf16x8 b0 = frag_b(B0s, li, k);
            acc = mma3_cr(acc, b0, g.b_exact ? b0 : frag_b(B1s, li, k), g.b_exact, v);
            if (g.t_full) {
                frag_f(Xr, li, k, v);
                const bf16x8 t0 = frag_b(T0s, li, k);
                acc = mma3_cr(acc, t0, g.t_exact ? t0 : frag_b(T1s, li, k), g.t_exact, v);
            }
        }
        __syncthreads();
    }
    const int j = c0 + li;
    if (j >= n) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = r0 + 4 * lg + e;
        if (r >= n) continue;
        const int64_t o = r + (int64_t)j * n;
        double v = (double)acc[e];
        if (g.s_full)
            v -= (double)Z[o];
        else
            v -= g.sd[r] * (double)x[o];
        if (!g.t_full) v -= (double)x[o] * g.td[j];
        epi((int32_t)o, 2.0 * v);
    }
}

// Q x with the result handed to epi(i, (Q x)_i): FP64 / split-BF16 x3
template <class Epi>
void qap_launch_f64(hipStream_t s, const hprqp_qop *q, const double *X, const Epi &epi) {
    const QapArgs g = qap_args(q);
    const int nt = (g.n + 15) / 16;
    const int64_t nn = (int64_t)g.n * g.n;
    double *Y = q->qtemp.p, *Z = q->qtemp.p + nn;
    const dim3 grid(nt, nt);
    launch_k(k_qap64_p1, grid, dim3(64), 0, s, g, X, Y, Z);
    HPRQP_HIP(hipGetLastError());
    launch_k(k_qap64_p2<Epi>, grid, dim3(64), 0, s, g, X, (const double *)Y, (const double *)Z, epi);
    HPRQP_HIP(hipGetLastError());
}

template <class IT, class Epi>
void qap_launch_bf16x3(hipStream_t s, const hprqp_qop *q, const IT *X, const Epi &epi) {
    const QapArgs g = qap_args(q);
    const int nt = (g.n + 15) / 16;
    const int64_t nn = (int64_t)g.n * g.n;
    float *Y = q->qtemp32.p, *Z = q->qtemp32.p + nn;
    const dim3 grid(nt, nt);
    launch_k(k_qapbf_p1<IT>, grid, dim3(64), 0, s, g, X, Y, Z);
    HPRQP_HIP(hipGetLastError());
    launch_k(k_qapbf_p2<IT, Epi>, grid, dim3(64), 0, s, g, X, (const float *)Y, (const float *)Z, epi);
    HPRQP_HIP(hipGetLastError());
}

}  // namespace hprqp
