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
// Both run the operator as two passes:
//   pass 1: Y = A X and, for a full S, Z = S X              -> temporaries
//   pass 2: R = Y B - X T (full T);
//           out = 2 (R - Z, or - s_i X_ij for diagonal S, - X_ij t_j for
//           diagonal T), handed to the caller's epilogue
// in one of three kernel families by size: n < 256, one 16 x 16 tile per
// block with its K range split over the block's four waves; 256 <= n < 1024
// (and any full S / T), register-blocked 32 x 32 blocks; n >= 1024 with
// diagonal S and T, LDS-tiled 64 x 64 blocks (split-BF16 from n = 2048:
// 64 x 128, swizzled LDS rows).  Staging loads are issued all
// together before their first use, with out-of-range lanes reading element 0
// and discarding it, so that no load waits behind a branch.
#pragma once
#include <cmath>
#include <cstring>
#include <type_traits>

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

// ---------------- n < 256: split-K one-tile passes ---------------------------
// The one-tile passes with the tile's K range split over four waves of a
// 256-thread block: wave w stages and multiplies the K slabs w, w+4, ... of
// width KS (16 FP64, 32 split-BF16) in its own LDS region, so a tile's loads
// are spread over four waves and each MFMA chain is a quarter as long; the
// four partial tiles are summed through LDS in wave order and wave 0 stores
// (pass 1) or runs the epilogue (pass 2).  Every wave makes the same number
// of slab rounds (a slab past n stages zeros and skips its MFMAs), so the
// block-wide barriers are uniform.
template <int KS>
__device__ __forceinline__ void stage_rows_k(const double *__restrict__ M, int n, int r0, int k0, double *L, bool neg) {
    constexpr int IT = 16 * KS / 64;
    const int lane = threadIdx.x & 63, i = lane & 15, r = r0 + i;
    double v[IT];
#pragma unroll
    for (int t = 0; t < IT; ++t) {
        const int k = k0 + t * 4 + (lane >> 4);
        const bool in = r < n && k < n;
        const double a = M[in ? r + (int64_t)k * n : 0];
        v[t] = in ? a : 0.0;
    }
#pragma unroll
    for (int t = 0; t < IT; ++t) L[(t * 4 + (lane >> 4)) * 16 + i] = neg ? -v[t] : v[t];
}
// columns c0..c0+15 of column-major M over rows k0..k0+KS-1 into L[j * ld + k]
template <int KS, class T, class D>
__device__ __forceinline__ void stage_cols_k(const T *__restrict__ M, int n, int c0, int k0, D *L, int ld) {
    constexpr int IT = 16 * KS / 64, CW = 64 / KS;  // columns per pass of the wave
    const int lane = threadIdx.x & 63, kk = lane % KS;
    D v[IT];
#pragma unroll
    for (int t = 0; t < IT; ++t) {
        const int c = c0 + t * CW + lane / KS, k = k0 + kk;
        const bool in = c < n && k < n;
        const D a = (D)M[in ? k + (int64_t)c * n : 0];
        v[t] = in ? a : (D)0;
    }
#pragma unroll
    for (int t = 0; t < IT; ++t) L[(t * CW + lane / KS) * ld + kk] = v[t];
}

constexpr int kSK64 = 16;          // FP64 K slab per wave
constexpr int kSLd64 = kSK64 + 1;  // double stride of staged columns

__device__ __forceinline__ doublex4 wave_sum1(doublex4 a, double *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int e = 0; e < 4; ++e) red[(w * 4 + e) * 64 + lane] = a[e];
    __syncthreads();
    doublex4 r = {0.0, 0.0, 0.0, 0.0};
    if (w == 0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            double v = red[e * 64 + lane];
#pragma unroll
            for (int u = 1; u < 4; ++u) v += red[(u * 4 + e) * 64 + lane];
            r[e] = v;
        }
    }
    __syncthreads();
    return r;
}

// Epilogues that set kTileSums return the value they produce (tempv) and
// carry tsum: the small-n pass 2 then writes each 16 x 16 tile's row sums
// (tsum[by * n + i]) and column sums (tsum[(nt + bx) * n + j]) of it, from
// which the next kernel forms A tempv of the doubly stochastic constraints.
template <class E, class = void>
struct TileSums : std::false_type {};
template <class E>
struct TileSums<E, std::void_t<decltype(E::kTileSums)>> : std::bool_constant<E::kTileSums> {};

// row and column sums of the tile held by wave 0: lane (g = lane / 16,
// c = lane % 16) has t[e] at column c0 + c and row r0 + g + 4e (ROWS4 false,
// the FP64 MFMA layout) or r0 + 4g + e (true, the BF16 layout); entries
// outside the matrix are 0.  Butterfly sums: every lane ends with the same
// bits.
template <bool ROWS4>
__device__ __forceinline__ void qap_tile_sums(const double (&t)[4], int n, int r0, int c0, double *__restrict__ tsum) {
    const int lane = threadIdx.x & 63, c = lane & 15, gq = lane >> 4;
    double cs = (t[0] + t[1]) + (t[2] + t[3]);
    cs += __shfl_xor(cs, 16);
    cs += __shfl_xor(cs, 32);
    double rs[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        double v = t[e];
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 8);
        rs[e] = v;
    }
    const int64_t nt = gridDim.x;
    if (gq == 0 && c0 + c < n) tsum[(nt + blockIdx.x) * n + c0 + c] = cs;
    if (c == 0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = r0 + (ROWS4 ? 4 * gq + e : gq + 4 * e);
            if (r < n) tsum[(int64_t)blockIdx.y * n + r] = rs[e];
        }
    }
}

static __global__ __launch_bounds__(256) void k_qap64s_p1(QapArgs g, const double *__restrict__ x,
                                                          double *__restrict__ Y, double *__restrict__ Z) {
    __shared__ double As[4][kSK64 * 16], Ss[4][kSK64 * 16], Xs[4][16 * kSLd64];
    __shared__ double red[4 * 4 * 64];
    const int n = g.n, r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    doublex4 ay = {0.0, 0.0, 0.0, 0.0}, az = {0.0, 0.0, 0.0, 0.0};
    for (int kb = 0; kb < n; kb += 4 * kSK64) {
        const int k0 = kb + w * kSK64;
        const int kwp = k0 < n ? (min(kSK64, n - k0) + 3) & ~3 : 0;
        stage_rows_k<kSK64>(g.A, n, r0, k0, As[w], false);
        if (g.s_full) stage_rows_k<kSK64>(g.S, n, r0, k0, Ss[w], false);
        stage_cols_k<kSK64>(x, n, c0, k0, Xs[w], kSLd64);
        __syncthreads();
        for (int kq = 0; kq < kwp; kq += 4) {
            const double b = Xs[w][li * kSLd64 + kq + lg];
            ay = __builtin_amdgcn_mfma_f64_16x16x4f64(As[w][(kq + lg) * 16 + li], b, ay, 0, 0, 0);
            if (g.s_full) az = __builtin_amdgcn_mfma_f64_16x16x4f64(Ss[w][(kq + lg) * 16 + li], b, az, 0, 0, 0);
        }
        __syncthreads();
    }
    ay = wave_sum1(ay, red);
    if (g.s_full) az = wave_sum1(az, red);
    const int j = c0 + li;
    if (w != 0 || j >= n) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = r0 + lg + 4 * e;
        if (r >= n) continue;
        Y[r + (int64_t)j * n] = ay[e];
        if (g.s_full) Z[r + (int64_t)j * n] = az[e];
    }
}

template <class Epi>
__global__ __launch_bounds__(256) void k_qap64s_p2(QapArgs g, const double *__restrict__ x,
                                                   const double *__restrict__ Y, const double *__restrict__ Z,
                                                   Epi epi) {
    __shared__ double Ls[4][kSK64 * 16], Xr[4][kSK64 * 16], Bs[4][16 * kSLd64], Ts[4][16 * kSLd64];
    __shared__ double red[4 * 4 * 64];
    const int n = g.n, r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    doublex4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int kb = 0; kb < n; kb += 4 * kSK64) {
        const int k0 = kb + w * kSK64;
        const int kwp = k0 < n ? (min(kSK64, n - k0) + 3) & ~3 : 0;
        stage_rows_k<kSK64>(Y, n, r0, k0, Ls[w], false);
        stage_cols_k<kSK64>(g.B, n, c0, k0, Bs[w], kSLd64);
        if (g.t_full) {
            stage_rows_k<kSK64>(x, n, r0, k0, Xr[w], true);
            stage_cols_k<kSK64>(g.T, n, c0, k0, Ts[w], kSLd64);
        }
        __syncthreads();
        for (int kq = 0; kq < kwp; kq += 4) {
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Ls[w][(kq + lg) * 16 + li], Bs[w][li * kSLd64 + kq + lg], acc,
                                                       0, 0, 0);
            if (g.t_full)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Xr[w][(kq + lg) * 16 + li], Ts[w][li * kSLd64 + kq + lg],
                                                           acc, 0, 0, 0);
        }
        __syncthreads();
    }
    acc = wave_sum1(acc, red);
    const int j = c0 + li;
    if (w != 0) return;
    double tv[4] = {0.0, 0.0, 0.0, 0.0};
    if (j < n) {
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
            if constexpr (TileSums<Epi>::value)
                tv[e] = epi((int32_t)o, 2.0 * v);
            else
                epi((int32_t)o, 2.0 * v);
        }
    }
    if constexpr (TileSums<Epi>::value) qap_tile_sums<false>(tv, n, r0, c0, epi.tsum);
}

// ---------------- FP64 passes, large n ----------------------------------------
// 256-thread blocks; a block owns a (16 kBT) x (16 kBT) output block held as
// kBT x kBT MFMA tiles with independent accumulators (kBT^2 MFMA chains in
// flight per wave instead of one), and its four waves split the K range
// (wave w takes the 4-wide K steps w, w+4, w+8, ...), so a block is four
// times the parallelism of one wave per block and no wave repeats another's
// loads.  The partial accumulators are summed through LDS in wave order
// (deterministic).  Operands are read straight from global memory: the
// n x n operands stay in L2 at these sizes.  Rows and columns past n are
// clamped to n-1 (their results are not stored); the last, partial K step
// zeroes the entries past n.
constexpr int kBT = 2;                 // MFMA tiles per block in each direction
constexpr int kBE = 16 * kBT;          // block tile edge
constexpr int kQapBlockedMin = 256;    // n from which the blocked passes run (measured faster at 256)

// sum of the four waves' accumulators; wave w then owns components
// q = 4w .. 4w+3 of the kBT x kBT x 4 per-lane results: the four elements
// of tile w (kBT = 2: one tile per wave)
__device__ __forceinline__ void wave_sum(doublex4 (&acc)[kBT][kBT], double *red, double (&out)[kBT * kBT]) {
    constexpr int NC = kBT * kBT * 4;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < kBT; ++a)
#pragma unroll
        for (int b = 0; b < kBT; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) red[((w * NC) + (a * kBT + b) * 4 + e) * 64 + lane] = acc[a][b][e];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NC / 4; ++t) {
        const int q = 4 * w + t;
        double v = red[q * 64 + lane];
#pragma unroll
        for (int u = 1; u < 4; ++u) v += red[((u * NC) + q) * 64 + lane];
        out[t] = v;
    }
    __syncthreads();
}

static __global__ __launch_bounds__(256) void k_qap64b_p1(QapArgs g, const double *__restrict__ x,
                                                          double *__restrict__ Y, double *__restrict__ Z) {
    __shared__ double red[4 * kBT * kBT * 4 * 64];
    const int n = g.n, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int r0 = blockIdx.x * kBE, c0 = blockIdx.y * kBE;
    const double *pa[kBT], *ps[kBT], *px[kBT];
#pragma unroll
    for (int t = 0; t < kBT; ++t) {
        const int r = min(r0 + 16 * t + li, n - 1), c = min(c0 + 16 * t + li, n - 1);
        pa[t] = g.A + r;  // A[r + k n]
        ps[t] = g.S + r;
        px[t] = x + (int64_t)c * n;  // X[k + c n]
    }
    doublex4 ay[kBT][kBT], az[kBT][kBT];
#pragma unroll
    for (int a = 0; a < kBT; ++a)
#pragma unroll
        for (int b = 0; b < kBT; ++b) {
            ay[a][b] = doublex4{0.0, 0.0, 0.0, 0.0};
            az[a][b] = doublex4{0.0, 0.0, 0.0, 0.0};
        }
    const int nk = n & ~3;
    auto step = [&](int k, bool tail) {
        const bool in = !tail || k < n;
        const int kk = in ? k : 0;
        const int64_t kn = (int64_t)kk * n;
        double fa[kBT], fs[kBT], fx[kBT];
#pragma unroll
        for (int t = 0; t < kBT; ++t) {
            fa[t] = in ? pa[t][kn] : 0.0;
            fs[t] = in && g.s_full ? ps[t][kn] : 0.0;
            fx[t] = in ? px[t][kk] : 0.0;
        }
#pragma unroll
        for (int a = 0; a < kBT; ++a)
#pragma unroll
            for (int b = 0; b < kBT; ++b) {
                ay[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fx[b], ay[a][b], 0, 0, 0);
                if (g.s_full) az[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fs[a], fx[b], az[a][b], 0, 0, 0);
            }
    };
#pragma unroll 4
    for (int k0 = 4 * w; k0 < nk; k0 += 16) step(k0 + lg, false);
    if (nk < n && w == ((nk >> 2) & 3)) step(nk + lg, true);
    double vy[kBT * kBT], vz[kBT * kBT];
    wave_sum(ay, red, vy);
    if (g.s_full) wave_sum(az, red, vz);
#pragma unroll
    for (int t = 0; t < kBT * kBT; ++t) {
        const int q = 4 * w + t, e = q & 3, ab = q >> 2, a = ab / kBT, b = ab % kBT;
        const int r = r0 + 16 * a + lg + 4 * e, j = c0 + 16 * b + li;
        if (r >= n || j >= n) continue;
        Y[r + (int64_t)j * n] = vy[t];
        if (g.s_full) Z[r + (int64_t)j * n] = vz[t];
    }
}

template <class Epi>
__global__ __launch_bounds__(256) void k_qap64b_p2(QapArgs g, const double *__restrict__ x,
                                                   const double *__restrict__ Y, const double *__restrict__ Z,
                                                   Epi epi) {
    __shared__ double red[4 * kBT * kBT * 4 * 64];
    const int n = g.n, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int r0 = blockIdx.x * kBE, c0 = blockIdx.y * kBE;
    const double *py[kBT], *pxr[kBT], *pb[kBT], *pt[kBT];
#pragma unroll
    for (int t = 0; t < kBT; ++t) {
        const int r = min(r0 + 16 * t + li, n - 1), c = min(c0 + 16 * t + li, n - 1);
        py[t] = Y + r;  // Y[r + k n]
        pxr[t] = x + r;
        pb[t] = g.B + (int64_t)c * n;  // B[k + c n]
        pt[t] = g.T + (int64_t)c * n;
    }
    doublex4 acc[kBT][kBT];
#pragma unroll
    for (int a = 0; a < kBT; ++a)
#pragma unroll
        for (int b = 0; b < kBT; ++b) acc[a][b] = doublex4{0.0, 0.0, 0.0, 0.0};
    const int nk = n & ~3;
    auto step = [&](int k, bool tail) {
        const bool in = !tail || k < n;
        const int kk = in ? k : 0;
        const int64_t kn = (int64_t)kk * n;
        double fy[kBT], fx[kBT], fb[kBT], ft[kBT];
#pragma unroll
        for (int t = 0; t < kBT; ++t) {
            fy[t] = in ? py[t][kn] : 0.0;
            fx[t] = in && g.t_full ? -pxr[t][kn] : 0.0;
            fb[t] = in ? pb[t][kk] : 0.0;
            ft[t] = in && g.t_full ? pt[t][kk] : 0.0;
        }
#pragma unroll
        for (int a = 0; a < kBT; ++a)
#pragma unroll
            for (int b = 0; b < kBT; ++b) {
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fy[a], fb[b], acc[a][b], 0, 0, 0);
                if (g.t_full) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fx[a], ft[b], acc[a][b], 0, 0, 0);
            }
    };
#pragma unroll 4
    for (int k0 = 4 * w; k0 < nk; k0 += 16) step(k0 + lg, false);
    if (nk < n && w == ((nk >> 2) & 3)) step(nk + lg, true);
    double v[kBT * kBT];
    wave_sum(acc, red, v);
#pragma unroll
    for (int t = 0; t < kBT * kBT; ++t) {
        const int q = 4 * w + t, e = q & 3, ab = q >> 2, a = ab / kBT, b = ab % kBT;
        const int r = r0 + 16 * a + lg + 4 * e, j = c0 + 16 * b + li;
        if (r >= n || j >= n) continue;
        const int64_t o = r + (int64_t)j * n;
        double u = v[t];
        if (g.s_full)
            u -= Z[o];
        else
            u -= g.sd[r] * x[o];
        if (!g.t_full) u -= x[o] * g.td[j];
        epi((int32_t)o, 2.0 * u);
    }
}

// ---------------- FP64 passes, n >= 1024, diagonal S and T -------------------
// LDS-tiled GEMM form of the FP64 passes: a block of 2 x 2 waves owns a
// (32 TW) x (32 TW) output, each wave TW x TW MFMA tiles.  Per DK-deep K step
// (32 for 64-wide blocks, TW = 2, the form the launcher uses; 16 for
// 128-wide, within 64 KB of LDS) the block
// stages R = 32 TW rows of the left operand (column-major, so each k is R
// contiguous doubles) and R columns of the right operand (16 contiguous k
// per load group) k-major in LDS, rows padded by one double so both the
// transposing stores and the fragment reads are conflict-free; the next K
// step's loads are issued before this step's MFMAs.  Rows and columns past
// n are clamped to n-1 (results not stored); K past n reads zero.
template <int TW, int DK>
struct LdsStage64 {
    static constexpr int R = 32 * TW;
    static constexpr int KQ = 256 / R;           // threads per left-operand row across k
    static constexpr int NL = DK / KQ;           // left-operand loads per thread
    static constexpr int NR = (R / 16) * (DK / 16);  // right-operand loads per thread
    double l[NL], r[NR];
    __device__ __forceinline__ void load(const double *__restrict__ L, const double *__restrict__ Rm, int n,
                                         int lrow0, int rcol0, int k0) {
        const int t = threadIdx.x;
        const int lr = min(lrow0 + t % R, n - 1), kq = t / R;
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const int k = k0 + kq + KQ * i;
            const double v = L[lr + (int64_t)(k < n ? k : 0) * n];
            l[i] = k < n ? v : 0.0;
        }
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int k = k0 + (i / (R / 16)) * 16 + t % 16;
            const int c = min(rcol0 + t / 16 + 16 * (i % (R / 16)), n - 1);
            const double v = Rm[(k < n ? k : 0) + (int64_t)c * n];
            r[i] = k < n ? v : 0.0;
        }
    }
    __device__ __forceinline__ void store(double *Ls, double *Rs) const {
        const int t = threadIdx.x;
        const int lr = t % R, kq = t / R;
#pragma unroll
        for (int i = 0; i < NL; ++i) Ls[(kq + KQ * i) * (R + 1) + lr] = l[i];
#pragma unroll
        for (int i = 0; i < NR; ++i)
            Rs[((i / (R / 16)) * 16 + t % 16) * (R + 1) + t / 16 + 16 * (i % (R / 16))] = r[i];
    }
};

template <int TW, int DK>
__device__ __forceinline__ void lds_step64(doublex4 (&acc)[TW][TW], const double *Ls, const double *Rs) {
    constexpr int R = 32 * TW;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lg = lane >> 4;
    const int lr = (w & 1) * 16 * TW + li, rc = (w >> 1) * 16 * TW + li;
#pragma unroll
    for (int kq = 0; kq < DK; kq += 4) {
        double fl[TW], fr[TW];
#pragma unroll
        for (int a = 0; a < TW; ++a) fl[a] = Ls[(kq + lg) * (R + 1) + lr + 16 * a];
#pragma unroll
        for (int b = 0; b < TW; ++b) fr[b] = Rs[(kq + lg) * (R + 1) + rc + 16 * b];
#pragma unroll
        for (int a = 0; a < TW; ++a)
#pragma unroll
            for (int b = 0; b < TW; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fl[a], fr[b], acc[a][b], 0, 0, 0);
    }
}

template <int TW, int DK>
__global__ __launch_bounds__(256) void k_qap64l_p1(QapArgs g, const double *__restrict__ x, double *__restrict__ Y) {
    constexpr int R = 32 * TW;
    __shared__ double Ls[DK * (R + 1)], Rs[DK * (R + 1)];
    const int n = g.n, lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lg = lane >> 4;
    const int r0 = blockIdx.x * R, c0 = blockIdx.y * R;
    doublex4 acc[TW][TW];
#pragma unroll
    for (int a = 0; a < TW; ++a)
#pragma unroll
        for (int b = 0; b < TW; ++b) acc[a][b] = doublex4{0.0, 0.0, 0.0, 0.0};
    LdsStage64<TW, DK> st;
    st.load(g.A, x, n, r0, c0, 0);
    for (int k0 = 0; k0 < n; k0 += DK) {
        st.store(Ls, Rs);
        __syncthreads();
        if (k0 + DK < n) st.load(g.A, x, n, r0, c0, k0 + DK);
        lds_step64<TW, DK>(acc, Ls, Rs);
        __syncthreads();
    }
    const int rr = r0 + (w & 1) * 16 * TW, cc = c0 + (w >> 1) * 16 * TW;
#pragma unroll
    for (int a = 0; a < TW; ++a)
#pragma unroll
        for (int b = 0; b < TW; ++b) {
            const int j = cc + 16 * b + li;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = rr + 16 * a + lg + 4 * e;
                if (r < n && j < n) Y[r + (int64_t)j * n] = acc[a][b][e];
            }
        }
}

template <int TW, int DK, class Epi>
__global__ __launch_bounds__(256) void k_qap64l_p2(QapArgs g, const double *__restrict__ x,
                                                   const double *__restrict__ Y, Epi epi) {
    constexpr int R = 32 * TW;
    __shared__ double Ls[DK * (R + 1)], Rs[DK * (R + 1)];
    const int n = g.n, lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lg = lane >> 4;
    const int r0 = blockIdx.x * R, c0 = blockIdx.y * R;
    doublex4 acc[TW][TW];
#pragma unroll
    for (int a = 0; a < TW; ++a)
#pragma unroll
        for (int b = 0; b < TW; ++b) acc[a][b] = doublex4{0.0, 0.0, 0.0, 0.0};
    LdsStage64<TW, DK> st;
    st.load(Y, g.B, n, r0, c0, 0);
    for (int k0 = 0; k0 < n; k0 += DK) {
        st.store(Ls, Rs);
        __syncthreads();
        if (k0 + DK < n) st.load(Y, g.B, n, r0, c0, k0 + DK);
        lds_step64<TW, DK>(acc, Ls, Rs);
        __syncthreads();
    }
    const int rr = r0 + (w & 1) * 16 * TW, cc = c0 + (w >> 1) * 16 * TW;
#pragma unroll
    for (int a = 0; a < TW; ++a)
#pragma unroll
        for (int b = 0; b < TW; ++b) {
            const int j = cc + 16 * b + li;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = rr + 16 * a + lg + 4 * e;
                if (r >= n || j >= n) continue;
                const int64_t o = r + (int64_t)j * n;
                const double xo = x[o];
                epi((int32_t)o, 2.0 * (acc[a][b][e] - g.sd[r] * xo - xo * g.td[j]));
            }
        }
}

// ---------------- split-BF16 x3, n < 256: split-K one-tile passes ------------
constexpr int kSKB = 32;           // K slab per wave: one MFMA K step
constexpr int kSLdB = kSKB + 8;    // u16 row stride of staged pieces
constexpr int kSLdF = kSKB + 4;    // float row stride of staged data

// rows b0..b0+15 of a padded piece over k0..k0+31 into L[i * kSLdB + k]
// (one 16-byte chunk per lane; zero past np)
__device__ __forceinline__ void stage_piece_s(const unsigned short *__restrict__ P, int np, int b0, int k0,
                                              unsigned short *L) {
    const int lane = threadIdx.x & 63, i = lane >> 2, kc = (lane & 3) * 8, k = k0 + kc;
    const bool in = k < np;
    const u16x8 a = *(const u16x8 *)(P + (in ? (int64_t)(b0 + i) * np + k : 0));
    const u16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
    *(u16x8 *)(L + i * kSLdB + kc) = in ? a : z;
}
// rows r0..r0+15 of column-major M over columns k0..k0+31 into L[i * ld + k]
template <class T>
__device__ __forceinline__ void stage_rows_f(const T *__restrict__ M, int n, int r0, int k0, float *L, bool neg) {
    const int lane = threadIdx.x & 63, i = lane & 15, r = r0 + i;
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int k = k0 + t * 4 + (lane >> 4);
        const bool in = r < n && k < n;
        const float a = (float)M[in ? r + (int64_t)k * n : 0];
        v[t] = in ? a : 0.f;
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) L[i * kSLdF + t * 4 + (lane >> 4)] = neg ? -v[t] : v[t];
}
__device__ __forceinline__ floatx4 wave_sum1f(floatx4 a, float *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int e = 0; e < 4; ++e) red[(w * 4 + e) * 64 + lane] = a[e];
    __syncthreads();
    floatx4 r = {0.f, 0.f, 0.f, 0.f};
    if (w == 0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float v = red[e * 64 + lane];
#pragma unroll
            for (int u = 1; u < 4; ++u) v += red[(u * 4 + e) * 64 + lane];
            r[e] = v;
        }
    }
    __syncthreads();
    return r;
}

template <class IT>
__global__ __launch_bounds__(256) void k_qapbfs_p1(QapArgs g, const IT *__restrict__ x, float *__restrict__ Y,
                                                   float *__restrict__ Z) {
    __shared__ __attribute__((aligned(16))) unsigned short A0s[4][16 * kSLdB], A1s[4][16 * kSLdB],
        S0s[4][16 * kSLdB], S1s[4][16 * kSLdB];
    __shared__ float Xs[4][16 * kSLdF];
    __shared__ float red[4 * 4 * 64];
    const int n = g.n, np = g.np, r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    floatx4 ay = {0.f, 0.f, 0.f, 0.f}, az = {0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < np; kb += 4 * kSKB) {
        const int k0 = kb + w * kSKB;
        stage_piece_s(g.A0, np, r0, k0, A0s[w]);
        if (!g.a_exact) stage_piece_s(g.A1, np, r0, k0, A1s[w]);
        if (g.s_full) {
            stage_piece_s(g.S0, np, r0, k0, S0s[w]);
            if (!g.s_exact) stage_piece_s(g.S1, np, r0, k0, S1s[w]);
        }
        stage_cols_k<kSKB>(x, n, c0, k0, Xs[w], kSLdF);
        __syncthreads();
        if (k0 < np) {
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = Xs[w][li * kSLdF + 8 * lg + e];
            const bf16x8 a0 = __builtin_bit_cast(bf16x8, *(const u16x8 *)(A0s[w] + li * kSLdB + 8 * lg));
            const bf16x8 a1 =
                g.a_exact ? a0 : __builtin_bit_cast(bf16x8, *(const u16x8 *)(A1s[w] + li * kSLdB + 8 * lg));
            ay = mma3_cl(ay, a0, a1, g.a_exact, v);
            if (g.s_full) {
                const bf16x8 s0 = __builtin_bit_cast(bf16x8, *(const u16x8 *)(S0s[w] + li * kSLdB + 8 * lg));
                const bf16x8 s1 =
                    g.s_exact ? s0 : __builtin_bit_cast(bf16x8, *(const u16x8 *)(S1s[w] + li * kSLdB + 8 * lg));
                az = mma3_cl(az, s0, s1, g.s_exact, v);
            }
        }
        __syncthreads();
    }
    ay = wave_sum1f(ay, red);
    if (g.s_full) az = wave_sum1f(az, red);
    const int j = c0 + li;
    if (w != 0 || j >= n) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = r0 + 4 * lg + e;
        if (r >= n) continue;
        Y[r + (int64_t)j * n] = ay[e];
        if (g.s_full) Z[r + (int64_t)j * n] = az[e];
    }
}

template <class IT, class Epi>
__global__ __launch_bounds__(256) void k_qapbfs_p2(QapArgs g, const IT *__restrict__ x, const float *__restrict__ Y,
                                                   const float *__restrict__ Z, Epi epi) {
    __shared__ __attribute__((aligned(16))) unsigned short B0s[4][16 * kSLdB], B1s[4][16 * kSLdB],
        T0s[4][16 * kSLdB], T1s[4][16 * kSLdB];
    __shared__ float Ys[4][16 * kSLdF], Xr[4][16 * kSLdF];
    __shared__ float red[4 * 4 * 64];
    const int n = g.n, np = g.np, r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < np; kb += 4 * kSKB) {
        const int k0 = kb + w * kSKB;
        stage_rows_f(Y, n, r0, k0, Ys[w], false);
        stage_piece_s(g.B0, np, c0, k0, B0s[w]);
        if (!g.b_exact) stage_piece_s(g.B1, np, c0, k0, B1s[w]);
        if (g.t_full) {
            stage_rows_f(x, n, r0, k0, Xr[w], true);
            stage_piece_s(g.T0, np, c0, k0, T0s[w]);
            if (!g.t_exact) stage_piece_s(g.T1, np, c0, k0, T1s[w]);
        }
        __syncthreads();
        if (k0 < np) {
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = Ys[w][li * kSLdF + 8 * lg + e];
            const bf16x8 b0 = __builtin_bit_cast(bf16x8, *(const u16x8 *)(B0s[w] + li * kSLdB + 8 * lg));
            const bf16x8 b1 =
                g.b_exact ? b0 : __builtin_bit_cast(bf16x8, *(const u16x8 *)(B1s[w] + li * kSLdB + 8 * lg));
            acc = mma3_cr(acc, b0, b1, g.b_exact, v);
            if (g.t_full) {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = Xr[w][li * kSLdF + 8 * lg + e];
                const bf16x8 t0 = __builtin_bit_cast(bf16x8, *(const u16x8 *)(T0s[w] + li * kSLdB + 8 * lg));
                const bf16x8 t1 =
                    g.t_exact ? t0 : __builtin_bit_cast(bf16x8, *(const u16x8 *)(T1s[w] + li * kSLdB + 8 * lg));
                acc = mma3_cr(acc, t0, t1, g.t_exact, v);
            }
        }
        __syncthreads();
    }
    acc = wave_sum1f(acc, red);
    const int j = c0 + li;
    if (w != 0) return;
    double tv[4] = {0.0, 0.0, 0.0, 0.0};
    if (j < n) {
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
            if constexpr (TileSums<Epi>::value)
                tv[e] = epi((int32_t)o, 2.0 * v);
            else
                epi((int32_t)o, 2.0 * v);
        }
    }
    if constexpr (TileSums<Epi>::value) qap_tile_sums<true>(tv, n, r0, c0, epi.tsum);
}

// ---------------- split-BF16 x3 passes, large n ------------------------------
// The blocking of the large-n FP64 passes (a block owns a 32 x 32 output as
// 2 x 2 MFMA tiles; its four waves split K and are summed through LDS), with
// both operands read as BF16 pieces: the data operand is split once per
// apply (X by k_qap_split, Y written as pieces by pass 1), so the MFMA loop
// is 16-byte fragment loads and MFMAs only.  Pass 1 forms the transpose
// Y^T = X^T A^T, whose result tiles store as row-major Y -- the layout pass 2
// reads as its left operand.  Every piece array is qnp x qnp with zero
// padding, so no load needs a bound check (qnp is a multiple of the 32-wide
// block edge and of the 32-deep K step).
__device__ __forceinline__ bf16x8 ld_frag(const unsigned short *__restrict__ P, int64_t o) {
    return __builtin_bit_cast(bf16x8, *(const u16x8 *)(P + o));
}
// acc += D C for data pieces d (d[2] used when c is exact) and coefficient c
__device__ __forceinline__ floatx4 mma_dc(floatx4 acc, const bf16x8 (&d)[3], bf16x8 c0, bf16x8 c1, bool exact) {
    if (exact) {
        acc = mfma_bf16(d[2], c0, acc);
        acc = mfma_bf16(d[1], c0, acc);
        acc = mfma_bf16(d[0], c0, acc);
    } else {
        acc = mfma_bf16(d[0], c1, acc);
        acc = mfma_bf16(d[1], c0, acc);
        acc = mfma_bf16(d[0], c0, acc);
    }
    return acc;
}
// acc[a][b] += D_a C_b for all a, b as three rounds, each issuing the
// TA x TB independent MFMAs back to back (rather than each accumulator's
// three dependent MFMAs in a row); per accumulator the terms are added in
// the order of mma_dc
template <int TA, int TB, bool EXACT>
__device__ __forceinline__ void mma_rounds_(floatx4 (&acc)[TA][TB], const bf16x8 (&d)[TA][3],
                                            const bf16x8 (&c0)[TB], const bf16x8 (&c1)[TB]) {
#pragma unroll
    for (int rd = 0; rd < 3; ++rd)
#pragma unroll
        for (int a = 0; a < TA; ++a)
#pragma unroll
            for (int b = 0; b < TB; ++b) {
                if constexpr (EXACT)
                    acc[a][b] = mfma_bf16(d[a][2 - rd], c0[b], acc[a][b]);
                else
                    acc[a][b] = mfma_bf16(d[a][rd == 1 ? 1 : 0], rd == 0 ? c1[b] : c0[b], acc[a][b]);
            }
}
template <int TA, int TB = TA>
__device__ __forceinline__ void mma_rounds(floatx4 (&acc)[TA][TB], const bf16x8 (&d)[TA][3], const bf16x8 (&c0)[TB],
                                           const bf16x8 (&c1)[TB], bool exact) {
    if (exact)
        mma_rounds_<TA, TB, true>(acc, d, c0, c1);
    else
        mma_rounds_<TA, TB, false>(acc, d, c0, c1);
}
__device__ __forceinline__ void wave_sum_f(floatx4 (&acc)[kBT][kBT], float *red, float (&out)[kBT * kBT]) {
    constexpr int NC = kBT * kBT * 4;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < kBT; ++a)
#pragma unroll
        for (int b = 0; b < kBT; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) red[((w * NC) + (a * kBT + b) * 4 + e) * 64 + lane] = acc[a][b][e];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NC / 4; ++t) {
        const int q = 4 * w + t;
        float v = red[q * 64 + lane];
#pragma unroll
        for (int u = 1; u < 4; ++u) v += red[((u * NC) + q) * 64 + lane];
        out[t] = v;
    }
    __syncthreads();
}

// v[0..3] as three BF16 pieces at P[o .. o+3] (+ pp, + 2 pp): one 8-byte
// store per piece (o is a multiple of 4)
typedef __attribute__((ext_vector_type(4))) unsigned short u16x4;
__device__ __forceinline__ void store_pieces4(unsigned short *__restrict__ P, int64_t pp, int64_t o, const float *v) {
    u16x4 p0, p1, p2;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const unsigned short h0 = bf16_rn(v[e]);
        const float r1 = v[e] - bf16_f(h0);
        const unsigned short h1 = bf16_rn(r1);
        p0[e] = h0;
        p1[e] = h1;
        p2[e] = bf16_rn(r1 - bf16_f(h1));
    }
    *(u16x4 *)(P + o) = p0;
    *(u16x4 *)(P + pp + o) = p1;
    *(u16x4 *)(P + 2 * pp + o) = p2;
}

// X (column-major n x n) into three BF16 pieces of X^T: P[j np + k] =
// piece(X[k + j n]); with R, also of X row-major: R[k np + j]
template <class IT>
__global__ __launch_bounds__(256) void k_qap_split(int n, int np, const IT *__restrict__ x,
                                                   unsigned short *__restrict__ P, unsigned short *__restrict__ R) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x, nn = (int64_t)n * n, pp = (int64_t)np * np;
    if (idx >= nn) return;
    const int k = (int)(idx % n), j = (int)(idx / n);
    const float v = (float)x[idx];
    const unsigned short h0 = bf16_rn(v);
    const float r1 = v - bf16_f(h0);
    const unsigned short h1 = bf16_rn(r1);
    const unsigned short h2 = bf16_rn(r1 - bf16_f(h1));
    const int64_t o = (int64_t)j * np + k;
    P[o] = h0;
    P[pp + o] = h1;
    P[2 * pp + o] = h2;
    if (R) {
        const int64_t t = (int64_t)k * np + j;
        R[t] = h0;
        R[pp + t] = h1;
        R[2 * pp + t] = h2;
    }
}

// pass 1: Y^T = X^T A^T (and Z^T = X^T S^T for a full S); Y stored as
// pieces, Z as FP32 column-major
static __global__ __launch_bounds__(256) void k_qapbfb_p1(QapArgs g, const unsigned short *__restrict__ XP,
                                                          unsigned short *__restrict__ YP, float *__restrict__ Z) {
    __shared__ float red[4 * kBT * kBT * 4 * 64];
    const int n = g.n, np = g.np, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int64_t pp = (int64_t)np * np;
    const int j0 = blockIdx.x * kBE, i0 = blockIdx.y * kBE;  // Y^T rows j, columns i
    floatx4 ay[kBT][kBT], az[kBT][kBT];
#pragma unroll
    for (int a = 0; a < kBT; ++a)
#pragma unroll
        for (int b = 0; b < kBT; ++b) {
            ay[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
            az[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
    for (int ks = w; ks * 32 < np; ks += 4) {
        const int k = ks * 32 + 8 * lg;
        bf16x8 d[kBT][3], ca0[kBT], ca1[kBT], cs0[kBT], cs1[kBT];
#pragma unroll
        for (int t = 0; t < kBT; ++t) {
            const int64_t od = (int64_t)(j0 + 16 * t + li) * np + k, oc = (int64_t)(i0 + 16 * t + li) * np + k;
            d[t][0] = ld_frag(XP, od);
            d[t][1] = ld_frag(XP, pp + od);
            d[t][2] = ld_frag(XP, 2 * pp + od);
            ca0[t] = ld_frag(g.A0, oc);
            ca1[t] = g.a_exact ? ca0[t] : ld_frag(g.A1, oc);
            if (g.s_full) {
                cs0[t] = ld_frag(g.S0, oc);
                cs1[t] = g.s_exact ? cs0[t] : ld_frag(g.S1, oc);
            }
        }
        mma_rounds<kBT>(ay, d, ca0, ca1, g.a_exact);
        if (g.s_full) mma_rounds<kBT>(az, d, cs0, cs1, g.s_exact);
    }
    float vy[kBT * kBT], vz[kBT * kBT];
    wave_sum_f(ay, red, vy);
    if (g.s_full) wave_sum_f(az, red, vz);
    // wave w holds tile w: four consecutive j per lane, stored as one 8-byte
    // write per piece.  Rows and columns past n are exact zeros (zero-padded
    // operands), so the whole padded tile is stored without a bound check.
    const int a = w / kBT, b = w % kBT;
    const int jb = j0 + 16 * a + 4 * lg, i = i0 + 16 * b + li;
    store_pieces4(YP, pp, (int64_t)i * np + jb, vy);
    if (g.s_full) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (i < n && jb + e < n) Z[i + (int64_t)(jb + e) * n] = vz[e];
    }
}

// pass 2: R = Y B - X T (full T), the diagonal terms and the epilogue
template <class IT, class Epi>
__global__ __launch_bounds__(256) void k_qapbfb_p2(QapArgs g, const IT *__restrict__ x,
                                                   const unsigned short *__restrict__ YP,
                                                   const unsigned short *__restrict__ XR,
                                                   const float *__restrict__ Z, Epi epi) {
    __shared__ float red[4 * kBT * kBT * 4 * 64];
    const int n = g.n, np = g.np, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int64_t pp = (int64_t)np * np;
    const int i0 = blockIdx.x * kBE, j0 = blockIdx.y * kBE;
    floatx4 acc[kBT][kBT], at[kBT][kBT];
#pragma unroll
    for (int a = 0; a < kBT; ++a)
#pragma unroll
        for (int b = 0; b < kBT; ++b) {
            acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
            at[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
    for (int ks = w; ks * 32 < np; ks += 4) {
        const int k = ks * 32 + 8 * lg;
        bf16x8 d[kBT][3], dx[kBT][3], cb0[kBT], cb1[kBT], ct0[kBT], ct1[kBT];
#pragma unroll
        for (int t = 0; t < kBT; ++t) {
            const int64_t od = (int64_t)(i0 + 16 * t + li) * np + k, oc = (int64_t)(j0 + 16 * t + li) * np + k;
            d[t][0] = ld_frag(YP, od);
            d[t][1] = ld_frag(YP, pp + od);
            d[t][2] = ld_frag(YP, 2 * pp + od);
            cb0[t] = ld_frag(g.B0, oc);
            cb1[t] = g.b_exact ? cb0[t] : ld_frag(g.B1, oc);
            if (g.t_full) {
                dx[t][0] = ld_frag(XR, od);
                dx[t][1] = ld_frag(XR, pp + od);
                dx[t][2] = ld_frag(XR, 2 * pp + od);
                ct0[t] = ld_frag(g.T0, oc);
                ct1[t] = g.t_exact ? ct0[t] : ld_frag(g.T1, oc);
            }
        }
        mma_rounds<kBT>(acc, d, cb0, cb1, g.b_exact);
        if (g.t_full) mma_rounds<kBT>(at, dx, ct0, ct1, g.t_exact);
    }
    float v[kBT * kBT], vt[kBT * kBT];
    wave_sum_f(acc, red, v);
    if (g.t_full) wave_sum_f(at, red, vt);
#pragma unroll
    for (int t = 0; t < kBT * kBT; ++t) {
        const int q = 4 * w + t, e = q & 3, ab = q >> 2, a = ab / kBT, b = ab % kBT;
        const int r = i0 + 16 * a + 4 * lg + e, j = j0 + 16 * b + li;
        if (r >= n || j >= n) continue;
        const int64_t o = r + (int64_t)j * n;
        double u = (double)v[t];
        if (g.t_full) u -= (double)vt[t];
        if (g.s_full)
            u -= (double)Z[o];
        else
            u -= g.sd[r] * (double)x[o];
        if (!g.t_full) u -= (double)x[o] * g.td[j];
        epi((int32_t)o, 2.0 * u);
    }
}

// ---------------- split-BF16 x3 passes, n >= 1024, diagonal S and T ---------
// LDS-tiled GEMM form of the two passes for the QAP relaxation's shifts
// (diagonal S, T): a block of 2 x 2 waves owns a (32 TA) x (32 TB) output
// (data rows x coefficient rows), each wave a (16 TA) x (16 TB) part as
// TA x TB MFMA tiles.  Per K step of KS the block stages its data rows
// (three pieces) and coefficient rows (one or two pieces) in LDS with
// 16-byte loads -- each staged fragment then feeds the MFMA tiles of two
// waves -- and every wave reads its fragments from LDS.  Per 32-deep step a
// wave reads 3 TA + TB (exact coefficients) or 3 TA + 2 TB fragments for
// 3 TA TB MFMAs, so the wider coefficient side (TB = 4) cuts LDS reads per
// MFMA.  Rows are padded to KS + 8 halfwords to spread the 16 lanes reading
// one fragment column over the banks.  Piece arrays are qnp x qnp with qnp
// a multiple of 128, so no access needs a bound check; results outside
// n x n are not stored.
constexpr int kQapLdsMin = 1024;       // n from which these passes run

// The staging of one K step, split into its global loads (into registers,
// issued before the previous step's MFMAs so they are in flight meanwhile)
// and its LDS stores: three data pieces of RD rows and the coefficient's one
// or two pieces of RC rows, 16-byte chunks, rows * KS / 8 chunks per piece.
// KS = 32: LDS rows of 32 halfwords (64 bytes, no padding) with chunk c of
// row r at position c ^ ((r >> 1) & 3), measured free of bank conflicts both
// for these stores (four lanes per row) and for the fragment reads (sixteen
// rows, one chunk each), where padded rows conflict two ways on gfx950.
// Other KS: rows padded to KS + 8 halfwords.
template <int KS>
__device__ __forceinline__ int lds_off(int row, int c) {
    if constexpr (KS == 32)
        return row * 32 + ((c ^ (row >> 1)) & 3) * 8;
    else
        return row * (KS + 8) + c * 8;
}
template <int KS>
constexpr int lds_row() { return KS == 32 ? KS : KS + 8; }

template <int TA, int TB, int KS>
struct LdsStage {
    static constexpr int RD = 32 * TA, RC = 32 * TB, CR = KS / 8, LS = lds_row<KS>();
    static constexpr int CPD = RD * CR / 256, CPC = RC * CR / 256;  // chunks per thread per piece
    static_assert(RD * CR % 256 == 0 && RC * CR % 256 == 0, "chunks per piece must be a multiple of the block");
    u16x8 d[3][CPD], c[2][CPC];
    __device__ __forceinline__ void load(const unsigned short *__restrict__ D, int64_t pp, const unsigned short *C0,
                                         const unsigned short *C1, bool two, int np, int drow0, int crow0, int k0) {
#pragma unroll
        for (int t = 0; t < CPD; ++t) {
            const int ch = t * 256 + threadIdx.x, row = ch / CR, kc = (ch % CR) * 8;
            const int64_t od = (int64_t)(drow0 + row) * np + k0 + kc;
#pragma unroll
            for (int p = 0; p < 3; ++p) d[p][t] = *(const u16x8 *)(D + p * pp + od);
        }
#pragma unroll
        for (int t = 0; t < CPC; ++t) {
            const int ch = t * 256 + threadIdx.x, row = ch / CR, kc = (ch % CR) * 8;
            const int64_t oc = (int64_t)(crow0 + row) * np + k0 + kc;
            c[0][t] = *(const u16x8 *)(C0 + oc);
            if (two) c[1][t] = *(const u16x8 *)(C1 + oc);
        }
    }
    __device__ __forceinline__ void store(unsigned short *Ds, unsigned short *Cs, bool two) const {
#pragma unroll
        for (int t = 0; t < CPD; ++t) {
            const int ch = t * 256 + threadIdx.x, row = ch / CR, kc = (ch % CR) * 8;
#pragma unroll
            for (int p = 0; p < 3; ++p) *(u16x8 *)(Ds + p * RD * LS + lds_off<KS>(row, kc / 8)) = d[p][t];
        }
#pragma unroll
        for (int t = 0; t < CPC; ++t) {
            const int ch = t * 256 + threadIdx.x, row = ch / CR, kc = (ch % CR) * 8;
            *(u16x8 *)(Cs + lds_off<KS>(row, kc / 8)) = c[0][t];
            if (two) *(u16x8 *)(Cs + RC * LS + lds_off<KS>(row, kc / 8)) = c[1][t];
        }
    }
};

// one K step of a wave: acc[a][b] += D_a C_b over the staged pieces, in
// 32-deep MFMA steps
template <int TA, int TB, int KS>
__device__ __forceinline__ void lds_step(floatx4 (&acc)[TA][TB], const unsigned short *Ds, const unsigned short *Cs,
                                         bool exact) {
    constexpr int RD = 32 * TA, RC = 32 * TB, LS = lds_row<KS>();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lg = lane >> 4;
    const int dr = (w & 1) * 16 * TA, cr = (w >> 1) * 16 * TB;
#pragma unroll
    for (int kq = 0; kq < KS / 32; ++kq) {
        bf16x8 c0[TB], c1[TB];
#pragma unroll
        for (int b = 0; b < TB; ++b) {
            const int o = lds_off<KS>(cr + 16 * b + li, 4 * kq + lg);
            c0[b] = __builtin_bit_cast(bf16x8, *(const u16x8 *)(Cs + o));
            c1[b] = exact ? c0[b] : __builtin_bit_cast(bf16x8, *(const u16x8 *)(Cs + RC * LS + o));
        }
        bf16x8 d[TA][3];
#pragma unroll
        for (int a = 0; a < TA; ++a) {
            const int o = lds_off<KS>(dr + 16 * a + li, 4 * kq + lg);
#pragma unroll
            for (int p = 0; p < 3; ++p) d[a][p] = __builtin_bit_cast(bf16x8, *(const u16x8 *)(Ds + p * RD * LS + o));
        }
        mma_rounds<TA, TB>(acc, d, c0, c1, exact);
    }
}

// The K loop of one block: acc += D rows drow0.. times C rows crow0.. over
// all np.  NB = 1: one LDS stage, two barriers per K step; NB = 2: two
// stages, the next step's pieces are stored into the other stage after this
// step's MFMAs and one barrier per step separates them (a stage is
// overwritten only after every wave has passed the barrier that ended the
// MFMAs reading it).
template <int TA, int TB, int KS, int NB>
__device__ __forceinline__ void lds_gemm(floatx4 (&acc)[TA][TB], const unsigned short *__restrict__ D, int64_t pp,
                                         const unsigned short *C0, const unsigned short *C1, bool exact, int np,
                                         int drow0, int crow0, unsigned short *Ds, unsigned short *Cs) {
    constexpr int DSZ = 3 * 32 * TA * lds_row<KS>(), CSZ = 2 * 32 * TB * lds_row<KS>();
    const bool two = !exact;
    LdsStage<TA, TB, KS> st;
    st.load(D, pp, C0, C1, two, np, drow0, crow0, 0);
    if constexpr (NB == 1) {
        for (int k0 = 0; k0 < np; k0 += KS) {
            st.store(Ds, Cs, two);
            __syncthreads();
            if (k0 + KS < np) st.load(D, pp, C0, C1, two, np, drow0, crow0, k0 + KS);
            lds_step<TA, TB, KS>(acc, Ds, Cs, exact);
            __syncthreads();
        }
    } else {
        st.store(Ds, Cs, two);
        __syncthreads();
        int cur = 0;
        for (int k0 = 0; k0 < np; k0 += KS) {
            const bool more = k0 + KS < np;
            if (more) st.load(D, pp, C0, C1, two, np, drow0, crow0, k0 + KS);
            lds_step<TA, TB, KS>(acc, Ds + cur * DSZ, Cs + cur * CSZ, exact);
            if (more) st.store(Ds + (cur ^ 1) * DSZ, Cs + (cur ^ 1) * CSZ, two);
            __syncthreads();
            cur ^= 1;
        }
    }
}

// pass 1: Y^T = X^T A^T, stored as row-major Y pieces; block (x, y) owns
// columns j0.. of Y (data rows, 32 TA) and rows i0.. (coefficient rows, 32 TB)
template <int TA, int TB, int KS, int NB>
__global__ __launch_bounds__(256) void k_qapbfl_p1(QapArgs g, const unsigned short *__restrict__ XP,
                                                   unsigned short *__restrict__ YP) {
    constexpr int RD = 32 * TA, RC = 32 * TB;
    __shared__ __attribute__((aligned(16))) unsigned short Ds[NB * 3 * RD * lds_row<KS>()],
        Cs[NB * 2 * RC * lds_row<KS>()];
    const int np = g.np, lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lg = lane >> 4;
    const int64_t pp = (int64_t)np * np;
    const int j0 = blockIdx.x * RD, i0 = blockIdx.y * RC;
    floatx4 acc[TA][TB];
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
        for (int b = 0; b < TB; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    lds_gemm<TA, TB, KS, NB>(acc, XP, pp, g.A0, g.A1, g.a_exact, np, j0, i0, Ds, Cs);
    // four consecutive j per lane: one 8-byte store per piece; rows and
    // columns past n are exact zeros, so the padded tile is stored whole
    const int jr = j0 + (w & 1) * 16 * TA, ic = i0 + (w >> 1) * 16 * TB;
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
        for (int b = 0; b < TB; ++b) {
            const float v[4] = {acc[a][b][0], acc[a][b][1], acc[a][b][2], acc[a][b][3]};
            store_pieces4(YP, pp, (int64_t)(ic + 16 * b + li) * np + jr + 16 * a + 4 * lg, v);
        }
}

// pass 2: R = Y B, then - s_i X_ij - X_ij t_j and the epilogue; block (x, y)
// owns rows i0.. (data rows, 32 TA) and columns j0.. (coefficient rows, 32 TB)
template <int TA, int TB, int KS, int NB, class IT, class Epi>
__global__ __launch_bounds__(256) void k_qapbfl_p2(QapArgs g, const IT *__restrict__ x,
                                                   const unsigned short *__restrict__ YP, Epi epi) {
    constexpr int RD = 32 * TA, RC = 32 * TB;
    __shared__ __attribute__((aligned(16))) unsigned short Ds[NB * 3 * RD * lds_row<KS>()],
        Cs[NB * 2 * RC * lds_row<KS>()];
    const int n = g.n, np = g.np, lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lg = lane >> 4;
    const int64_t pp = (int64_t)np * np;
    const int i0 = blockIdx.x * RD, j0 = blockIdx.y * RC;
    floatx4 acc[TA][TB];
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
        for (int b = 0; b < TB; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    lds_gemm<TA, TB, KS, NB>(acc, YP, pp, g.B0, g.B1, g.b_exact, np, i0, j0, Ds, Cs);
    const int ir = i0 + (w & 1) * 16 * TA, jc = j0 + (w >> 1) * 16 * TB;
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
        for (int b = 0; b < TB; ++b) {
            const int j = jc + 16 * b + li;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = ir + 16 * a + 4 * lg + e;
                if (r >= n || j >= n) continue;
                const int64_t o = r + (int64_t)j * n;
                const double xo = (double)x[o];
                const double u = (double)acc[a][b][e] - g.sd[r] * xo - xo * g.td[j];
                epi((int32_t)o, 2.0 * u);
            }
        }
}

// Q x with the result handed to epi(i, (Q x)_i): FP64 / split-BF16 x3
// (tmp, when given, replaces the operator's two n x n intermediates: 2 n^2
// doubles)
// kernel families of the passes (hprqp_qop::qfam64 / qfam16)
enum QapFamily { kQapBySize = 0, kQapFamSmall = 1, kQapFamBlocked = 2, kQapFamLds = 3, kQapFamLdsWide = 4 };

// the family a pass runs: the operator's chosen one where it applies, else
// by size.  The one-tile passes serve n < 256 (their pieces are padded to 32,
// and the solver's one-apply step needs their tile sums there); the
// LDS-tiled ones need diagonal S and T; the 64 x 128 split-BF16 blocks need
// two blocks per CU (the grid np/64 x np/128 >= 512).
static inline int qap_family(const QapArgs &g, int chosen, bool bf16) {
    const bool diag = !g.s_full && !g.t_full;
    if (g.n < kQapBlockedMin) return kQapFamSmall;
    if (chosen == kQapFamBlocked) return kQapFamBlocked;
    if (chosen == kQapFamLds && diag) return kQapFamLds;
    if (chosen == kQapFamLdsWide && diag && bf16 && (int64_t)(g.np / 64) * (g.np / 128) >= 512)
        return kQapFamLdsWide;
    if (g.n >= kQapLdsMin && diag)
        return bf16 && (int64_t)(g.np / 64) * (g.np / 128) >= 512 ? kQapFamLdsWide : kQapFamLds;
    return kQapFamBlocked;
}

template <class Epi>
void qap_launch_f64(hipStream_t s, const hprqp_qop *q, const double *X, const Epi &epi, double *tmp = nullptr) {
    const QapArgs g = qap_args(q);
    const int nt = (g.n + 15) / 16;
    const int64_t nn = (int64_t)g.n * g.n;
    double *const base = tmp ? tmp : q->qtemp.p;
    double *Y = base, *Z = base + nn;
    const int fam = qap_family(g, q->qfam64, false);
    if (fam == kQapFamLds) {
        // 64-wide blocks (measured faster than 128-wide up to n = 4096:
        // four blocks per CU hide each other's K-step barriers)
        const int nb = (g.n + 63) / 64;
        launch_k(k_qap64l_p1<2, 32>, dim3(nb, nb), dim3(256), 0, s, g, X, Y);
        HPRQP_HIP(hipGetLastError());
        launch_k(k_qap64l_p2<2, 32, Epi>, dim3(nb, nb), dim3(256), 0, s, g, X, (const double *)Y, epi);
        HPRQP_HIP(hipGetLastError());
        return;
    }
    if (fam == kQapFamBlocked) {
        const int nb = (g.n + kBE - 1) / kBE;
        launch_k(k_qap64b_p1, dim3(nb, nb), dim3(256), 0, s, g, X, Y, Z);
        HPRQP_HIP(hipGetLastError());
        launch_k(k_qap64b_p2<Epi>, dim3(nb, nb), dim3(256), 0, s, g, X, (const double *)Y, (const double *)Z, epi);
        HPRQP_HIP(hipGetLastError());
        return;
    }
    const dim3 grid(nt, nt);
    launch_k(k_qap64s_p1, grid, dim3(256), 0, s, g, X, Y, Z);
    HPRQP_HIP(hipGetLastError());
    launch_k(k_qap64s_p2<Epi>, grid, dim3(256), 0, s, g, X, (const double *)Y, (const double *)Z, epi);
    HPRQP_HIP(hipGetLastError());
}

template <class IT, class Epi>
void qap_launch_bf16x3(hipStream_t s, const hprqp_qop *q, const IT *X, const Epi &epi) {
    const QapArgs g = qap_args(q);
    const int nt = (g.n + 15) / 16;
    const int64_t nn = (int64_t)g.n * g.n;
    float *Y = q->qtemp32.p, *Z = q->qtemp32.p + nn;
    const int fam = qap_family(g, q->qfam16, true);
    if (fam == kQapFamLds || fam == kQapFamLdsWide) {
        launch_k(k_qap_split<IT>, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, g.n, g.np, X, q->qxp.p,
                 (unsigned short *)nullptr);
        HPRQP_HIP(hipGetLastError());
        const unsigned short *xp = q->qxp.p;
        if (fam == kQapFamLdsWide) {
            // 64 x 128 blocks, each wave 2 x 4 tiles: fewer LDS reads per
            // MFMA (measured 13 % faster than 64 x 64 at n = 2048), used
            // once they make two blocks per CU
            const dim3 grid(g.np / 64, g.np / 128);
            launch_k(k_qapbfl_p1<2, 4, 32, 1>, grid, dim3(256), 0, s, g, xp, q->qyp.p);
            HPRQP_HIP(hipGetLastError());
            launch_k(k_qapbfl_p2<2, 4, 32, 1, IT, Epi>, grid, dim3(256), 0, s, g, X, (const unsigned short *)q->qyp.p,
                     epi);
        } else {
            const dim3 grid(g.np / 64, g.np / 64);
            launch_k(k_qapbfl_p1<2, 2, 64, 1>, grid, dim3(256), 0, s, g, xp, q->qyp.p);
            HPRQP_HIP(hipGetLastError());
            launch_k(k_qapbfl_p2<2, 2, 64, 1, IT, Epi>, grid, dim3(256), 0, s, g, X, (const unsigned short *)q->qyp.p,
                     epi);
        }
        HPRQP_HIP(hipGetLastError());
        return;
    }
    if (fam == kQapFamBlocked) {
        const int nb = g.np / kBE;
        launch_k(k_qap_split<IT>, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, g.n, g.np, X, q->qxp.p,
                 g.t_full ? q->qxr.p : (unsigned short *)nullptr);
        HPRQP_HIP(hipGetLastError());
        launch_k(k_qapbfb_p1, dim3(nb, nb), dim3(256), 0, s, g, (const unsigned short *)q->qxp.p, q->qyp.p, Z);
        HPRQP_HIP(hipGetLastError());
        launch_k(k_qapbfb_p2<IT, Epi>, dim3(nb, nb), dim3(256), 0, s, g, X, (const unsigned short *)q->qyp.p,
                 (const unsigned short *)q->qxr.p, (const float *)Z, epi);
        HPRQP_HIP(hipGetLastError());
        return;
    }
    const dim3 grid(nt, nt);
    launch_k(k_qapbfs_p1<IT>, grid, dim3(256), 0, s, g, X, Y, Z);
    HPRQP_HIP(hipGetLastError());
    launch_k(k_qapbfs_p2<IT, Epi>, grid, dim3(256), 0, s, g, X, (const float *)Y, (const float *)Z, epi);
    HPRQP_HIP(hipGetLastError());
}

}  // namespace hprqp
