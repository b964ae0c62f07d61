// hprqp_qap.hip -- QAP relaxation operator Q(X) = 2(A X B - S X - X T)
// (src/Q_operators/QAP_operator.jl:52-76) on dense column-major nq x nq
// matrices, X = reshape(x, nq, nq).
//
// The operator is the only dense contraction of the solver, so it runs on
// the matrix cores.  One workgroup owns 16 rows of the result:
//   phase 1: Y = A[r,:] X (and Z = S[r,:] X when S is a full matrix) into LDS;
//   phase 2: R = Y B - X[r,:] T, then the epilogue subtracts Z (or s_i X_ij
//            for diagonal S) and X_ij t_j (diagonal T) and writes 2 R.
// Two precisions:
//   * FP64: v_mfma_f64_16x16x4f64, FP64 accumulation (plain FP64 solves and
//     every residual evaluation);
//   * split-BF16 x3: v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  Each
//     product a*b becomes three BF16 products: when the coefficient matrix
//     (A, B, S or T) is exactly representable in BF16 (integer flows and
//     distances are), the data operand is split into three BF16 pieces
//     b0 + b1 + b2 (24 significant bits, FP32 accuracy) and the terms are
//     a*b0 + a*b1 + a*b2; otherwise both are split in two and the terms are
//     a0*b0 + a0*b1 + a1*b0 (16 significant bits).
// Sizes above kMfmaMaxN fall back to LDS-tiled GEMMs (four launches).
#include <cmath>
#include <cstring>
#include <vector>

#include "hprqp_internal.hpp"

namespace hprqp {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) unsigned short u16x8;
typedef __attribute__((ext_vector_type(4))) float floatx4;
typedef __attribute__((ext_vector_type(4))) double doublex4;

constexpr int kMfmaMaxN = 512;  // padded size limit of the MFMA kernels (LDS)
constexpr int kRB = 16;         // result rows per workgroup

// ---------------- BF16 splitting ---------------------------------------
__device__ __forceinline__ unsigned short bf16_rn(float f) {
    const uint32_t u = __float_as_uint(f);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

static unsigned short bf16_rn_host(float f) {
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
__device__ __forceinline__ bf16x8 ld8(const unsigned short *p) {
    return __builtin_bit_cast(bf16x8, *(const u16x8 *)p);
}

struct QapArgs {
    int32_t n, np;
    const double *A, *Bt, *S, *Tt, *sd, *td;
    const unsigned short *A0, *A1, *B0, *B1, *S0, *S1, *T0, *T1;
    int a_exact, b_exact, s_exact, t_exact, s_full, t_full;
};

static QapArgs qap_args(const hprqp_qop *q) {
    QapArgs g{};
    g.n = q->nq;
    g.np = q->qnp;
    g.A = q->qa.p;
    g.Bt = q->qbt.p;
    g.S = q->qs.p;
    g.Tt = q->qtt.p;
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

// MFMA operand layouts (16 x 16 tiles, lane l, group g = l / 16, measured
// on gfx950 with lane-index probes):
//   bf16 16x16x32: left operand row l%16, k = 8g..8g+7; right operand column
//   l%16, same k; result element e (0..3) at row 4g + e, column l%16.
//   f64 16x16x4: left row l%16, k = g; right column l%16, k = g; result
//   element e at row g + 4e, column l%16.

// ---------------- split-BF16 x3 kernel ----------------------------------
// dynamic LDS: Ys, then Zs (S full), then Xs (T full); each kRB x (np + 4) floats
template <class IT, class OT>
__global__ __launch_bounds__(256) void k_qap_bf16x3(QapArgs g, const IT *__restrict__ x, OT *__restrict__ out) {
    extern __shared__ float lds[];
    const int n = g.n, np = g.np, ld = np + 4;
    float *Ys = lds;
    float *Zs = Ys + kRB * ld;
    float *Xs = Zs + (g.s_full ? kRB * ld : 0);
    const int r0 = blockIdx.x * kRB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int ntile = np / 16;
    if (g.t_full) {
        for (int e = threadIdx.x; e < kRB * np; e += 256) {
            const int i = e / np, k = e - i * np, r = r0 + i;
            Xs[i * ld + k] = (r < n && k < n) ? -(float)x[r + (int64_t)k * n] : 0.f;
        }
    }
    // phase 1: Y = A[r,:] X, Z = S[r,:] X
    for (int ct = wave; ct < ntile; ct += 4) {
        const int j = ct * 16 + li;
        floatx4 ay = {0.f, 0.f, 0.f, 0.f}, az = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < np; k0 += 32) {
            const int kk = k0 + 8 * lg;
            float xv[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = kk + e;
                xv[e] = (k < n && j < n) ? (float)x[k + (int64_t)j * n] : 0.f;
            }
            const int64_t ao = (int64_t)(r0 + li) * np + kk;
            ay = mma3_cl(ay, ld8(g.A0 + ao), ld8(g.A1 + ao), g.a_exact, xv);
            if (g.s_full) az = mma3_cl(az, ld8(g.S0 + ao), ld8(g.S1 + ao), g.s_exact, xv);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            Ys[(4 * lg + e) * ld + j] = ay[e];
            if (g.s_full) Zs[(4 * lg + e) * ld + j] = az[e];
        }
    }
    __syncthreads();
    // phase 2: R = Y B - X[r,:] T (Xs holds -X[r,:])
    for (int ct = wave; ct < ntile; ct += 4) {
        const int j = ct * 16 + li;
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < np; k0 += 32) {
            const int kk = k0 + 8 * lg;
            float yv[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) yv[e] = Ys[li * ld + kk + e];
            const int64_t bo = (int64_t)j * np + kk;
            acc = mma3_cr(acc, ld8(g.B0 + bo), ld8(g.B1 + bo), g.b_exact, yv);
            if (g.t_full) {
#pragma unroll
                for (int e = 0; e < 8; ++e) yv[e] = Xs[li * ld + kk + e];
                acc = mma3_cr(acc, ld8(g.T0 + bo), ld8(g.T1 + bo), g.t_exact, yv);
            }
        }
        if (j >= n) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = 4 * lg + e, r = r0 + i;
            if (r >= n) continue;
            const int64_t o = r + (int64_t)j * n;
            double v = (double)acc[e];
            if (g.s_full)
                v -= (double)Zs[i * ld + j];
            else
                v -= g.sd[r] * (double)x[o];
            if (!g.t_full) v -= (double)x[o] * g.td[j];
            out[o] = (OT)(2.0 * v);
        }
    }
}

// ---------------- FP64 kernel -------------------------------------------
// dynamic LDS: Ys, then Zs (S full); each kRB x (np + 1) doubles
__global__ __launch_bounds__(256) void k_qap_f64(QapArgs g, const double *__restrict__ x, double *__restrict__ out) {
    extern __shared__ double ldsd[];
    const int n = g.n, np = g.np, ld = np + 1;
    double *Ys = ldsd;
    double *Zs = Ys + kRB * ld;
    const int r0 = blockIdx.x * kRB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int ntile = np / 16;
    const int ri = r0 + li;
    // phase 1
    for (int ct = wave; ct < ntile; ct += 4) {
        const int j = ct * 16 + li;
        doublex4 ay = {0.0, 0.0, 0.0, 0.0}, az = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < n; k0 += 4) {
            const int k = k0 + lg;
            const bool kin = k < n;
            const double xv = (kin && j < n) ? x[k + (int64_t)j * n] : 0.0;
            const double av = (kin && ri < n) ? g.A[ri + (int64_t)k * n] : 0.0;
            ay = __builtin_amdgcn_mfma_f64_16x16x4f64(av, xv, ay, 0, 0, 0);
            if (g.s_full) {
                const double sv = (kin && ri < n) ? g.S[ri + (int64_t)k * n] : 0.0;
                az = __builtin_amdgcn_mfma_f64_16x16x4f64(sv, xv, az, 0, 0, 0);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            Ys[(lg + 4 * e) * ld + j] = ay[e];
            if (g.s_full) Zs[(lg + 4 * e) * ld + j] = az[e];
        }
    }
    __syncthreads();
    // phase 2
    for (int ct = wave; ct < ntile; ct += 4) {
        const int j = ct * 16 + li;
        doublex4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < n; k0 += 4) {
            const int k = k0 + lg;
            const bool kin = k < n;
            const double yv = kin ? Ys[li * ld + k] : 0.0;
            const double bv = (kin && j < n) ? g.Bt[(int64_t)k * n + j] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(yv, bv, acc, 0, 0, 0);
            if (g.t_full) {
                const double xv = (kin && ri < n) ? -x[ri + (int64_t)k * n] : 0.0;
                const double tv = (kin && j < n) ? g.Tt[(int64_t)k * n + j] : 0.0;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xv, tv, acc, 0, 0, 0);
            }
        }
        if (j >= n) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = lg + 4 * e, r = r0 + i;
            if (r >= n) continue;
            const int64_t o = r + (int64_t)j * n;
            double v = acc[e];
            if (g.s_full)
                v -= Zs[i * ld + j];
            else
                v -= g.sd[r] * x[o];
            if (!g.t_full) v -= x[o] * g.td[j];
            out[o] = 2.0 * v;
        }
    }
}

// ---------------- large sizes: LDS-tiled GEMMs ---------------------------
constexpr int kT = 32;  // output tile edge; 256 threads compute 32 x 32

// C = alpha * X Y + beta * C, all nq x nq column-major; accumulation in AT.
template <class AT, class TX, class TY, class TC>
__global__ __launch_bounds__(256) void k_gemm_nn(int nq, AT alpha, const TX *__restrict__ X,
                                                 const TY *__restrict__ Y, AT beta, TC *__restrict__ C) {
    __shared__ AT sX[kT][kT + 1];  // sX[k][i]
    __shared__ AT sY[kT][kT + 1];  // sY[k][j]
    const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;  // ty in [0, 8)
    const int i0 = blockIdx.x * kT, j0 = blockIdx.y * kT;
    AT acc[4] = {0, 0, 0, 0};
    for (int k0 = 0; k0 < nq; k0 += kT) {
        for (int r = ty; r < kT; r += 8) {
            const int i = i0 + tx, k = k0 + r;
            sX[r][tx] = (i < nq && k < nq) ? (AT)X[i + (int64_t)k * nq] : (AT)0;
            const int kk = k0 + tx, j = j0 + r;
            sY[tx][r] = (kk < nq && j < nq) ? (AT)Y[kk + (int64_t)j * nq] : (AT)0;
        }
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < kT; ++k) {
            const AT a = sX[k][tx];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] += a * sY[k][ty + 8 * q];
        }
        __syncthreads();
    }
    const int i = i0 + tx;
    if (i >= nq) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int j = j0 + ty + 8 * q;
        if (j < nq) {
            const int64_t o = i + (int64_t)j * nq;
            const AT prev = beta == (AT)0 ? (AT)0 : beta * (AT)C[o];
            C[o] = (TC)(alpha * acc[q] + prev);
        }
    }
}

template <class AT, class TX, class TY, class TC>
static void gemm(hipStream_t s, int nq, AT alpha, const TX *X, const TY *Y, AT beta, TC *C) {
    dim3 grid((nq + kT - 1) / kT, (nq + kT - 1) / kT);
    k_gemm_nn<AT, TX, TY, TC><<<grid, 256, 0, s>>>(nq, alpha, X, Y, beta, C);
    HPRQP_HIP(hipGetLastError());
}

template <class AT, class IT, class OT>
static void qap_tiled(hipStream_t s, const hprqp_qop *q, const IT *X, AT *tmp, OT *QX) {
    const int nq = q->nq;
    gemm<AT>(s, nq, (AT)1, q->qa.p, X, (AT)0, tmp);
    gemm<AT>(s, nq, (AT)2, tmp, q->qb.p, (AT)0, QX);
    gemm<AT>(s, nq, (AT)-2, q->qs.p, X, (AT)1, QX);
    gemm<AT>(s, nq, (AT)-2, X, q->qt.p, (AT)1, QX);
}

// ---------------- host side ---------------------------------------------
void qap_prepare(hipStream_t s, hprqp_qop *q, const double *A, const double *B, const double *S, const double *T) {
    const int n = q->nq;
    const int np = (n + 31) / 32 * 32;
    q->qnp = np;
    const int64_t nn = (int64_t)n * n, pp = (int64_t)np * np;
    auto is_diag = [&](const double *M) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                if (i != j && M[i + (int64_t)j * n] != 0.0) return false;
        return true;
    };
    q->qs_diag = is_diag(S);
    q->qt_diag = is_diag(T);
    std::vector<double> tr(nn);
    auto transpose = [&](const double *M) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) tr[(int64_t)i * n + j] = M[i + (int64_t)j * n];
    };
    transpose(B);
    q->qbt.alloc(nn);
    HPRQP_HIP(hipMemcpy(q->qbt.p, tr.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
    if (!q->qt_diag) {
        transpose(T);
        q->qtt.alloc(nn);
        HPRQP_HIP(hipMemcpy(q->qtt.p, tr.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
    }
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
    // split-BF16 pieces; row_major: piece[i * np + k] = M[i, k], else
    // piece[k + j * np] = M[k, j]; the matrix is exact when piece 1 is zero
    // and piece 0 reproduces it
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
                uint32_t hb = (uint32_t)h << 16;
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
    (void)s;
}

void qap_apply_f64(hipStream_t s, const hprqp_qop *q, const double *X, double *QX) {
    if (q->qnp > kMfmaMaxN) {
        qap_tiled<double>(s, q, X, q->qtemp.p, QX);
        return;
    }
    const QapArgs g = qap_args(q);
    const size_t lds = sizeof(double) * kRB * (g.np + 1) * (g.s_full ? 2 : 1);
    if (lds > 65536)
        HPRQP_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_qap_f64),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_qap_f64<<<(g.n + kRB - 1) / kRB, 256, lds, s>>>(g, X, QX);
    HPRQP_HIP(hipGetLastError());
}

template <class IT, class OT>
void qap_apply_bf16x3(hipStream_t s, const hprqp_qop *q, const IT *X, OT *QX) {
    if (q->qnp > kMfmaMaxN) {
        qap_tiled<float>(s, q, X, q->qtemp32.p, QX);
        return;
    }
    const QapArgs g = qap_args(q);
    const size_t lds = sizeof(float) * kRB * (g.np + 4) * (1 + g.s_full + g.t_full);
    if (lds > 65536)
        HPRQP_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_qap_bf16x3<IT, OT>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_qap_bf16x3<IT, OT><<<(g.n + kRB - 1) / kRB, 256, lds, s>>>(g, X, QX);
    HPRQP_HIP(hipGetLastError());
}

template void qap_apply_bf16x3<double, double>(hipStream_t, const hprqp_qop *, const double *, double *);
template void qap_apply_bf16x3<float, float>(hipStream_t, const hprqp_qop *, const float *, float *);

}  // namespace hprqp
