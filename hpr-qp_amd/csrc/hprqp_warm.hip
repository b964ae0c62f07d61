// hprqp_warm.hip -- the learned warm start of a whole batch, on the device.
//
// For every instance of a packed batch (Q_i = diag(d_i) + U_i U_i^T), one
// workgroup computes the 12 per-variable features the warm-start MLP was
// trained on, normalises them with the training statistics and writes them
// as rows of one FP32 matrix; the MLP runs over all rows at once; a second
// per-instance kernel turns the outputs into the packed (x0, y0) that the
// batched solver's process_initial_points! step (src/algorithm.jl:1783-1836)
// receives.  Nothing travels through the host.
//
// Features of variable i of an instance with n variables, k factors
// (cm, cs = mean and population standard deviation of c, cs += 1e-12;
// ubar = column means of U; q_ii = d_i + |U_i.|^2; r_i = -c_i / q_ii;
// rank(v)_i = #{j : v_j < v_i} + #{j < i : v_j == v_i}, divided by
// max(n - 1, 1)):
//   (c_i - cm)/cs, 20 c_i, 10 d_i, 3 (q_ii - d_i), 30 U_i. ubar, 0.5 r_i,
//   rank(c)_i, rank(-r)_i, 100/n, k/50, min(n u_i, 100)/100 (u_i := 1 if
//   infinite), 50 cs.
// Outputs o (rows x 2): x0_i = clamp(o_i0, 0, 1) u_i (u_i := 1 if
// infinite); y0 = mean_i(o_i1) / y_scale on every row of the instance.
#include <cmath>

#include "hprqp_internal.hpp"

namespace hprqp {

constexpr int kWB = 256;       // threads per instance workgroup
constexpr int kRankTile = 1024;  // keys staged in LDS per pass of the rank count

__device__ __forceinline__ double wsum(double v, double *sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
}

struct WarmArgs {
    const int32_t *n, *k, *order;
    const int64_t *noff, *uoff;
    const double *U, *d, *c, *u;
    const float *mean, *stdv;
    double *ratio;  // Ntot scratch: r_i
    float *F;       // Ntot x 12 feature rows
};

__global__ __launch_bounds__(kWB) void k_warm_features(WarmArgs a) {
    __shared__ double sh[4];
    __shared__ double ubar[64];
    __shared__ double kc[kRankTile], kr[kRankTile];
    const int b = a.order[blockIdx.x];
    const int32_t n = a.n[b], k = a.k[b];
    const int64_t o = a.noff[b];
    const double *U = a.U + a.uoff[b], *d = a.d + o, *c = a.c + o, *u = a.u + o;
    double *ratio = a.ratio + o;
    float *F = a.F + o * kWarmFeatures;
    // instance summaries
    double s = 0.0;
    for (int32_t i = threadIdx.x; i < n; i += kWB) s += c[i];
    const double cm = wsum(s, sh) / n;
    s = 0.0;
    for (int32_t i = threadIdx.x; i < n; i += kWB) {
        const double t = c[i] - cm;
        s += t * t;
    }
    const double cs = sqrt(wsum(s, sh) / n) + 1e-12;
    for (int32_t j = 0; j < k; ++j) {
        s = 0.0;
        for (int32_t i = threadIdx.x; i < n; i += kWB) s += U[(int64_t)j * n + i];
        const double t = wsum(s, sh);
        if (threadIdx.x == 0) ubar[j] = t / n;
    }
    __syncthreads();
    // per-variable features except the two ranks
    const double inv_nm1 = 1.0 / (double)(n > 1 ? n - 1 : 1);
    for (int32_t i = threadIdx.x; i < n; i += kWB) {
        double q2 = 0.0, ex = 0.0;
        for (int32_t j = 0; j < k; ++j) {
            const double v = U[(int64_t)j * n + i];
            q2 += v * v;
            ex += v * ubar[j];
        }
        const double qii = d[i] + q2;
        const double r = -c[i] / qii;
        ratio[i] = r;
        const double ui = isfinite(u[i]) ? u[i] : 1.0;
        double f[kWarmFeatures];
        f[0] = (c[i] - cm) / cs;
        f[1] = c[i] * 20.0;
        f[2] = d[i] * 10.0;
        f[3] = (qii - d[i]) * 3.0;
        f[4] = ex * 30.0;
        f[5] = r * 0.5;
        f[6] = 0.0;
        f[7] = 0.0;
        f[8] = 100.0 / n;
        f[9] = k / 50.0;
        f[10] = fmin(ui * n, 100.0) / 100.0;
        f[11] = cs * 50.0;
        float *row = F + (int64_t)i * kWarmFeatures;
#pragma unroll
        for (int q = 0; q < kWarmFeatures; ++q) {
            if (q == 6 || q == 7) continue;
            row[q] = ((float)f[q] - a.mean[q]) / a.stdv[q];
        }
    }
    __syncthreads();
    // ranks by counting, keys staged through LDS a tile at a time
    for (int32_t i0 = 0; i0 < n; i0 += kWB) {
        const int32_t i = i0 + threadIdx.x;
        const bool live = i < n;
        const double ci = live ? c[i] : 0.0, ri = live ? -ratio[i] : 0.0;
        int32_t cc = 0, cr = 0;
        for (int32_t j0 = 0; j0 < n; j0 += kRankTile) {
            const int32_t cnt = min(kRankTile, n - j0);
            __syncthreads();
            for (int32_t t = threadIdx.x; t < cnt; t += kWB) {
                kc[t] = c[j0 + t];
                kr[t] = -ratio[j0 + t];
            }
            __syncthreads();
            if (live) {
                for (int32_t t = 0; t < cnt; ++t) {
                    const int32_t j = j0 + t;
                    const double vc = kc[t], vr = kr[t];
                    cc += (vc < ci) || (vc == ci && j < i);
                    cr += (vr < ri) || (vr == ri && j < i);
                }
            }
        }
        if (live) {
            float *row = F + (int64_t)i * kWarmFeatures;
            row[6] = ((float)(cc * inv_nm1) - a.mean[6]) / a.stdv[6];
            row[7] = ((float)(cr * inv_nm1) - a.mean[7]) / a.stdv[7];
        }
    }
}

__global__ __launch_bounds__(kWB) void k_warm_finish(const int32_t *nv, const int32_t *mv, const int64_t *noff,
                                                     const int64_t *moff, const double *u, const float *O,
                                                     double y_scale, double *x0, double *y0) {
    __shared__ double sh[4];
    const int b = blockIdx.x;
    const int32_t n = nv[b], m = mv[b];
    const int64_t o = noff[b];
    double s = 0.0;
    for (int32_t i = threadIdx.x; i < n; i += kWB) {
        const double ui = isfinite(u[o + i]) ? u[o + i] : 1.0;
        const double v = (double)O[2 * (o + i)];
        x0[o + i] = fmin(fmax(v, 0.0), 1.0) * ui;
        s += (double)O[2 * (o + i) + 1];
    }
    const double ybar = wsum(s, sh) / n / y_scale;
    for (int32_t r = threadIdx.x; r < m; r += kWB) y0[moff[b] + r] = ybar;
}

void batch_predict(hprqp_ctx *ctx, hprqp_batch *B, hprqp_mlp *M, const float *feat_mean, const float *feat_std,
                   double y_scale, double *x0, double *y0) {
    HPRQP_REQUIRE(B->qkind == 0, HPRQP_E_UNSUP,
                  "the learned warm start takes diagonal-plus-low-rank batches (hprqp_batch_create)");
    HPRQP_REQUIRE(M->nin == kWarmFeatures && M->nout == 2, HPRQP_E_ARG,
                  "warm-start MLP must take 12 features and give 2 outputs");
    HPRQP_REQUIRE(y_scale > 0.0, HPRQP_E_ARG, "y_scale must be positive");
    for (int q = 0; q < kWarmFeatures; ++q)
        HPRQP_REQUIRE(std::isfinite(feat_mean[q]) && std::isfinite(feat_std[q]) && feat_std[q] > 0.f, HPRQP_E_ARG,
                      "feature statistics must be finite with positive standard deviations");
    hipStream_t s = ctx->stream;
    const int64_t N = B->Ntot;
    DevVec<float> stats(2 * kWarmFeatures), F(N * kWarmFeatures), O(2 * N);
    DevVec<double> ratio(N);
    HPRQP_HIP(hipMemcpyAsync(stats.p, feat_mean, sizeof(float) * kWarmFeatures, hipMemcpyHostToDevice, s));
    HPRQP_HIP(hipMemcpyAsync(stats.p + kWarmFeatures, feat_std, sizeof(float) * kWarmFeatures,
                             hipMemcpyHostToDevice, s));
    WarmArgs a{B->dn.p, B->dk.p, B->dorder.p, B->dnoff.p, B->duoff.p, B->U.p, B->d.p, B->c.p, B->u.p,
               stats.p, stats.p + kWarmFeatures, ratio.p, F.p};
    k_warm_features<<<B->count, kWB, 0, s>>>(a);
    HPRQP_HIP(hipGetLastError());
    mlp_forward_dev(s, M, F.p, N, O.p);
    k_warm_finish<<<B->count, kWB, 0, s>>>(B->dn.p, B->dm.p, B->dnoff.p, B->dmoff.p, B->u.p, O.p, y_scale, x0, y0);
    HPRQP_HIP(hipGetLastError());
    // the temporaries are freed on return: wait for the kernels that use them
    HPRQP_HIP(hipStreamSynchronize(s));
}

}  // namespace hprqp
