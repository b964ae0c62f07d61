// hprqp_mlp.hip -- the learned warm start: a per-variable MLP
//
//   features (n_in) -> 128 -> 128 -> 128 -> 2,  ReLU, FP32
//
// applied to every variable of an instance (shared weights, so the map is
// permutation-equivariant in the variables).  Output 0 is the variable's
// primal value (scaled), output 1 is averaged over the instance's variables
// to give the instance's dual value (permutation-invariant).  The
// prediction enters the solver only as initial_x / initial_y
// (process_initial_points!, src/algorithm.jl:1783-1836); the KKT residual
// still decides termination, so the tolerance is never relaxed.
//
// Training runs on the GPU with the kernels below: an LDS-tiled FP32 GEMM
// for forward (A W), weight gradients (A^T dH) and input gradients (dH W^T),
// bias/ReLU epilogues, the loss and its derivative, and Adam.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "hprqp_internal.hpp"

namespace hprqp {

constexpr int kMB = 64, kNB = 64, kKB = 16;

// C (M x N) = op(A) (M x K) * op(B) (K x N) + beta * C, row-major;
// op(A)(i,k) = TA ? A[k*lda + i] : A[i*lda + k], same for B.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_sgemm(int M, int N, int K, const float *__restrict__ A, int lda,
                                               const float *__restrict__ B, int ldb, float *__restrict__ C, int ldc,
                                               float beta) {
    __shared__ float As[kKB][kMB + 4];
    __shared__ float Bs[kKB][kNB + 4];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int i0 = blockIdx.y * kMB, j0 = blockIdx.x * kNB;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += kKB) {
        for (int e = threadIdx.x; e < kMB * kKB; e += 256) {
            const int ii = e % kMB, kk = e / kMB;
            const int i = i0 + ii, k = k0 + kk;
            float v = 0.f;
            if (i < M && k < K) v = TA ? A[(int64_t)k * lda + i] : A[(int64_t)i * lda + k];
            As[kk][ii] = v;
        }
        for (int e = threadIdx.x; e < kNB * kKB; e += 256) {
            const int jj = e % kNB, kk = e / kNB;
            const int j = j0 + jj, k = k0 + kk;
            float v = 0.f;
            if (j < N && k < K) v = TB ? B[(int64_t)j * ldb + k] : B[(int64_t)k * ldb + j];
            Bs[kk][jj] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kKB; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = As[kk][ty * 4 + q];
                b[q] = Bs[kk][tx * 4 + q];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] = fmaf(a[p], b[q], acc[p][q]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int i = i0 + ty * 4 + p;
        if (i >= M) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = j0 + tx * 4 + q;
            if (j >= N) continue;
            float *c = C + (int64_t)i * ldc + j;
            *c = acc[p][q] + (beta != 0.f ? beta * *c : 0.f);
        }
    }
}

template <bool TA, bool TB>
static void sgemm(hipStream_t s, int M, int N, int K, const float *A, int lda, const float *B, int ldb, float *C,
                  int ldc, float beta) {
    dim3 grid((N + kNB - 1) / kNB, (M + kMB - 1) / kMB);
    k_sgemm<TA, TB><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, beta);
    HPRQP_HIP(hipGetLastError());
}

__global__ void k_bias_act(int64_t rows, int w, float *__restrict__ H, const float *__restrict__ b, int relu) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= rows * w) return;
    const float v = H[e] + b[e % w];
    H[e] = relu ? fmaxf(v, 0.f) : v;
}
// dH *= (H > 0)
__global__ void k_relu_back(int64_t cnt, float *__restrict__ dH, const float *__restrict__ H) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < cnt && !(H[e] > 0.f)) dH[e] = 0.f;
}
// db[j] = sum_r dH[r, j]: one block per column, fixed order
__global__ void k_colsum(int64_t rows, int w, const float *__restrict__ dH, float *__restrict__ db) {
    const int j = blockIdx.x;
    double acc = 0.0;
    for (int64_t r = threadIdx.x; r < rows; r += 256) acc += dH[r * w + j];
    __shared__ double sh[256];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[j] = (float)sh[0];
}
// loss and dL/dO for O (rows x 2): MSE on column 0 vs tx; group means of
// column 1 vs ty (groups of `group` consecutive rows).  One thread per group.
__global__ void k_loss(int64_t groups, int group, const float *__restrict__ O, const float *__restrict__ tx,
                       const float *__restrict__ ty, float *__restrict__ dO, double *__restrict__ part,
                       float wy) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double lx = 0.0, ly = 0.0;
    if (g < groups) {
        const int64_t rows = groups * group;
        double ym = 0.0;
        for (int q = 0; q < group; ++q) {
            const int64_t r = g * group + q;
            const double ex = (double)O[2 * r] - (double)tx[r];
            lx += ex * ex;
            dO[2 * r] = (float)(2.0 * ex / (double)rows);
            ym += O[2 * r + 1];
        }
        ym /= group;
        const double ey = ym - (double)ty[g];
        ly = ey * ey;
        const float dy = (float)(wy * 2.0 * ey / ((double)groups * group));
        for (int q = 0; q < group; ++q) dO[2 * (g * group + q) + 1] = dy;
    }
    __shared__ double sx[256], sy[256];
    sx[threadIdx.x] = lx;
    sy[threadIdx.x] = ly;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            sx[threadIdx.x] += sx[threadIdx.x + st];
            sy[threadIdx.x] += sy[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sx[0];
        part[2 * blockIdx.x + 1] = sy[0];
    }
}
__global__ void k_adam(int64_t cnt, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= cnt) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    w[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
// gather rows (X, tx) and groups (ty) of a minibatch in a given group order
__global__ void k_gather(int64_t gcount, int group, int nin, const int32_t *__restrict__ order,
                         const float *__restrict__ X, const float *__restrict__ tx, const float *__restrict__ ty,
                         float *__restrict__ Xb, float *__restrict__ txb, float *__restrict__ tyb) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= gcount * group) return;
    const int64_t g = r / group, q = r % group;
    const int64_t src = (int64_t)order[g] * group + q;
    for (int f = 0; f < nin; ++f) Xb[r * nin + f] = X[src * nin + f];
    txb[r] = tx[src];
    if (q == 0) tyb[g] = ty[order[g]];
}

static inline int nb256(int64_t n) { return (int)((n + 255) / 256); }

}  // namespace hprqp

using hprqp::DevVec;

namespace hprqp {

static void mlp_layout(hprqp_mlp *M) {
    const int w[5] = {M->nin, M->hid, M->hid, M->hid, M->nout};
    int64_t o = 0;
    for (int l = 0; l < 4; ++l) {
        M->in_w[l] = w[l];
        M->out_w[l] = w[l + 1];
        M->off_W[l] = o;
        o += (int64_t)w[l] * w[l + 1];
        M->off_b[l] = o;
        o += w[l + 1];
    }
    M->np = o;
}

static void mlp_reserve(hprqp_mlp *M, int64_t rows) {
    if (rows <= M->cap) return;
    const int w[5] = {M->nin, M->hid, M->hid, M->hid, M->nout};
    for (int l = 0; l < 5; ++l) {
        M->act[l].alloc(rows * w[l]);
        M->dact[l].alloc(rows * w[l]);
    }
    M->part.alloc(2 * ((rows + 255) / 256 + 1));
    M->cap = rows;
}

// forward on rows already in act[0] (or in X when given)
static void mlp_forward(hipStream_t s, hprqp_mlp *M, int64_t rows, const float *X = nullptr) {
    for (int l = 0; l < 4; ++l) {
        const float *in = (l == 0 && X) ? X : M->act[l].p;
        sgemm<false, false>(s, (int)rows, M->out_w[l], M->in_w[l], in, M->in_w[l], M->p.p + M->off_W[l],
                            M->out_w[l], M->act[l + 1].p, M->out_w[l], 0.f);
        k_bias_act<<<nb256(rows * M->out_w[l]), 256, 0, s>>>(rows, M->out_w[l], M->act[l + 1].p,
                                                             M->p.p + M->off_b[l], l < 3);
        HPRQP_HIP(hipGetLastError());
    }
}

void mlp_forward_dev(hipStream_t s, hprqp_mlp *M, const float *X, int64_t rows, float *out) {
    const int64_t chunk = 1 << 20;
    mlp_reserve(M, std::min(rows, chunk));
    for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
        const int64_t r = std::min(chunk, rows - r0);
        mlp_forward(s, M, r, X + r0 * M->nin);
        HPRQP_HIP(hipMemcpyAsync(out + 2 * r0, M->act[4].p, sizeof(float) * 2 * r, hipMemcpyDeviceToDevice, s));
    }
}

// loss + full gradient into M->g for rows = groups * group in act[0]
static double mlp_backward(hipStream_t s, hprqp_mlp *M, int64_t groups, int group, const float *tx, const float *ty,
                           double *loss_parts, float wy) {
    const int64_t rows = groups * group;
    mlp_forward(s, M, rows);
    const int nbk = nb256(groups);
    k_loss<<<nbk, 256, 0, s>>>(groups, group, M->act[4].p, tx, ty, M->dact[4].p, M->part.p, wy);
    HPRQP_HIP(hipGetLastError());
    for (int l = 3; l >= 0; --l) {
        // dW_l = act_l^T dact_{l+1}; db_l = colsum(dact_{l+1})
        sgemm<true, false>(s, M->in_w[l], M->out_w[l], (int)rows, M->act[l].p, M->in_w[l], M->dact[l + 1].p,
                           M->out_w[l], M->g.p + M->off_W[l], M->out_w[l], 0.f);
        k_colsum<<<M->out_w[l], 256, 0, s>>>(rows, M->out_w[l], M->dact[l + 1].p, M->g.p + M->off_b[l]);
        HPRQP_HIP(hipGetLastError());
        if (l > 0) {
            // dact_l = dact_{l+1} W_l^T, masked by ReLU'(act_l)
            sgemm<false, true>(s, (int)rows, M->in_w[l], M->out_w[l], M->dact[l + 1].p, M->out_w[l],
                               M->p.p + M->off_W[l], M->out_w[l], M->dact[l].p, M->in_w[l], 0.f);
            k_relu_back<<<nb256(rows * M->in_w[l]), 256, 0, s>>>(rows * M->in_w[l], M->dact[l].p, M->act[l].p);
            HPRQP_HIP(hipGetLastError());
        }
    }
    std::vector<double> h(2 * nbk);
    HPRQP_HIP(hipMemcpyAsync(h.data(), M->part.p, sizeof(double) * 2 * nbk, hipMemcpyDeviceToHost, s));
    HPRQP_HIP(hipStreamSynchronize(s));
    double lx = 0, ly = 0;
    for (int b = 0; b < nbk; ++b) {
        lx += h[2 * b];
        ly += h[2 * b + 1];
    }
    if (loss_parts) {
        loss_parts[0] = lx / (double)rows;
        loss_parts[1] = ly / (double)groups;
    }
    return lx / (double)rows + wy * ly / (double)groups;
}

}  // namespace hprqp

#define MLP_GUARD(ctx, ...)                               \
    do {                                                  \
        if (!(ctx)) return HPRQP_E_ARG;                   \
        (void)hipSetDevice((ctx)->device);                \
        try {                                             \
            __VA_ARGS__;                                  \
            return HPRQP_OK;                              \
        } catch (const hprqp::Error &e) {                 \
            (ctx)->err = e.what();                        \
            return e.code;                                \
        } catch (const std::exception &e) {               \
            (ctx)->err = e.what();                        \
            return HPRQP_E_STATE;                         \
        }                                                 \
    } while (0)

extern "C" {

int hprqp_mlp_create(hprqp_ctx *ctx, int32_t n_in, int32_t hidden, uint64_t seed, hprqp_mlp **out) {
    if (!out || n_in <= 0 || hidden <= 0) return HPRQP_E_ARG;
    *out = nullptr;
    MLP_GUARD(ctx, {
        hprqp_mlp *M = new hprqp_mlp();
        M->nin = n_in;
        M->hid = hidden;
        M->nout = 2;
        hprqp::mlp_layout(M);
        // He-uniform weights, zero biases, from a seeded host generator
        std::vector<float> h(M->np, 0.f);
        std::mt19937_64 rng(seed);
        for (int l = 0; l < 4; ++l) {
            const float a = std::sqrt(6.0f / (float)M->in_w[l]) * (l == 3 ? 0.1f : 1.0f);
            std::uniform_real_distribution<float> U(-a, a);
            for (int64_t e = 0; e < (int64_t)M->in_w[l] * M->out_w[l]; ++e) h[M->off_W[l] + e] = U(rng);
        }
        M->p.alloc(M->np);
        M->g.alloc(M->np);
        M->am.alloc(M->np);
        M->av.alloc(M->np);
        HPRQP_HIP(hipMemcpy(M->p.p, h.data(), sizeof(float) * M->np, hipMemcpyHostToDevice));
        HPRQP_HIP(hipMemset(M->am.p, 0, sizeof(float) * M->np));
        HPRQP_HIP(hipMemset(M->av.p, 0, sizeof(float) * M->np));
        *out = M;
    });
}

int hprqp_mlp_destroy(hprqp_ctx *ctx, hprqp_mlp *mlp) {
    MLP_GUARD(ctx, {
        HPRQP_HIP(hipStreamSynchronize(ctx->stream));
        delete mlp;
    });
}

int64_t hprqp_mlp_param_count(const hprqp_mlp *mlp) { return mlp ? mlp->np : 0; }

int hprqp_mlp_get_params(hprqp_ctx *ctx, const hprqp_mlp *mlp, float *host) {
    if (!mlp || !host) return HPRQP_E_ARG;
    MLP_GUARD(ctx, HPRQP_HIP(hipMemcpy(host, mlp->p.p, sizeof(float) * mlp->np, hipMemcpyDeviceToHost)));
}

int hprqp_mlp_set_params(hprqp_ctx *ctx, hprqp_mlp *mlp, const float *host) {
    if (!mlp || !host) return HPRQP_E_ARG;
    MLP_GUARD(ctx, {
        HPRQP_HIP(hipMemcpy(mlp->p.p, host, sizeof(float) * mlp->np, hipMemcpyHostToDevice));
        HPRQP_HIP(hipMemset(mlp->am.p, 0, sizeof(float) * mlp->np));
        HPRQP_HIP(hipMemset(mlp->av.p, 0, sizeof(float) * mlp->np));
        mlp->adam_t = 0;
    });
}

int hprqp_mlp_forward(hprqp_ctx *ctx, hprqp_mlp *mlp, int64_t rows, const float *X, float *out) {
    if (!mlp || !X || !out || rows < 0) return HPRQP_E_ARG;
    MLP_GUARD(ctx, {
        if (rows == 0) return HPRQP_OK;
        hipStream_t s = ctx->stream;
        const int64_t chunk = 1 << 20;
        hprqp::mlp_reserve(mlp, std::min(rows, chunk));
        for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
            const int64_t r = std::min(chunk, rows - r0);
            HPRQP_HIP(hipMemcpyAsync(mlp->act[0].p, X + r0 * mlp->nin, sizeof(float) * r * mlp->nin,
                                     hipMemcpyHostToDevice, s));
            hprqp::mlp_forward(s, mlp, r);
            HPRQP_HIP(hipMemcpyAsync(out + 2 * r0, mlp->act[4].p, sizeof(float) * 2 * r, hipMemcpyDeviceToHost, s));
            HPRQP_HIP(hipStreamSynchronize(s));
        }
    });
}

int hprqp_mlp_loss_grad(hprqp_ctx *ctx, hprqp_mlp *mlp, int64_t groups, int32_t group, const float *X,
                        const float *tx, const float *ty, float wy, double *loss, float *grad) {
    if (!mlp || !X || !tx || !ty || groups <= 0 || group <= 0) return HPRQP_E_ARG;
    MLP_GUARD(ctx, {
        hipStream_t s = ctx->stream;
        const int64_t rows = groups * group;
        hprqp::mlp_reserve(mlp, rows);
        DevVec<float> dtx(rows), dty(groups);
        HPRQP_HIP(hipMemcpyAsync(mlp->act[0].p, X, sizeof(float) * rows * mlp->nin, hipMemcpyHostToDevice, s));
        HPRQP_HIP(hipMemcpyAsync(dtx.p, tx, sizeof(float) * rows, hipMemcpyHostToDevice, s));
        HPRQP_HIP(hipMemcpyAsync(dty.p, ty, sizeof(float) * groups, hipMemcpyHostToDevice, s));
        const double L = hprqp::mlp_backward(s, mlp, groups, group, dtx.p, dty.p, nullptr, wy);
        if (loss) *loss = L;
        if (grad) HPRQP_HIP(hipMemcpy(grad, mlp->g.p, sizeof(float) * mlp->np, hipMemcpyDeviceToHost));
    });
}

int hprqp_mlp_train(hprqp_ctx *ctx, hprqp_mlp *mlp, int64_t groups, int32_t group, const float *X,
                    const float *tx, const float *ty, float wy, int32_t batch_groups, int32_t epochs, float lr,
                    uint64_t seed, double *loss_history) {
    if (!mlp || !X || !tx || !ty || groups <= 0 || group <= 0 || batch_groups <= 0 || epochs < 0)
        return HPRQP_E_ARG;
    MLP_GUARD(ctx, {
        hipStream_t s = ctx->stream;
        const int64_t rows = groups * group, nin = mlp->nin;
        DevVec<float> dX(rows * nin), dtx(rows), dty(groups);
        HPRQP_HIP(hipMemcpyAsync(dX.p, X, sizeof(float) * rows * nin, hipMemcpyHostToDevice, s));
        HPRQP_HIP(hipMemcpyAsync(dtx.p, tx, sizeof(float) * rows, hipMemcpyHostToDevice, s));
        HPRQP_HIP(hipMemcpyAsync(dty.p, ty, sizeof(float) * groups, hipMemcpyHostToDevice, s));
        const int64_t bg = std::min<int64_t>(batch_groups, groups);
        hprqp::mlp_reserve(mlp, bg * group);
        DevVec<float> btx(bg * group), bty(bg);
        DevVec<int32_t> dorder(groups);
        std::vector<int32_t> order(groups);
        for (int64_t g = 0; g < groups; ++g) order[g] = (int32_t)g;
        std::mt19937_64 rng(seed);
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        for (int e = 0; e < epochs; ++e) {
            std::shuffle(order.begin(), order.end(), rng);
            HPRQP_HIP(hipMemcpyAsync(dorder.p, order.data(), sizeof(int32_t) * groups, hipMemcpyHostToDevice, s));
            double ep_loss = 0.0;
            int64_t nb = 0;
            for (int64_t g0 = 0; g0 + bg <= groups; g0 += bg) {
                hprqp::k_gather<<<hprqp::nb256(bg * group), 256, 0, s>>>(bg, group, (int)nin, dorder.p + g0, dX.p,
                                                                       dtx.p, dty.p, mlp->act[0].p, btx.p, bty.p);
                HPRQP_HIP(hipGetLastError());
                ep_loss += hprqp::mlp_backward(s, mlp, bg, group, btx.p, bty.p, nullptr, wy);
                ++nb;
                ++mlp->adam_t;
                const float c1 = 1.f - std::pow(b1, (float)mlp->adam_t), c2 = 1.f - std::pow(b2, (float)mlp->adam_t);
                hprqp::k_adam<<<hprqp::nb256(mlp->np), 256, 0, s>>>(mlp->np, mlp->p.p, mlp->g.p, mlp->am.p,
                                                                  mlp->av.p, lr, b1, b2, eps, c1, c2);
                HPRQP_HIP(hipGetLastError());
            }
            if (loss_history) loss_history[e] = nb ? ep_loss / nb : 0.0;
        }
        HPRQP_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
