// hprqp_qap.hip -- QAP relaxation operator Q(X) = 2(A X B - S X - X T)
// (src/Q_operators/QAP_operator.jl:52-76) on dense column-major nq x nq
// matrices.  This file holds the LDS-tiled FP64/FP32 products; each call is
// four launches: TMP = A X; QX = 2 TMP B; QX -= 2 S X; QX -= 2 X T.
#include "hprqp_internal.hpp"


namespace hprqp {

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

void qap_apply_fp64(hipStream_t s, int32_t nq, const double *A, const double *B, const double *S,
                    const double *T, const double *X, double *tmp, double *QX) {
    gemm<double>(s, nq, 1.0, A, X, 0.0, tmp);
    gemm<double>(s, nq, 2.0, tmp, B, 0.0, QX);
    gemm<double>(s, nq, -2.0, S, X, 1.0, QX);
    gemm<double>(s, nq, -2.0, X, T, 1.0, QX);
}

void qap_apply_fp32(hipStream_t s, int32_t nq, const double *A, const double *B, const double *S,
                    const double *T, const float *X, float *tmp, float *QX) {
    gemm<float>(s, nq, 1.0f, A, X, 0.0f, tmp);
    gemm<float>(s, nq, 2.0f, tmp, B, 0.0f, QX);
    gemm<float>(s, nq, -2.0f, S, X, 1.0f, QX);
    gemm<float>(s, nq, -2.0f, X, T, 1.0f, QX);
}

}  // namespace hprqp
