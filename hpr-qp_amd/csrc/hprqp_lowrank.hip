// hprqp_lowrank.hip -- the "diagonal plus low rank" Q operator,
// Qx = diag .* x + U (U^T x), U n x k column-major: the custom operator the
// reference documents for AbstractQOperator (docs/src/guide/
// q_operators_overview.md:121-139, interface src/Q_operators/
// Q_operator_interface.jl:5-92).  Two launches: k column dots into t (one
// block per column, two-pass free because each block owns its column), then
// one thread per row for d_i x_i + U_i. t (columns of U read coalesced).
#include "hprqp_internal.hpp"
#include "hprqp_kernels.hpp"

namespace hprqp {

template <class MT, class XT>
__global__ __launch_bounds__(kBlock) void k_lr_colsdot(int32_t n, const MT *__restrict__ U,
                                                       const XT *__restrict__ x, double *__restrict__ t) {
    const int j = blockIdx.x;
    const MT *col = U + (int64_t)j * n;
    double acc = 0.0;
    for (int32_t i = threadIdx.x; i < n; i += kBlock) acc += (double)col[i] * (double)x[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    __shared__ double sh[kBlock / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) r += sh[w];
        t[j] = r;
    }
}

template <class MT, class XT, class YT>
__global__ __launch_bounds__(kBlock) void k_lr_rows(int32_t n, int32_t k, const MT *__restrict__ U,
                                                    const MT *__restrict__ d, const XT *__restrict__ x,
                                                    const double *__restrict__ t, YT *__restrict__ Qx) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double acc = (double)d[i] * (double)x[i];
    for (int32_t j = 0; j < k; ++j) acc += (double)U[i + (int64_t)j * n] * t[j];
    Qx[i] = (YT)acc;
}

template <class MT, class XT, class YT>
void lowrank_apply(hipStream_t s, int32_t n, int32_t k, const MT *U, const MT *d, const XT *x, double *t, YT *Qx) {
    if (k > 0) {
        k_lr_colsdot<MT, XT><<<k, kBlock, 0, s>>>(n, U, x, t);
        HPRQP_HIP(hipGetLastError());
    }
    k_lr_rows<MT, XT, YT><<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(n, k, U, d, x, t, Qx);
    HPRQP_HIP(hipGetLastError());
}

template void lowrank_apply<double, double, double>(hipStream_t, int32_t, int32_t, const double *, const double *,
                                                    const double *, double *, double *);
template void lowrank_apply<double, float, double>(hipStream_t, int32_t, int32_t, const double *, const double *,
                                                   const float *, double *, double *);
template void lowrank_apply<float, double, double>(hipStream_t, int32_t, int32_t, const float *, const float *,
                                                   const double *, double *, double *);
template void lowrank_apply<float, float, float>(hipStream_t, int32_t, int32_t, const float *, const float *,
                                                 const float *, double *, float *);

}  // namespace hprqp
