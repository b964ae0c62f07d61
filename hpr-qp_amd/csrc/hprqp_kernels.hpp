// hprqp_kernels.hpp -- gfx950 device kernels of the HPR-QP iteration.
//
// Design (CDNA4, wave64):
//  * CSR SpMV: a group of L lanes (L = 1..64, a power of two, chosen per
//    matrix by the autotuner) owns one row; lanes stride the row's nonzeros
//    and combine with xor-shuffles inside the group.  The transpose product
//    uses a stored transpose, never atomics, so every kernel is race-free and
//    deterministic.
//  * Every HPR half-step is ONE kernel: the SpMV's epilogue (run by the
//    group's lane 0) applies the projection / prox / Halpern update for that
//    row, so the product vector never makes a round trip through HBM unless a
//    later phase needs it.  Equivalent of the reference's "customized" fused
//    kernels (src/kernels.jl:422-991) with a wave-cooperative row product.
//  * Block index remap: the hardware deals workgroups round-robin over the
//    8 XCDs; xcd_block() gives each XCD a contiguous range of rows so the
//    gathered vector segments stay in that XCD's L2.
//  * Reductions are two-pass (per-block partials, then one block) with a
//    fixed grid, so results are bitwise reproducible run to run.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hprqp {

constexpr int kBlock = 256;
constexpr int kMaxReduceBlocks = 1024;

__device__ __forceinline__ int xcd_block(int b, int nblocks) {
    constexpr int X = 8;
    const int per = nblocks / X, rem = nblocks % X;
    const int xcd = b % X, k = b / X;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

template <class T>
__device__ __forceinline__ T dmin(T a, T b) { return a < b ? a : b; }
template <class T>
__device__ __forceinline__ T dmax(T a, T b) { return a > b ? a : b; }
// Julia min(max(z, lo), hi)
template <class T>
__device__ __forceinline__ T proj(T z, T lo, T hi) { return dmin(dmax(z, lo), hi); }

// ---------------------------------------------------------------------
// CSR row product with fused epilogue
// ---------------------------------------------------------------------
template <int L, class MT, class XT, class AT, class Epi>
__global__ __launch_bounds__(kBlock) void k_csr(int32_t nrows,
                                                const int32_t *__restrict__ rp,
                                                const int32_t *__restrict__ ci,
                                                const MT *__restrict__ v,
                                                const XT *__restrict__ x, Epi epi) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int64_t t = (int64_t)b * kBlock + threadIdx.x;
    const int32_t row = (int32_t)(t / L);
    const int lane = threadIdx.x & (L - 1);
    if (row >= nrows) return;  // whole L-group leaves together
    const int32_t s = rp[row], e = rp[row + 1];
    AT acc0 = 0, acc1 = 0;
    int32_t k = s + lane;
    for (; k + L < e; k += 2 * L) {
        acc0 += (AT)v[k] * (AT)x[ci[k]];
        acc1 += (AT)v[k + L] * (AT)x[ci[k + L]];
    }
    if (k < e) acc0 += (AT)v[k] * (AT)x[ci[k]];
    AT acc = acc0 + acc1;
    if constexpr (L > 1) {
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, L);
    }
    if (lane == 0) epi(row, acc);
}

// long rows (hundreds of nonzeros and up): one block per row, every thread
// strides the row with four independent accumulators, so a row's loads are
// in flight together; wave shuffles and one LDS exchange finish the sum in a
// fixed order
template <class MT, class XT, class AT, class Epi>
__global__ __launch_bounds__(kBlock) void k_csr_rowblock(int32_t nrows, const int32_t *__restrict__ rp,
                                                         const int32_t *__restrict__ ci,
                                                         const MT *__restrict__ v, const XT *__restrict__ x,
                                                         Epi epi) {
    __shared__ AT part[kBlock / 64];
    const int32_t row = blockIdx.x;
    const int32_t s = rp[row], e = rp[row + 1];
    AT a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int32_t k = s + threadIdx.x;
    for (; k + 3 * kBlock < e; k += 4 * kBlock) {
        a0 += (AT)v[k] * (AT)x[ci[k]];
        a1 += (AT)v[k + kBlock] * (AT)x[ci[k + kBlock]];
        a2 += (AT)v[k + 2 * kBlock] * (AT)x[ci[k + 2 * kBlock]];
        a3 += (AT)v[k + 3 * kBlock] * (AT)x[ci[k + 3 * kBlock]];
    }
    for (; k < e; k += kBlock) a0 += (AT)v[k] * (AT)x[ci[k]];
    AT acc = (a0 + a1) + (a2 + a3);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) epi(row, (part[0] + part[1]) + (part[2] + part[3]));
}

// Column-chunked row product for long rows over many columns (the LASSO data
// matrix: 3,308 rows of ~1,400 nonzeros over 150,360 columns).  There, one
// block per row gathers x[col] at random from L2, one request per nonzero.
// Here block (g, c) stages x[c W .. c W + W) in LDS and sums the nonzeros of
// rows 128 g .. 128 g + 127 that fall in that chunk; a 16-lane group owns a
// row, and the first 64 nonzeros of four rows are loaded together.  The
// per-chunk sums go to part[c * nrow + r], and k_chunk_sum adds them in a
// fixed order, so the result does not depend on scheduling.  seg[c * nrow +
// r] is the first nonzero of row r with column >= c W (c = 0 .. nchunk).
constexpr int kChunkW = 4096;       // columns per chunk
constexpr int kChunkRows = 128;     // rows per block (32 per wave)
constexpr int kLanesChunked = 1 << 20;  // row-product shape code of this form
constexpr int kLanesDiagonal = -2000;   // shape code of Q's diagonal form (k_dia)

static __global__ __launch_bounds__(kBlock) void k_chunk_seg(int32_t nrow, int32_t nchunk, const int32_t *__restrict__ rp,
                                                             const int32_t *__restrict__ ci, int32_t *__restrict__ seg) {
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= (int64_t)(nchunk + 1) * nrow) return;
    const int32_t c = (int32_t)(idx / nrow), r = (int32_t)(idx % nrow);
    // first entry with column >= c W; at c = nchunk (c W >= ncol) that is
    // the row's end
    int32_t lo = rp[r], hi = rp[r + 1];
    const int64_t col = (int64_t)c * kChunkW;
    while (lo < hi) {
        const int32_t mid = lo + (hi - lo) / 2;
        if ((int64_t)ci[mid] < col)
            lo = mid + 1;
        else
            hi = mid;
    }
    seg[idx] = lo;
}

template <class MT, class XT, class AT>
__global__ __launch_bounds__(kBlock) void k_csr_chunk(int32_t nrow, int32_t ncol, const int32_t *__restrict__ seg,
                                                      const int32_t *__restrict__ ci, const MT *__restrict__ v,
                                                      const XT *__restrict__ x, double *__restrict__ part) {
    __shared__ XT xs[kChunkW];
    // 1-D grid of (row group, chunk) pairs, chunk-major over each XCD's
    // contiguous share of the grid: an XCD's blocks take a few whole chunks,
    // so the chunk of x each block stages is read from HBM about once and
    // from that XCD's L2 by the chunk's other row groups (row-group-major
    // dealing spread every chunk over all eight L2s: 1.6x the algorithmic
    // bytes fetched)
    const int32_t ng = (nrow + kChunkRows - 1) / kChunkRows;
    const int32_t vb = xcd_block(blockIdx.x, gridDim.x), c = vb / ng, grp = vb % ng;
    const int32_t c0 = c * kChunkW, cw = min(kChunkW, ncol - c0);
    for (int i = threadIdx.x; i < cw; i += kBlock) xs[i] = x[c0 + i];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, sub = lane >> 4, l16 = lane & 15;
    for (int32_t rg = grp * kChunkRows + 16 * w; rg < grp * kChunkRows + kChunkRows; rg += 64) {
        const int32_t rbase = rg + sub;  // rows rbase + 4 p, p < 4
        int32_t sg[4], eg[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int32_t r = rbase + 4 * p;
            sg[p] = r < nrow ? seg[(int64_t)c * nrow + r] : 0;
            eg[p] = r < nrow ? seg[(int64_t)(c + 1) * nrow + r] : 0;
        }
        AT acc[4];
        {
            // the first 64 nonzeros of the four rows, all loads issued
            // together; lanes past a segment's end load entry 0 and add nothing
            MT vv[4][4];
            int32_t cc[4][4];
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int32_t q = sg[p] + l16 + 16 * u, at = q < eg[p] ? q : 0;
                    vv[p][u] = v[at];
                    cc[p][u] = ci[at];
                }
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                acc[p] = (AT)0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const bool ok = sg[p] + l16 + 16 * u < eg[p];
                    const AT pr = (AT)vv[p][u] * (AT)xs[ok ? cc[p][u] - c0 : 0];
                    acc[p] += ok ? pr : (AT)0;
                }
            }
        }
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            for (int32_t q = sg[p] + 64 + l16; q < eg[p]; q += 16) acc[p] += (AT)v[q] * (AT)xs[ci[q] - c0];
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) acc[p] += __shfl_xor(acc[p], off);
            const int32_t r = rbase + 4 * p;
            if (l16 == 0 && r < nrow) part[(int64_t)c * nrow + r] = (double)acc[p];
        }
    }
}

// 16 lanes per row: lane l adds chunks l, l + 16, ..., then the 16 sums are
// combined by a fixed shuffle tree
template <class AT, class Epi>
__global__ __launch_bounds__(kBlock) void k_chunk_sum(int32_t nrow, int32_t nchunk, const double *__restrict__ part,
                                                      Epi epi) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int32_t r = (int32_t)(t >> 4), l16 = (int32_t)(t & 15);
    double s = 0.0;
    if (r < nrow)
        for (int32_t c = l16; c < nchunk; c += 16) s += part[(int64_t)c * nrow + r];
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (l16 == 0 && r < nrow) epi(r, (AT)s);
}

// CSR-stream row product for short rows: a block owns RB consecutive rows;
// its nonzeros are read once, fully coalesced, and their products staged in
// LDS; then thread t sums row r0+t in order and runs the epilogue, so every
// vector access of the epilogue is coalesced as well.  A block whose rows
// hold more than kStreamCap nonzeros falls back to one thread per row.
constexpr int kStreamCap = 2048;
constexpr int kStreamU = 4;  // nonzeros per thread per pass of the product phase

template <class MT, class XT, class AT, class Epi>
__global__ __launch_bounds__(kBlock) void k_csr_stream(int32_t nrows, int rb, const int32_t *__restrict__ rp,
                                                       const int32_t *__restrict__ ci,
                                                       const MT *__restrict__ v, const XT *__restrict__ x,
                                                       Epi epi) {
    __shared__ AT prod[kStreamCap];
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int32_t r0 = b * rb;
    if (r0 >= nrows) return;
    const int32_t r1 = min(r0 + rb, nrows);
    const int32_t base = rp[r0];
    const int32_t cnt = rp[r1] - base;
    const int t = threadIdx.x;
    const int32_t row = r0 + t;
    if (cnt <= kStreamCap) {
        // kStreamU nonzeros per thread per pass, all loads issued before the
        // first use; lanes past cnt load the block's first nonzero (cnt >= 1
        // inside the loop, so it exists) and discard it, so no load waits
        // behind a branch
        for (int32_t k0 = 0; k0 < cnt; k0 += kStreamU * kBlock) {
            MT vv[kStreamU];
            int32_t cc[kStreamU];
#pragma unroll
            for (int j = 0; j < kStreamU; ++j) {
                const int32_t k = k0 + j * kBlock + t;
                const int32_t at = base + (k < cnt ? k : 0);
                vv[j] = v[at];
                cc[j] = ci[at];
            }
            AT pr[kStreamU];
#pragma unroll
            for (int j = 0; j < kStreamU; ++j) pr[j] = (AT)vv[j] * (AT)x[cc[j]];
#pragma unroll
            for (int j = 0; j < kStreamU; ++j) {
                const int32_t k = k0 + j * kBlock + t;
                if (k < cnt) prod[k] = pr[j];
            }
        }
        __syncthreads();
        if (row < r1) {
            const int32_t s = rp[row] - base, e = rp[row + 1] - base;
            AT acc = 0;
            for (int32_t k = s; k < e; ++k) acc += prod[k];
            epi(row, acc);
        }
    } else if (row < r1) {
        AT acc = 0;
        for (int32_t k = rp[row]; k < rp[row + 1]; ++k) acc += (AT)v[k] * (AT)x[ci[k]];
        epi(row, acc);
    }
}

// The CSR-stream product of a matrix in value-table form: entry k of row r
// is dict[code[k]] * rsc[r] * csc[col] -- the unscaled value (one byte per
// nonzero instead of eight) with the row and column scaling factors
// (src/utils.jl:350-804 scales by diagonal factors, so the scaled matrix is
// exactly this product).  The table (at most 256 values) is staged in LDS;
// csc[col] is gathered beside x[col], and rsc[r] multiplies the row's sum.
template <class XT, class AT, class Epi>
__global__ __launch_bounds__(kBlock) void k_csr_stream_dict(int32_t nrows, int rb, const int32_t *__restrict__ rp,
                                                            const int32_t *__restrict__ ci,
                                                            const uint8_t *__restrict__ code,
                                                            const double *__restrict__ dict, int ndict,
                                                            const double *__restrict__ rsc,
                                                            const double *__restrict__ csc, const XT *__restrict__ x,
                                                            Epi epi) {
    __shared__ AT prod[kStreamCap];
    __shared__ double dl[256];
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int32_t r0 = b * rb;
    if (r0 >= nrows) return;
    if ((int)threadIdx.x < ndict) dl[threadIdx.x] = dict[threadIdx.x];
    const int32_t r1 = min(r0 + rb, nrows);
    const int32_t base = rp[r0];
    const int32_t cnt = rp[r1] - base;
    const int t = threadIdx.x;
    const int32_t row = r0 + t;
    __syncthreads();
    if (cnt <= kStreamCap) {
        for (int32_t k0 = 0; k0 < cnt; k0 += kStreamU * kBlock) {
            uint8_t cd[kStreamU];
            int32_t cc[kStreamU];
#pragma unroll
            for (int j = 0; j < kStreamU; ++j) {
                const int32_t k = k0 + j * kBlock + t;
                const int32_t at = base + (k < cnt ? k : 0);
                cd[j] = code[at];
                cc[j] = ci[at];
            }
            AT pr[kStreamU];
#pragma unroll
            for (int j = 0; j < kStreamU; ++j) pr[j] = (AT)(dl[cd[j]] * csc[cc[j]]) * (AT)x[cc[j]];
#pragma unroll
            for (int j = 0; j < kStreamU; ++j) {
                const int32_t k = k0 + j * kBlock + t;
                if (k < cnt) prod[k] = pr[j];
            }
        }
        __syncthreads();
        if (row < r1) {
            const int32_t s = rp[row] - base, e = rp[row + 1] - base;
            AT acc = 0;
            for (int32_t k = s; k < e; ++k) acc += prod[k];
            epi(row, acc * (AT)rsc[row]);
        }
    } else if (row < r1) {
        AT acc = 0;
        for (int32_t k = rp[row]; k < rp[row + 1]; ++k) acc += (AT)(dl[code[k]] * csc[ci[k]]) * (AT)x[ci[k]];
        epi(row, acc * (AT)rsc[row]);
    }
}

// Diagonal form: a matrix whose nonzeros lie on at most kDiaMax diagonals --
// entry (i, j) on diagonal d when j = base(i) + off[d], base(i) =
// floor(i * ncol / nrow) (i itself for a square matrix; for a rectangular one
// the diagonals follow its slope, as an edge-cell operator's rows do on a
// grid) -- keeps each diagonal as an array indexed by row, zero where the
// pattern has no entry.  There are no row pointers or column indices to read,
// and every load of a wave -- the diagonal values and x[base(i) + off] -- is
// contiguous (the neighbouring diagonals' x segments are L2 hits).  Row i adds
// its diagonals in ascending column order (off ascending), each product
// rounded before the add, as the CSR-stream kernel does, so the sum is
// bitwise that kernel's: the padding adds exact zeros.
constexpr int kDiaMax = 8;
struct DiaOff {
    int32_t o[kDiaMax];
};

__device__ __host__ __forceinline__ int32_t dia_base(int32_t i, int32_t nrow, int32_t ncol) {
    return nrow == ncol ? i : (int32_t)(((int64_t)i * ncol) / nrow);
}

// PAIR (an instance of its own, launched only for matrices with pairs):
// bit d of pairs set when diagonals d and d + 1 read x[j] and x[j + 1] with
// j even in every row (an even slope ncol / nrow, an even off[d], off[d + 1]
// = off[d] + 1, x aligned to two elements) -- one two-element load serves
// both, halving the cache lines a wave's strided x loads touch
template <class XT>
struct Vec2;
template <>
struct Vec2<double> {
    using T = double2;
};
template <>
struct Vec2<float> {
    using T = float2;
};

template <bool PAIR, class MT, class XT, class AT, class Epi>
__global__ __launch_bounds__(kBlock) void k_dia(int32_t nrow, int32_t ncol, int nd, DiaOff off, unsigned pairs,
                                               const MT *__restrict__ dv, const XT *__restrict__ x, Epi epi) {
#pragma clang fp contract(off)
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int32_t i = b * kBlock + (int32_t)threadIdx.x;
    if (i >= nrow) return;
    const int32_t base = dia_base(i, nrow, ncol);
    AT acc = 0, carry = 0;
#pragma unroll
    for (int d = 0; d < kDiaMax; ++d) {
        if (d < nd) {
            const int64_t j = (int64_t)base + off.o[d];
            const bool in = j >= 0 && j < ncol;
            const AT v = (AT)dv[(int64_t)d * nrow + i];
            AT xv;
            if (PAIR && d > 0 && ((pairs >> (d - 1)) & 1u)) {
                xv = carry;  // the second element of the previous diagonal's load
            } else if (PAIR && ((pairs >> d) & 1u)) {
                if (in && j + 1 < ncol) {
                    const typename Vec2<XT>::T t = *reinterpret_cast<const typename Vec2<XT>::T *>(x + j);
                    xv = (AT)t.x;
                    carry = (AT)t.y;
                } else {
                    xv = in ? (AT)x[j] : (AT)0;
                    carry = (j + 1 >= 0 && j + 1 < ncol) ? (AT)x[j + 1] : (AT)0;
                }
            } else {
                xv = in ? (AT)x[in ? j : 0] : (AT)0;
            }
            const AT pr = v * xv;
            acc += pr;
        }
    }
    epi(i, acc);
}

// elementwise form of the same epilogue (operator Q: product precomputed)
template <class ST, class AT, class Epi>
__global__ __launch_bounds__(kBlock) void k_vec(int32_t n, const ST *__restrict__ src, Epi epi) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) epi((int32_t)i, (AT)src[i]);
}

// plain store epilogue: y = A x
template <class YT>
struct EpiStore {
    YT *y;
    template <class AT>
    __device__ void operator()(int32_t i, AT acc) const { y[i] = (YT)acc; }
};

// y = (A x) * s  (s read from device memory; used by the power method)
template <class YT>
struct EpiStoreScaled {
    YT *y;
    const double *s;
    template <class AT>
    __device__ void operator()(int32_t i, AT acc) const { y[i] = (YT)(acc * (AT)(*s)); }
};

// ---------------------------------------------------------------------
// HPR half-step epilogues.  VT = bulk vector type, CT = compute type.
// Anchors (last_*) are always FP64.
// ---------------------------------------------------------------------

// z/x/w1 step: update_zxw1 (src/kernels.jl:422-537, CPU :2301-2397)
template <class VT, class DT, bool FULL, bool DIAG>
struct EpiZxw1 {
    VT *Qw;  // written when non-null (the product Q w)
    const VT *ATy, *w, *fact1, *fact2;
    const DT *c, *l, *u;
    const double *last_x;
    VT *x, *x_hat, *w_bar, *dx, *x_bar, *z_bar;
    double sigma, f1s, f2s, h1, h2;
    double lS = 0.0, uS = 0.0;  // the bounds when l / u are null (one value for all)
    template <class AT>
    __device__ void operator()(int32_t i, AT qw_) const {
        const double qw = (double)qw_;
        if (Qw) Qw[i] = (VT)qw_;
        const double xi = (double)x[i], wi = (double)w[i];
        const double z_raw = xi + sigma * (-qw + (double)ATy[i] - (double)c[i]);
        const double xb = proj(z_raw, l ? (double)l[i] : lS, u ? (double)u[i] : uS);
        const double xh = 2.0 * xb - xi;
        const double xn = fma(h2, xh, h1 * last_x[i]);
        const double wb = DIAG ? fma((double)fact1[i], wi, (double)fact2[i] * xh)
                               : fma(f1s, wi, f2s * xh);
        if constexpr (FULL) {
            dx[i] = (VT)(xb - xi);
            x_bar[i] = (VT)xb;
            z_bar[i] = (VT)((xb - z_raw) / sigma);
        }
        x[i] = (VT)xn;
        x_hat[i] = (VT)xh;
        w_bar[i] = (VT)wb;
    }
};

// tempv = x_hat + sigma (Qw - Qw_bar)  (src/kernels.jl:674-705)
template <class VT>
struct EpiTempv {
    const VT *x_hat, *Qw;
    VT *tempv, *Qw_bar;  // Qw_bar written when non-null
    double sigma;
    template <class AT>
    __device__ void operator()(int32_t i, AT qwb_) const {
        const double qwb = (double)qwb_;
        if (Qw_bar) Qw_bar[i] = (VT)qwb_;
        tempv[i] = (VT)((double)x_hat[i] + sigma * ((double)Qw[i] - qwb));
    }
};

// dual step: update_y (src/kernels.jl:760-862, CPU :2682-2743)
template <class VT, class DT, bool FULL>
struct EpiY {
    const DT *AL, *AU;
    const double *last_y;
    VT *y, *y_bar, *s, *dy;
    double fact1, fact2, h1, h2;
    template <class AT>
    __device__ void operator()(int32_t i, AT ax_) const {
        const double ax = (double)ax_, yi = (double)y[i];
        const double s_raw = ax - fact1 * yi;
        const double lo = (double)AL[i];
        const double sp = proj(s_raw, lo, AU ? (double)AU[i] : lo);  // AU null: AL = AU
        const double yb = fact2 * (sp - s_raw);
        const double yh = 2.0 * yb - yi;
        const double yn = fma(h2, yh, h1 * last_y[i]);
        if constexpr (FULL) {
            s[i] = (VT)sp;
            dy[i] = (VT)(yb - yi);
        }
        y_bar[i] = (VT)yb;
        y[i] = (VT)yn;
    }
};

// second w step: update_w2 (src/kernels.jl:878-991, CPU :2747-2832)
template <class VT, bool FULL, bool DIAG>
struct EpiW2 {
    const VT *fact;
    const double *last_w, *last_ATy;
    VT *w, *w_bar, *ATy, *dw, *ATdy, *ATy_bar;
    double fs, h1, h2;
    template <class AT>
    __device__ void operator()(int32_t i, AT atyb_) const {
        const double atyb = (double)atyb_;
        const double wi = (double)w[i], wbi = (double)w_bar[i], atyi = (double)ATy[i];
        const double f = DIAG ? (double)fact[i] : fs;
        const double wbn = wbi + f * (atyb - atyi);
        const double wn = h1 * last_w[i] + h2 * (2.0 * wbn - wi);
        const double atyn = h1 * last_ATy[i] + h2 * (2.0 * atyb - atyi);
        w[i] = (VT)wn;
        ATy[i] = (VT)atyn;
        if constexpr (FULL) {
            w_bar[i] = (VT)wbn;
            dw[i] = (VT)(wbn - wi);
            ATdy[i] = (VT)(atyb - atyi);
            ATy_bar[i] = (VT)atyb;
        }
    }
};

// noC z/x/w step: product is A^T y (src/kernels.jl:540-669, CPU :2117-2209)
template <class VT, class DT, bool FULL, bool DIAG>
struct EpiNocZxw {
    VT *ATy;
    const VT *Qw, *fact1, *fact2;
    const DT *c, *l, *u;
    const double *last_x, *last_w;
    VT *x, *x_hat, *w, *w_bar, *dx, *x_bar, *z_bar, *dw;
    double sigma, f1s, f2s, h1, h2;
    double lS = 0.0, uS = 0.0;
    template <class AT>
    __device__ void operator()(int32_t i, AT aty_) const {
        const double aty = (double)aty_;
        ATy[i] = (VT)aty_;
        const double xi = (double)x[i], wi = (double)w[i];
        const double z_raw = xi + sigma * (-(double)Qw[i] + aty - (double)c[i]);
        const double xb = proj(z_raw, l ? (double)l[i] : lS, u ? (double)u[i] : uS);
        const double xh = 2.0 * xb - xi;
        const double xn = fma(h2, xh, h1 * last_x[i]);
        const double wb = DIAG ? fma((double)fact1[i], wi, (double)fact2[i] * xh)
                               : fma(f1s, wi, f2s * xh);
        const double wn = fma(h2, 2.0 * wb - wi, h1 * last_w[i]);
        if constexpr (FULL) {
            dx[i] = (VT)(xb - xi);
            x_bar[i] = (VT)xb;
            z_bar[i] = (VT)((xb - z_raw) / sigma);
            dw[i] = (VT)(wb - wi);
        }
        x[i] = (VT)xn;
        x_hat[i] = (VT)xh;
        w_bar[i] = (VT)wb;
        w[i] = (VT)wn;
    }
};

// noC tempv/Qw Halpern step: product is Q w_bar (src/kernels.jl:710-748)
template <class VT>
struct EpiNocTempv {
    const VT *x_hat;
    const double *last_Qw;
    VT *Qw, *Qw_bar, *tempv;
    double sigma, h1, h2;
    template <class AT>
    __device__ void operator()(int32_t i, AT qwb_) const {
        const double qwb = (double)qwb_, qw = (double)Qw[i];
        Qw_bar[i] = (VT)qwb_;
        Qw[i] = (VT)fma(h2, 2.0 * qwb - qw, h1 * last_Qw[i]);
        tempv[i] = (VT)((double)x_hat[i] + sigma * (qw - qwb));
    }
};

// LASSO step: product is Q w = A^T (A w) (src/kernels.jl:261-376)
template <class VT, class DT, bool FULL>
struct EpiLasso {
    VT *Qw;
    const VT *ATy;
    const DT *c;
    const double *last_x, *last_w;
    VT *x, *x_hat, *w, *w_bar, *dx, *x_bar, *z_bar, *dw;
    double sigma, lambda_sigma, f1s, f2s, h1, h2;
    template <class AT>
    __device__ void operator()(int32_t i, AT qw_) const {
        const double qw = (double)qw_;
        if (Qw) Qw[i] = (VT)qw_;
        const double xi = (double)x[i], wi = (double)w[i];
        const double z_raw = xi + sigma * (-qw + (double)ATy[i] - (double)c[i]);
        const double shrink = dmax(fabs(z_raw) - lambda_sigma, 0.0);
        const double xb = copysign(shrink, z_raw);
        const double xh = 2.0 * xb - xi;
        const double xn = fma(h2, xh, h1 * last_x[i]);
        const double wb = fma(f1s, wi, f2s * xh);
        const double wn = fma(h2, 2.0 * wb - wi, h1 * last_w[i]);
        if constexpr (FULL) {
            dx[i] = (VT)(xb - xi);
            x_bar[i] = (VT)xb;
            z_bar[i] = (VT)((xb - z_raw) / sigma);
            dw[i] = (VT)(wb - wi);
        }
        x[i] = (VT)xn;
        x_hat[i] = (VT)xh;
        w_bar[i] = (VT)wb;
        w[i] = (VT)wn;
    }
};

// LP z/x step: product is A^T y (src/kernels.jl:1243-1351)
template <class VT, class DT, bool FULL>
struct EpiLpZx {
    VT *ATy;
    const DT *c, *l, *u;
    const double *last_x;
    VT *x, *x_hat, *dx, *x_bar, *z_bar;
    double sigma, h1, h2;
    double lS = 0.0, uS = 0.0;
    template <class AT>
    __device__ void operator()(int32_t i, AT aty_) const {
        const double aty = (double)aty_;
        ATy[i] = (VT)aty_;
        const double xi = (double)x[i];
        const double z_raw = xi + sigma * (aty - (double)c[i]);
        const double xb = proj(z_raw, l ? (double)l[i] : lS, u ? (double)u[i] : uS);
        const double xh = 2.0 * xb - xi;
        const double xn = fma(h2, xh, h1 * last_x[i]);
        if constexpr (FULL) {
            dx[i] = (VT)(xb - xi);
            x_bar[i] = (VT)xb;
            z_bar[i] = (VT)((xb - z_raw) / sigma);
        }
        x[i] = (VT)xn;
        x_hat[i] = (VT)xh;
    }
};

// ---------------------------------------------------------------------
// Offset form of the standard half-steps (FP32 bulk).  The bulk iterates
// are FP32 offsets from the FP64 Halpern anchors X0 = last_x, W0 = last_w,
// Y0 = last_y, T0 = last_ATy: dX = x - X0, dW = w - W0, dY = y - Y0,
// dT = ATy - T0, and likewise for x_hat, w_bar, y_bar, A^T y_bar and
// tempv.  Between two restarts the anchors and sigma are fixed, so the
// anchors' products are FP64 vectors formed once per restart:
// G0 = -Q W0 + T0 - c, H0 = A X0 - lambda_A sigma Y0, and D0 = X0 - W0
// (FP32: it enters only as f2 D0, and shrinks with the anchors).  The
// row products act on offsets only (FP32 values, FP32 accumulation), so
// their rounding shrinks with the offsets as the iteration converges.
// Full steps also write the FP64 iterates the residuals, the M-norm, the
// sigma update and restarts read (x_bar, z_bar, dx; y_bar, s, dy; w_bar,
// dw, ATy_bar, ATdy).  Formulas: update_zxw1 / update_y / update_w2
// (src/kernels.jl:422-537, 760-862, 878-991) rewritten for offsets.
// ---------------------------------------------------------------------

// z/x/w1, product Q dW: z = X0 + dX + sigma (G0 - Q dW + dT),
// x_bar = proj(z), dXh = 2 (x_bar - X0) - dX, dX <- h2 dXh,
// dWb = f1 dW + f2 (dXh + D0)
template <bool FULL>
struct EpiOffZxw1 {
    float *qd;  // Q dW, read again by the tempv step
    const float *dW, *dT, *D0;
    const double *G0, *X0, *l, *u;
    float *dX, *dXh, *dWb;
    double *x_bar, *z_bar, *dx;
    double sigma, f1s, f2s, h2;
    double lS = 0.0, uS = 0.0;
    template <class AT>
    __device__ void operator()(int32_t i, AT q_) const {
        const double q = (double)q_, dxi = (double)dX[i], x0 = X0[i];
        qd[i] = (float)q_;
        const double z_raw = x0 + dxi + sigma * (G0[i] - q + (double)dT[i]);
        const double xb = proj(z_raw, l ? l[i] : lS, u ? u[i] : uS);
        const double dxb = xb - x0;
        const double dxh = 2.0 * dxb - dxi;
        if constexpr (FULL) {
            x_bar[i] = xb;
            z_bar[i] = (xb - z_raw) / sigma;
            dx[i] = dxb - dxi;
        }
        dX[i] = (float)(h2 * dxh);
        dXh[i] = (float)dxh;
        dWb[i] = (float)fma(f1s, (double)dW[i], f2s * (dxh + (double)D0[i]));
    }
};

// tempv offset: dTv = dXh + sigma (Q dW - Q dWb), product Q dWb
struct EpiOffTempv {
    const float *dXh, *qd;
    float *dTv;
    double sigma;
    template <class AT>
    __device__ void operator()(int32_t i, AT q_) const {
        dTv[i] = (float)((double)dXh[i] + sigma * ((double)qd[i] - (double)q_));
    }
};

// y step, product A dTv: s = H0 + A dTv - lambda_A sigma dY,
// y_bar = (proj(s) - s) / (lambda_A sigma), dYb = y_bar - Y0,
// dY <- h2 (2 dYb - dY)
template <bool FULL>
struct EpiOffY {
    const double *H0, *Y0, *AL, *AU;
    float *dY, *dYb;
    double *y_bar, *s, *dy;
    double fact1, fact2, h2;
    template <class AT>
    __device__ void operator()(int32_t i, AT a_) const {
        const double dyi = (double)dY[i];
        const double s_raw = H0[i] + (double)a_ - fact1 * dyi;
        const double lo = AL[i];
        const double sp = proj(s_raw, lo, AU ? AU[i] : lo);  // AU null: AL = AU
        const double yb = fact2 * (sp - s_raw);
        const double dyb = yb - Y0[i];
        if constexpr (FULL) {
            y_bar[i] = yb;
            s[i] = sp;
            dy[i] = dyb - dyi;
        }
        dYb[i] = (float)dyb;
        dY[i] = (float)(h2 * (2.0 * dyb - dyi));
    }
};

// w2 step, product A^T dYb = A^T y_bar - T0 (= dTb):
// dWbn = dWb + fs (dTb - dT), dW <- h2 (2 dWbn - dW), dT <- h2 (2 dTb - dT)
template <bool FULL>
struct EpiOffW2 {
    const double *W0, *T0;
    float *dW, *dWb, *dT;
    double *w_bar, *dw, *ATy_bar, *ATdy;
    double fs, h2;
    template <class AT>
    __device__ void operator()(int32_t i, AT a_) const {
        const double dtb = (double)a_, dwi = (double)dW[i], dti = (double)dT[i];
        const double dwbn = (double)dWb[i] + fs * (dtb - dti);
        if constexpr (FULL) {
            dWb[i] = (float)dwbn;
            w_bar[i] = W0[i] + dwbn;
            dw[i] = dwbn - dwi;
            ATy_bar[i] = T0[i] + dtb;
            ATdy[i] = dtb - dti;
        }
        dW[i] = (float)(h2 * (2.0 * dwbn - dwi));
        dT[i] = (float)(h2 * (2.0 * dtb - dti));
    }
};

// ---------------------------------------------------------------------
// multi-value reductions: NS sums and NM maxima, two passes
// ---------------------------------------------------------------------
template <int NS, int NM, class F>
__global__ __launch_bounds__(kBlock) void k_reduce(int64_t n, F f, double *__restrict__ part) {
    constexpr int NT = NS + NM;
    double s[NS > 0 ? NS : 1], mx[NM > 0 ? NM : 1];
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) mx[k] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        f(i, s, mx);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < NS; ++k) s[k] += __shfl_xor(s[k], off);
#pragma unroll
        for (int k = 0; k < NM; ++k) mx[k] = dmax(mx[k], __shfl_xor(mx[k], off));
    }
    __shared__ double sh[kBlock / 64][NT > 0 ? NT : 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NS; ++k) sh[wave][k] = s[k];
#pragma unroll
        for (int k = 0; k < NM; ++k) sh[wave][NS + k] = mx[k];
    }
    __syncthreads();
    if (threadIdx.x < NT) {
        const int k = threadIdx.x;
        double r = sh[0][k];
        for (int w = 1; w < kBlock / 64; ++w) r = k < NS ? r + sh[w][k] : dmax(r, sh[w][k]);
        part[(int64_t)blockIdx.x * NT + k] = r;
    }
}

// final pass: one block; out[k] for k < NS + NM
template <int NS, int NM>
__device__ __forceinline__ void reduce_final_body(int nb, const double *__restrict__ part,
                                                  double *__restrict__ out) {
    constexpr int NT = NS + NM;
    __shared__ double sh[kBlock / 64][NT > 0 ? NT : 1];
    double r[NT > 0 ? NT : 1];
#pragma unroll
    for (int k = 0; k < NT; ++k) r[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) {
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const double v = part[(int64_t)b * NT + k];
            r[k] = k < NS ? r[k] + v : dmax(r[k], v);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const double o = __shfl_xor(r[k], off);
            r[k] = k < NS ? r[k] + o : dmax(r[k], o);
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NT; ++k) sh[wave][k] = r[k];
    }
    __syncthreads();
    if (threadIdx.x < NT) {
        const int k = threadIdx.x;
        double a = sh[0][k];
        for (int w = 1; w < kBlock / 64; ++w) a = k < NS ? a + sh[w][k] : dmax(a, sh[w][k]);
        out[k] = a;
    }
}

template <int NS, int NM>
__global__ __launch_bounds__(kBlock) void k_reduce_final(int nb, const double *__restrict__ part,
                                                         double *__restrict__ out) {
    reduce_final_body<NS, NM>(nb, part, out);
}

// a block's sums of s0 and s1 (every thread's) into dst[0], dst[1], in a
// fixed order; every thread of the block calls it
__device__ __forceinline__ void pow_block_partials(double s0, double s1, double *dst) {
    static_assert(kBlock == 256, "the block reduction sums four waves");
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off);
        s1 += __shfl_xor(s1, off);
    }
    __shared__ double sh[kBlock / 64][2];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[wave][0] = s0;
        sh[wave][1] = s1;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int q = threadIdx.x;
        dst[q] = (sh[0][q] + sh[1][q]) + (sh[2][q] + sh[3][q]);
    }
}

// ---------------------------------------------------------------------
// power method on CSR matrices (src/utils.jl:1335-1461)
// ---------------------------------------------------------------------
// Every product of the power method is one launch, a block per row group
// (L lanes per row: 256 / L rows a group; L = 0: CSR-stream groups of rb
// rows), and where the product feeds the stop test each block writes one
// partial pair of (z'.t, |z'|^2).  The next launch totals them itself
// (hence at most kPowGrid of them): every block reads all of them in the
// same fixed order, so every block holds the same totals, evaluates the
// previous iteration's stop test and scales its own product by the new
// 1/|z|.  The scalar update needs no launch of its own and nothing crosses
// the XCDs inside a launch: block 0 only records what every block computed.
// The loads of the totals, of res[0] and of the block's first row group
// are issued before the first barrier, so their latencies overlap; the
// product is formed unscaled and multiplied by 1/|z| once that is known
// (y = g (M x) rather than M (g x): a rounding apart).
//   Q run, one launch per iteration (kPowQ): z_{t+1} = Q (z_t / |z_t|);
//   A A^T run, two (kPowAT: q = A^T z_t / |z_t| with the stop test of
//   iteration t-1; kPowA: z_{t+1} = A q with its pairs (z_{t+1}.z_t, ...)).
constexpr int kPowGrid = 1024;
static_assert(kPowGrid == 4 * kBlock, "four pairs per thread in the totals");
enum PowMode { kPowQ = 0, kPowAT = 1, kPowA = 2 };

//   prev, pnb: the previous product's pairs (for t = 0 the pair (0, |z_0|^2));
//   part: this launch's pairs; zz: ring of |z_t|^2 (zz[t & 3], written by
//   block 0 of the launch that tests iteration t-1); res: {stopped, lambda,
//   iterations, err} of the last test; scaled: lambda = z'.z / |z| (A A^T,
//   whose pairs are taken against the unscaled z) rather than z'.q (Q).
// A block that reads res[0] as set by block 0 of its own launch returns as
// it would after computing the same test; thread 0 reads res[0] for the
// whole block, so the block returns as one.
struct PowArgs {
    const double *prev;
    int pnb;
    double *part;
    double *zz;
    double *res;
    int64_t t;
    double tol;
    int scaled;
};

// first half: res[0] into *stopped (thread 0), this thread's share of the
// previous pairs summed over its wave into sh (the caller's barrier follows)
__device__ __forceinline__ void pow_issue(const PowArgs &a, double *stopped, double (*sh)[2]) {
    static_assert(kBlock == 256, "the totals sum four waves");
    if (threadIdx.x == 0) *stopped = a.res[0];
    const double2 *p2 = (const double2 *)a.prev;
    double2 u[kPowGrid / kBlock];
#pragma unroll
    for (int k = 0; k < kPowGrid / kBlock; ++k) {
        const int i = threadIdx.x + k * kBlock;
        u[k] = i < a.pnb ? p2[i] : double2{0.0, 0.0};
    }
    double s0 = (u[0].x + u[1].x) + (u[2].x + u[3].x);
    double s1 = (u[0].y + u[1].y) + (u[2].y + u[3].y);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off);
        s1 += __shfl_xor(s1, off);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6][0] = s0;
        sh[threadIdx.x >> 6][1] = s1;
    }
}

// second half, after a barrier: false when the run has stopped (the launch
// stores nothing); else g = 1/|z_t|
__device__ __forceinline__ bool pow_test(const PowArgs &a, double stopped, const double (*sh)[2], double &g) {
    if (stopped != 0.0) return false;
    const double dot = (sh[0][0] + sh[1][0]) + (sh[2][0] + sh[3][0]);
    const double nn = (sh[0][1] + sh[1][1]) + (sh[2][1] + sh[3][1]);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (a.t > 0) {
        // the reference's test for iteration t-1 (src/utils.jl:1357-1376):
        // lambda = q.z', err = |z' - lambda q| / (|z'| + lambda), with
        // |z' - lambda q|^2 = |z'|^2 - lambda^2 (2 - |q|^2)
        const double zz = a.zz[(a.t - 1) & 3], rz = 1.0 / sqrt(zz);
        const double lam = a.scaled ? dot * rz : dot, qn = zz * rz * rz;
        const double e2 = fmax(nn - lam * lam * (2.0 - qn), 0.0);
        const double err = sqrt(e2) / (sqrt(nn) + lam);
        if (lead) {
            a.res[1] = lam;
            a.res[2] = (double)a.t;
            a.res[3] = err;
            if (err < a.tol) a.res[0] = 1.0;
        }
        if (err < a.tol) return false;
    }
    if (lead) a.zz[a.t & 3] = nn;
    g = 1.0 / sqrt(nn);
    return true;
}

// a matrix as the power method's products read it: CSR (rp, ci, v) in
// row groups (L lanes per row: 256 / L rows; L = 0: CSR-stream groups of rb
// rows), or the value-table form (code, dict of ndict entries, row and
// column factors rsc, csc) in CSR-stream groups
struct PowMat {
    int32_t nrows, rb, ngroups;
    const int32_t *rp, *ci;
    const double *v;
    const uint8_t *code;
    const double *dict;
    int ndict;
    const double *rsc, *csc;
};

// y = g (M x) over the block's row groups, with (y.t', |y|^2) pairs in
// kPowQ (t' = x g) and kPowA (t' = t, g = 1) modes
template <int L, int MODE, bool DICT>
__global__ __launch_bounds__(kBlock) void k_pow(PowMat M, const double *__restrict__ x,
                                                const double *__restrict__ t, double *__restrict__ y, PowArgs a) {
    static_assert(!DICT || L == 0, "the value-table form runs in CSR-stream groups");
    constexpr bool TEST = MODE != kPowA;
    __shared__ double sh[kBlock / 64][2];
    __shared__ double stopped;
    __shared__ double prod[L == 0 ? kStreamCap : 1];
    __shared__ double dl[DICT ? 256 : 1];
    const int32_t nrows = M.nrows, rb = M.rb, ngroups = M.ngroups;
    const int32_t *__restrict__ rp = M.rp;
    const int32_t *__restrict__ ci = M.ci;
    const double *__restrict__ v = M.v;
    if constexpr (DICT) {
        if ((int)threadIdx.x < M.ndict) dl[threadIdx.x] = M.dict[threadIdx.x];
        __syncthreads();
    }
    if constexpr (TEST)
        pow_issue(a, &stopped, sh);
    else if (threadIdx.x == 0)
        stopped = a.res[0];
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int32_t per = (ngroups + (int32_t)gridDim.x - 1) / (int32_t)gridDim.x;
    const int32_t g0 = min(b * per, ngroups), g1 = min(g0 + per, ngroups);
    double g = 1.0, s0 = 0.0, s1 = 0.0;
    // the block's barrier before the stop decision (block-uniform)
    auto settle = [&]() -> bool {
        __syncthreads();
        if constexpr (TEST) return pow_test(a, stopped, sh, g);
        return stopped == 0.0;
    };
    if (g1 == g0 && !settle()) return;
    for (int32_t grp = g0; grp < g1; ++grp) {
        double acc = 0.0;
        int32_t row;
        bool lead;
        if constexpr (L > 0) {
            row = grp * (kBlock / L) + (int32_t)(threadIdx.x / L);
            const int lane = threadIdx.x & (L - 1);
            lead = lane == 0 && row < nrows;
            if (row < nrows) {  // whole L-group in or out together
                const int32_t s = rp[row], e = rp[row + 1];
                double acc0 = 0.0, acc1 = 0.0;
                int32_t k = s + lane;
                for (; k + L < e; k += 2 * L) {
                    acc0 += v[k] * x[ci[k]];
                    acc1 += v[k + L] * x[ci[k + L]];
                }
                if (k < e) acc0 += v[k] * x[ci[k]];
                acc = acc0 + acc1;
                if constexpr (L > 1) {
#pragma unroll
                    for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, L);
                }
            }
        } else {
            // CSR-stream: grp < ngroups, so the group has a row and every
            // thread of the block reaches its barriers
            const int32_t r0 = grp * rb, r1 = min(r0 + rb, nrows);
            const int tt = threadIdx.x;
            row = r0 + tt;
            lead = row < r1;
            const int32_t base = rp[r0], cnt = rp[r1] - base;
            if (cnt <= kStreamCap) {
                for (int32_t k0 = 0; k0 < cnt; k0 += kStreamU * kBlock) {
                    double vv[kStreamU];
                    int32_t cc[kStreamU];
#pragma unroll
                    for (int j = 0; j < kStreamU; ++j) {
                        const int32_t k = k0 + j * kBlock + tt;
                        const int32_t at = base + (k < cnt ? k : 0);
                        cc[j] = ci[at];
                        if constexpr (DICT)
                            vv[j] = dl[M.code[at]];
                        else
                            vv[j] = v[at];
                    }
                    double pr[kStreamU];
#pragma unroll
                    for (int j = 0; j < kStreamU; ++j) {
                        if constexpr (DICT)
                            pr[j] = (vv[j] * M.csc[cc[j]]) * x[cc[j]];
                        else
                            pr[j] = vv[j] * x[cc[j]];
                    }
#pragma unroll
                    for (int j = 0; j < kStreamU; ++j) {
                        const int32_t k = k0 + j * kBlock + tt;
                        if (k < cnt) prod[k] = pr[j];
                    }
                }
                __syncthreads();
                if (lead) {
                    const int32_t s = rp[row] - base, e = rp[row + 1] - base;
                    for (int32_t k = s; k < e; ++k) acc += prod[k];
                }
            } else if (lead) {
                for (int32_t k = rp[row]; k < rp[row + 1]; ++k) {
                    if constexpr (DICT)
                        acc += (dl[M.code[k]] * M.csc[ci[k]]) * x[ci[k]];
                    else
                        acc += v[k] * x[ci[k]];
                }
            }
            if constexpr (DICT)
                if (lead) acc *= M.rsc[row];
        }
        if (grp == g0 && !settle()) return;
        if (lead) {
            const double yr = acc * g;
            y[row] = yr;
            if constexpr (MODE != kPowAT) {
                const double tr = MODE == kPowQ ? x[row] * g : t[row];
                s0 += yr * tr;
                s1 += yr * yr;
            }
        }
        if constexpr (L == 0) __syncthreads();  // prod is rewritten by the next group
    }
    if constexpr (MODE != kPowAT) pow_block_partials(s0, s1, a.part + 2 * (int64_t)b);
}

// the test of the last iteration enqueued (one block), when the run ends
// at its iteration limit
static __global__ __launch_bounds__(kBlock) void k_pow_tail(PowArgs a) {
    __shared__ double sh[kBlock / 64][2];
    __shared__ double stopped;
    pow_issue(a, &stopped, sh);
    __syncthreads();
    double g;
    (void)pow_test(a, stopped, sh, g);
}

// ---------------------------------------------------------------------
// generic elementwise kernel
// ---------------------------------------------------------------------
template <class F>
__global__ __launch_bounds__(kBlock) void k_each(int64_t n, F f) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) f(i);
}

// ---------------------------------------------------------------------
// scaling kernels (src/kernels.jl:1746-2048)
// ---------------------------------------------------------------------
// column statistic over A's columns (rows of the stored transpose) and Q's
// rows: sqrt(max) (Ruiz) or sqrt(sum) (Pock-Chambolle), 0 -> 1.
// L lanes per row (L divides 64; the row's lanes stride over its nonzeros
// and combine by xor shuffles, so a wave reads its rows coalesced), a whole
// L-group in or out of range together
template <int L>
__device__ __forceinline__ double row_stat_lanes(const int32_t *__restrict__ rp, const double *__restrict__ v,
                                                 int64_t row, int lane, int ruiz) {
    double s = 0.0;
    for (int32_t k = rp[row] + lane; k < rp[row + 1]; k += L) s = ruiz ? dmax(s, fabs(v[k])) : s + fabs(v[k]);
    if constexpr (L > 1) {
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) {
            const double o = __shfl_xor(s, off, L);
            s = ruiz ? dmax(s, o) : s + o;
        }
    }
    return s;
}

// column statistic over A's columns (rows of the stored transpose) and Q's
// rows: sqrt(max) (Ruiz) or sqrt(sum) (Pock-Chambolle), 0 -> 1.
template <int L>
static __global__ __launch_bounds__(kBlock) void k_col_stat(int32_t n, const int32_t *__restrict__ at_rp,
                                                     const double *__restrict__ at_v,
                                                     const int32_t *__restrict__ q_rp,
                                                     const double *__restrict__ q_v, int ruiz,
                                                     double *__restrict__ out) {
    const int64_t j = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / L;
    const int lane = threadIdx.x & (L - 1);
    if (j >= n) return;
    const double sa = at_rp ? row_stat_lanes<L>(at_rp, at_v, j, lane, ruiz) : 0.0;
    const double sq = q_rp ? row_stat_lanes<L>(q_rp, q_v, j, lane, ruiz) : 0.0;
    const double t = ruiz ? dmax(sq, sa) : sq + sa;
    if (lane == 0) out[j] = t > 0.0 ? sqrt(t) : 1.0;
}

template <int L>
static __global__ __launch_bounds__(kBlock) void k_row_stat(int32_t m, const int32_t *__restrict__ rp,
                                                     const double *__restrict__ v, int ruiz,
                                                     double *__restrict__ out) {
    const int64_t i = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / L;
    const int lane = threadIdx.x & (L - 1);
    if (i >= m) return;
    const double s = row_stat_lanes<L>(rp, v, i, lane, ruiz);
    if (lane == 0) out[i] = s > 0.0 ? sqrt(s) : 1.0;
}

// v_ij = (v_ij * (1/r_i)) / c_j
template <int L>
static __global__ __launch_bounds__(kBlock) void k_scale_csr(int32_t nrow, const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ ci,
                                                      double *__restrict__ v,
                                                      const double *__restrict__ r,
                                                      const double *__restrict__ c) {
    const int64_t i = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / L;
    const int lane = threadIdx.x & (L - 1);
    if (i >= nrow) return;
    const double inv = 1.0 / r[i];
    for (int32_t k = rp[i] + lane; k < rp[i + 1]; k += L) v[k] = (v[k] * inv) / c[ci[k]];
}

}  // namespace hprqp
