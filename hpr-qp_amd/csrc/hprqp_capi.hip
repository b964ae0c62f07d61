// hprqp_capi.hip -- extern "C" entry points of libhprqp (include/hprqp.h).
// Each entry catches every exception and turns it into a status code plus a
// per-context message; nothing is thrown across the ABI.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <unordered_map>

#include "hprqp_internal.hpp"

using hprqp::DevCsr;
using hprqp::Error;
using hprqp::HostCsr;

namespace hprqp {

HostCsr host_csr_from(int32_t nrow, int32_t ncol, int64_t nnz, const int32_t *rowptr, const int32_t *colidx,
                      const double *vals, int index_base) {
    HPRQP_REQUIRE(nrow >= 0 && ncol >= 0 && nnz >= 0, HPRQP_E_ARG, "negative CSR dimension");
    HPRQP_REQUIRE(index_base == 0 || index_base == 1, HPRQP_E_ARG, "index_base must be 0 or 1");
    HPRQP_REQUIRE(nnz < INT32_MAX, HPRQP_E_ARG, "nnz exceeds int32 indexing");
    HostCsr h;
    h.nrow = nrow;
    h.ncol = ncol;
    h.rp.assign((size_t)nrow + 1, 0);
    if (nrow > 0) HPRQP_REQUIRE(rowptr != nullptr, HPRQP_E_ARG, "NULL rowptr");
    for (int32_t i = 0; i <= nrow && nrow > 0; ++i) h.rp[i] = rowptr[i] - index_base;
    HPRQP_REQUIRE(nrow == 0 || (h.rp[0] == 0 && h.rp[nrow] == nnz), HPRQP_E_ARG, "rowptr inconsistent with nnz");
    // the whole rowptr first: with rp[0] = 0, rp[nrow] = nnz and no descent,
    // every row's range lies inside colidx / vals
    for (int32_t i = 0; i < nrow; ++i) HPRQP_REQUIRE(h.rp[i] <= h.rp[i + 1], HPRQP_E_ARG, "rowptr not monotone");
    h.ci.resize((size_t)nnz);
    h.v.resize((size_t)nnz);
    for (int32_t i = 0; i < nrow; ++i) {
        std::vector<std::pair<int32_t, double>> row;
        row.reserve((size_t)(h.rp[i + 1] - h.rp[i]));
        for (int32_t k = h.rp[i]; k < h.rp[i + 1]; ++k) {
            const int32_t j = colidx[k] - index_base;
            HPRQP_REQUIRE(j >= 0 && j < ncol, HPRQP_E_ARG, "column index out of range");
            HPRQP_REQUIRE(std::isfinite(vals[k]), HPRQP_E_DATA, "matrix contains NaN or Inf values");
            row.emplace_back(j, vals[k]);
        }
        std::sort(row.begin(), row.end(), [](auto &a, auto &b) { return a.first < b.first; });
        for (size_t t = 1; t < row.size(); ++t)
            HPRQP_REQUIRE(row[t].first != row[t - 1].first, HPRQP_E_ARG, "duplicate entry in CSR row");
        for (size_t t = 0; t < row.size(); ++t) {
            h.ci[h.rp[i] + t] = row[t].first;
            h.v[h.rp[i] + t] = row[t].second;
        }
    }
    return h;
}

// the pattern of a square a made symmetric: where only one of (i, j),
// (j, i) is stored, the other is added as an explicit zero
HostCsr host_csr_symmetric_pattern(const HostCsr &q) {
    const int32_t n = q.nrow;
    HostCsr qt = host_csr_transpose(q);
    HostCsr s;
    s.nrow = s.ncol = n;
    s.rp.assign((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i) {
        int32_t a = q.rp[i], ae = q.rp[i + 1], b = qt.rp[i], be = qt.rp[i + 1];
        while (a < ae || b < be) {
            const int32_t ca = a < ae ? q.ci[a] : INT32_MAX, cb = b < be ? qt.ci[b] : INT32_MAX;
            if (ca == cb) {
                s.ci.push_back(ca);
                s.v.push_back(q.v[a]);
                ++a;
                ++b;
            } else if (ca < cb) {
                s.ci.push_back(ca);
                s.v.push_back(q.v[a]);
                ++a;
            } else {
                s.ci.push_back(cb);
                s.v.push_back(0.0);
                ++b;
            }
        }
        s.rp[i + 1] = (int32_t)s.ci.size();
    }
    return s;
}

// the value table of v: its distinct values (exact comparison, first
// occurrence order) and one code per entry; false (and nothing kept) when
// v is empty or has more than 256 distinct values
bool dict_encode(const std::vector<double> &v, std::vector<double> &dict, std::vector<uint8_t> &code) {
    dict.clear();
    code.clear();
    if (v.empty()) return false;
    std::unordered_map<uint64_t, uint8_t> idx;
    code.resize(v.size());
    for (size_t k = 0; k < v.size(); ++k) {
        uint64_t key;
        std::memcpy(&key, &v[k], 8);
        auto it = idx.find(key);
        if (it == idx.end()) {
            if (dict.size() == 256) {
                dict.clear();
                code.clear();
                return false;
            }
            it = idx.emplace(key, (uint8_t)dict.size()).first;
            dict.push_back(v[k]);
        }
        code[k] = it->second;
    }
    return true;
}

// for a symmetric pattern: the position of each entry's transpose partner
std::vector<int32_t> host_csr_transpose_positions(const HostCsr &s) {
    std::vector<int32_t> tpos(s.ci.size());
    std::vector<int32_t> next(s.rp.begin(), s.rp.end() - 1);
    for (int32_t i = 0; i < s.nrow; ++i)
        for (int32_t k = s.rp[i]; k < s.rp[i + 1]; ++k) {
            // rows are visited in increasing i, so the entries (j, i) of row
            // j appear in the same increasing order
            const int32_t j = s.ci[k];
            tpos[k] = next[j]++;
        }
    return tpos;
}

HostCsr host_csr_transpose(const HostCsr &a) {
    HostCsr t;
    t.nrow = a.ncol;
    t.ncol = a.nrow;
    const int64_t nnz = a.v.size();
    t.rp.assign((size_t)t.nrow + 1, 0);
    t.ci.resize((size_t)nnz);
    t.v.resize((size_t)nnz);
    for (int64_t k = 0; k < nnz; ++k) t.rp[a.ci[k] + 1]++;
    for (int32_t i = 0; i < t.nrow; ++i) t.rp[i + 1] += t.rp[i];
    std::vector<int32_t> next(t.rp.begin(), t.rp.end());
    for (int32_t i = 0; i < a.nrow; ++i)
        for (int32_t k = a.rp[i]; k < a.rp[i + 1]; ++k) {
            const int32_t p = next[a.ci[k]]++;
            t.ci[p] = i;
            t.v[p] = a.v[k];
        }
    return t;
}

int default_lanes(const HostCsr &h) {
    if (h.nrow <= 0) return 1;
    const double mean = (double)h.v.size() / (double)h.nrow;
    if (mean >= 256.0) return 256;  // long rows: one block per row
    // short rows: CSR-stream with 256/128/64 rows per block (negative code)
    if (mean <= 6.0) return -256;
    if (mean <= 12.0) return -128;
    if (mean <= 24.0) return -64;
    int L = 1;
    while (L * 2 <= mean && L < 64) L *= 2;
    return L;
}

void upload_csr(hipStream_t s, const HostCsr &h, DevCsr &d) {
    d.nrow = h.nrow;
    d.ncol = h.ncol;
    d.nnz = (int64_t)h.v.size();
    d.rp.alloc((int64_t)h.rp.size());
    if (!h.rp.empty())
        HPRQP_HIP(hipMemcpyAsync(d.rp.p, h.rp.data(), sizeof(int32_t) * h.rp.size(), hipMemcpyHostToDevice, s));
    d.ci.alloc(d.nnz);
    d.v.alloc(d.nnz);
    d.v32.release();
    if (d.nnz > 0) {
        HPRQP_HIP(hipMemcpyAsync(d.ci.p, h.ci.data(), sizeof(int32_t) * d.nnz, hipMemcpyHostToDevice, s));
        HPRQP_HIP(hipMemcpyAsync(d.v.p, h.v.data(), sizeof(double) * d.nnz, hipMemcpyHostToDevice, s));
    }
    d.lanes = default_lanes(h);
    d.cseg.release();  // a new pattern: chunk segments are rebuilt on use
    d.cpart.release();
    d.nchunk = 0;
    HPRQP_HIP(hipStreamSynchronize(s));
}

}  // namespace hprqp

// ---------------------------------------------------------------------
// Static scratch for hprqp_qop_apply_noalloc: a file-scope device array, so
// that entry point allocates nothing.  Each GPU the library runs on holds
// its own copy (8 MiB) until the process ends.  In host code the array's
// name is only a stand-in; its device address on each device comes from
// hipGetSymbolAddress, resolved once per device when the first context is
// created there and kept in g_scratch.
constexpr int64_t kScratchDoubles = int64_t(1) << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// with `device` current: this device's address of the scratch array
double *scratch_on(int device) {
    HPRQP_REQUIRE(device >= 0 && device < kMaxDevices, HPRQP_E_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        HPRQP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        HPRQP_REQUIRE(p != nullptr, HPRQP_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch[device] = static_cast<double *>(p);
    }
    return g_scratch[device];
}
}  // namespace

// ---------------------------------------------------------------------
template <class F>
static int guarded(hprqp_ctx *ctx, F &&f) {
    try {
        f();
        return HPRQP_OK;
    } catch (const Error &e) {
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->err = "host allocation failed";
        return HPRQP_E_ALLOC;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        return HPRQP_E_STATE;
    } catch (...) {
        if (ctx) ctx->err = "unknown error";
        return HPRQP_E_STATE;
    }
}

#define CTX_CHECK(ctx)                      \
    do {                                    \
        if (!(ctx)) return HPRQP_E_ARG;     \
        (void)hipSetDevice((ctx)->device);  \
    } while (0)

extern "C" {

int hprqp_abi_version(void) { return HPRQP_ABI_VERSION; }

void hprqp_params_default(hprqp_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->stoptol = 1e-6;
    p->sigma = -1;
    p->max_iter = 2147483647LL;
    p->sigma_fixed = 0;
    p->problem_type = HPRQP_PROBLEM_QP;
    p->time_limit = 3600.0;
    p->eig_factor = 1.05;
    p->check_iter = 100;
    p->print_frequency = -1;
    p->use_Ruiz_scaling = 1;
    p->use_Pock_Chambolle_scaling = 1;
    p->use_bc_scaling = 0;
    p->precision = HPRQP_PREC_FP64;
    p->verbose = 0;
    p->spmv_lanes = 0;
    p->time_kernels = 0;
    p->polish_stall_checks = 0;
    p->polish_switch = 5.0;
    p->qap_recompute_Qw = 0;
    p->qap_family = 0;
}

int hprqp_ctx_create(int device, hprqp_ctx **out) {
    if (!out) return HPRQP_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return HPRQP_E_HIP;
    if (device < 0 || device >= count) return HPRQP_E_ARG;
    hprqp_ctx *c = new (std::nothrow) hprqp_ctx();
    if (!c) return HPRQP_E_ALLOC;
    c->device = device;
    int rc = guarded(c, [&] {
        HPRQP_HIP(hipSetDevice(device));
        HPRQP_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HPRQP_HIP(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        HPRQP_HIP(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
        HPRQP_HIP(hipHostMalloc((void **)&c->pinned, sizeof(double) * 256, hipHostMallocDefault));
        c->pinned_count = 256;
        c->scratch = scratch_on(device);
    });
    if (rc != HPRQP_OK) {
        if (c->fork) (void)hipEventDestroy(c->fork);
        if (c->stream2) (void)hipStreamDestroy(c->stream2);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return HPRQP_OK;
}

int hprqp_ctx_destroy(hprqp_ctx *ctx) {
    if (!ctx) return HPRQP_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    if (ctx->stream2) {
        (void)hipStreamSynchronize(ctx->stream2);
        (void)hipStreamDestroy(ctx->stream2);
    }
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    delete ctx;
    return HPRQP_OK;
}

const char *hprqp_last_error(const hprqp_ctx *ctx) { return ctx ? ctx->err.c_str() : "NULL context"; }

int hprqp_ctx_sync(hprqp_ctx *ctx) {
    CTX_CHECK(ctx);
    return guarded(ctx, [&] { HPRQP_HIP(hipStreamSynchronize(ctx->stream)); });
}

// ---- buffers ---------------------------------------------------------
int hprqp_buf_alloc(hprqp_ctx *ctx, size_t bytes, hprqp_buf **out) {
    CTX_CHECK(ctx);
    if (!out) return HPRQP_E_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        hprqp_buf *b = new hprqp_buf();
        b->bytes = bytes;
        if (bytes) {
            hipError_t e = hipMalloc(&b->p, bytes);
            if (e != hipSuccess) {
                delete b;
                throw Error(HPRQP_E_ALLOC, std::string("hipMalloc failed: ") + hipGetErrorString(e));
            }
        }
        *out = b;
    });
}

int hprqp_buf_free(hprqp_ctx *ctx, hprqp_buf *buf) {
    CTX_CHECK(ctx);
    if (!buf) return HPRQP_OK;
    return guarded(ctx, [&] {
        HPRQP_HIP(hipStreamSynchronize(ctx->stream));
        if (buf->p) HPRQP_HIP(hipFree(buf->p));
        delete buf;
    });
}

int hprqp_buf_upload(hprqp_ctx *ctx, hprqp_buf *buf, const void *host, size_t bytes) {
    CTX_CHECK(ctx);
    if (!buf || (!host && bytes) || bytes > buf->bytes) return HPRQP_E_ARG;
    return guarded(ctx, [&] {
        if (bytes) HPRQP_HIP(hipMemcpyAsync(buf->p, host, bytes, hipMemcpyHostToDevice, ctx->stream));
        HPRQP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int hprqp_buf_download(hprqp_ctx *ctx, const hprqp_buf *buf, void *host, size_t bytes) {
    CTX_CHECK(ctx);
    if (!buf || (!host && bytes) || bytes > buf->bytes) return HPRQP_E_ARG;
    return guarded(ctx, [&] {
        if (bytes) HPRQP_HIP(hipMemcpyAsync(host, buf->p, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HPRQP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

void *hprqp_buf_device_ptr(hprqp_buf *buf) { return buf ? buf->p : nullptr; }
size_t hprqp_buf_size(const hprqp_buf *buf) { return buf ? buf->bytes : 0; }

// ---- CSR -------------------------------------------------------------
int hprqp_csr_create(hprqp_ctx *ctx, int32_t nrow, int32_t ncol, int64_t nnz, const int32_t *rowptr,
                     const int32_t *colidx, const double *vals, int index_base, int build_transpose,
                     hprqp_csr **out) {
    CTX_CHECK(ctx);
    if (!out) return HPRQP_E_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        HostCsr h = hprqp::host_csr_from(nrow, ncol, nnz, rowptr, colidx, vals, index_base);
        hprqp_csr *c = new hprqp_csr();
        try {
            hprqp::upload_csr(ctx->stream, h, c->a);
            if (build_transpose) {
                hprqp::upload_csr(ctx->stream, hprqp::host_csr_transpose(h), c->at);
                c->has_t = true;
            }
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

int hprqp_csr_destroy(hprqp_ctx *ctx, hprqp_csr *a) {
    CTX_CHECK(ctx);
    return guarded(ctx, [&] {
        HPRQP_HIP(hipStreamSynchronize(ctx->stream));
        delete a;
    });
}

int hprqp_csr_spmv(hprqp_ctx *ctx, const hprqp_csr *a, int transpose, const hprqp_buf *x, hprqp_buf *y) {
    CTX_CHECK(ctx);
    if (!a || !x || !y) return HPRQP_E_ARG;
    return guarded(ctx, [&] {
        const DevCsr &m = transpose ? a->at : a->a;
        HPRQP_REQUIRE(!transpose || a->has_t, HPRQP_E_STATE, "matrix was created without its transpose");
        HPRQP_REQUIRE(x->bytes >= sizeof(double) * (size_t)m.ncol, HPRQP_E_ARG, "x buffer too small");
        HPRQP_REQUIRE(y->bytes >= sizeof(double) * (size_t)m.nrow, HPRQP_E_ARG, "y buffer too small");
        hprqp::csr_spmv_fp64(ctx, m, (const double *)x->p, (double *)y->p);
    });
}

// ---- Q operators -------------------------------------------------------
int hprqp_qop_create_csr(hprqp_ctx *ctx, int32_t n, int64_t nnz, const int32_t *rowptr, const int32_t *colidx,
                         const double *vals, int index_base, hprqp_qop **out) {
    CTX_CHECK(ctx);
    if (!out) return HPRQP_E_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        hprqp_qop *q = new hprqp_qop();
        try {
            q->kind = QOP_CSR;
            q->n = n;
            q->name = "CSR";
            q->hq = hprqp::host_csr_from(n, n, nnz, rowptr, colidx, vals, index_base);
            hprqp::upload_csr(ctx->stream, q->hq, q->q);
        } catch (...) {
            delete q;
            throw;
        }
        *out = q;
    });
}

int hprqp_qop_create_lasso(hprqp_ctx *ctx, int32_t rows, int32_t n, int64_t nnz, const int32_t *rowptr,
                           const int32_t *colidx, const double *vals, int index_base, hprqp_qop **out) {
    CTX_CHECK(ctx);
    if (!out) return HPRQP_E_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        hprqp_qop *q = new hprqp_qop();
        try {
            q->kind = QOP_LASSO;
            q->n = n;
            q->name = "LASSO";
            HostCsr h = hprqp::host_csr_from(rows, n, nnz, rowptr, colidx, vals, index_base);
            hprqp::upload_csr(ctx->stream, h, q->la);
            hprqp::upload_csr(ctx->stream, hprqp::host_csr_transpose(h), q->lat);
            q->ltemp.alloc(rows);
        } catch (...) {
            delete q;
            throw;
        }
        *out = q;
    });
}

int hprqp_qop_create_qap(hprqp_ctx *ctx, int32_t nq, const double *A, const double *B, const double *S,
                         const double *T, hprqp_qop **out) {
    CTX_CHECK(ctx);
    if (!out || nq <= 0 || !A || !B || !S || !T) return HPRQP_E_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        hprqp_qop *q = new hprqp_qop();
        try {
            q->kind = QOP_QAP;
            q->nq = nq;
            q->n = nq * nq;
            q->name = "QAP";
            const int64_t nn = (int64_t)nq * nq;
            hprqp::DevVec<double> *dst[4] = {&q->qa, &q->qb, &q->qs, &q->qt};
            const double *src[4] = {A, B, S, T};
            for (int t = 0; t < 4; ++t) {
                for (int64_t k = 0; k < nn; ++k)
                    HPRQP_REQUIRE(std::isfinite(src[t][k]), HPRQP_E_DATA, "QAP matrix has NaN/Inf");
                dst[t]->alloc(nn);
                HPRQP_HIP(hipMemcpyAsync(dst[t]->p, src[t], sizeof(double) * nn, hipMemcpyHostToDevice,
                                         ctx->stream));
            }
            q->qtemp.alloc(nn);
            q->qtemp32.alloc(nn);
            HPRQP_HIP(hipStreamSynchronize(ctx->stream));
            hprqp::qap_prepare(ctx->stream, q, A, B, S, T);
        } catch (...) {
            delete q;
            throw;
        }
        *out = q;
    });
}

int hprqp_qop_create_lowrank(hprqp_ctx *ctx, int32_t n, int32_t k, const double *U, const double *diag,
                             hprqp_qop **out) {
    CTX_CHECK(ctx);
    if (!out || n <= 0 || k < 0 || !diag || (k > 0 && !U)) return HPRQP_E_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        hprqp_qop *q = new hprqp_qop();
        try {
            q->kind = QOP_LOWRANK;
            q->n = n;
            q->lr_k = k;
            q->name = "DiagonalPlusLowRank";
            const int64_t nk = (int64_t)n * k;
            for (int64_t i = 0; i < nk; ++i)
                HPRQP_REQUIRE(std::isfinite(U[i]), HPRQP_E_DATA, "U has NaN/Inf");
            for (int32_t i = 0; i < n; ++i)
                HPRQP_REQUIRE(std::isfinite(diag[i]), HPRQP_E_DATA, "diag has NaN/Inf");
            q->lr_U.alloc(nk);
            q->lr_d.alloc(n);
            q->lr_t.alloc(std::max(1, k));
            if (nk)
                HPRQP_HIP(hipMemcpyAsync(q->lr_U.p, U, sizeof(double) * nk, hipMemcpyHostToDevice, ctx->stream));
            HPRQP_HIP(hipMemcpyAsync(q->lr_d.p, diag, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
            HPRQP_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            delete q;
            throw;
        }
        *out = q;
    });
}

int hprqp_qop_create_user(hprqp_ctx *ctx, const hprqp_qop_vtable *vt, void *user, hprqp_qop **out) {
    CTX_CHECK(ctx);
    if (!out || !vt || !vt->apply || !vt->problem_size) return HPRQP_E_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        hprqp_qop *q = new hprqp_qop();
        q->kind = QOP_USER;
        q->vt = *vt;
        q->user = user;
        const int64_t n = vt->problem_size(user);
        if (n <= 0 || n >= INT32_MAX) {
            delete q;
            throw Error(HPRQP_E_ARG, "user operator problem_size out of range");
        }
        q->n = (int32_t)n;
        q->name = vt->name ? vt->name(user) : "USER";
        *out = q;
    });
}

int hprqp_qop_destroy(hprqp_ctx *ctx, hprqp_qop *q) {
    CTX_CHECK(ctx);
    return guarded(ctx, [&] {
        HPRQP_HIP(hipStreamSynchronize(ctx->stream));
        delete q;
    });
}

int hprqp_qop_apply(hprqp_ctx *ctx, hprqp_qop *q, const hprqp_buf *x, hprqp_buf *Qx) {
    CTX_CHECK(ctx);
    if (!q || !x || !Qx) return HPRQP_E_ARG;
    return guarded(ctx, [&] {
        HPRQP_REQUIRE(x->bytes >= sizeof(double) * (size_t)q->n && Qx->bytes >= sizeof(double) * (size_t)q->n,
                      HPRQP_E_ARG, "buffer smaller than the operator size");
        hprqp::qop_apply_fp64(ctx, q, (const double *)x->p, (double *)Qx->p);
    });
}

int hprqp_qop_apply_noalloc(hprqp_ctx *ctx, hprqp_qop *q, const hprqp_buf *x, hprqp_buf *Qx) {
    CTX_CHECK(ctx);
    if (!q || !x || !Qx) return HPRQP_E_ARG;
    return guarded(ctx, [&] {
        HPRQP_REQUIRE(x->bytes >= sizeof(double) * (size_t)q->n && Qx->bytes >= sizeof(double) * (size_t)q->n,
                      HPRQP_E_ARG, "buffer smaller than the operator size");
        const int64_t need = hprqp::qop_scratch_doubles(q);
        HPRQP_REQUIRE(need <= kScratchDoubles, HPRQP_E_ARG,
                      "operator temporaries (" + std::to_string(need) +
                          " doubles) exceed the static scratch of 1048576 doubles");
        HPRQP_REQUIRE(ctx->scratch != nullptr, HPRQP_E_STATE, "no scratch address for this device");
        hprqp::qop_apply_fp64(ctx, q, (const double *)x->p, (double *)Qx->p, ctx->scratch);
    });
}

int hprqp_qop_apply_prec(hprqp_ctx *ctx, hprqp_qop *q, const hprqp_buf *x, hprqp_buf *Qx, int precision) {
    CTX_CHECK(ctx);
    if (!q || !x || !Qx) return HPRQP_E_ARG;
    if (precision != HPRQP_PREC_FP64 && precision != HPRQP_PREC_MIXED) return HPRQP_E_ARG;
    return guarded(ctx, [&] {
        HPRQP_REQUIRE(x->bytes >= sizeof(double) * (size_t)q->n && Qx->bytes >= sizeof(double) * (size_t)q->n,
                      HPRQP_E_ARG, "buffer smaller than the operator size");
        if (precision == HPRQP_PREC_FP64)
            hprqp::qop_apply_fp64(ctx, q, (const double *)x->p, (double *)Qx->p);
        else
            hprqp::qop_apply_mixed(ctx, q, (const double *)x->p, (double *)Qx->p);
    });
}

int hprqp_qop_set_qap_family(hprqp_ctx *ctx, hprqp_qop *q, int32_t family, int32_t *fp64_family,
                             int32_t *bf16_family) {
    CTX_CHECK(ctx);
    if (!q) return HPRQP_E_ARG;
    return guarded(ctx, [&] {
        HPRQP_REQUIRE(q->kind == QOP_QAP, HPRQP_E_ARG, "not a QAP operator");
        HPRQP_REQUIRE(family == 0 || family == 2 || family == 3 || family == 4, HPRQP_E_ARG,
                      "qap_family must be 0, 2, 3 or 4");
        q->qfam64 = q->qfam16 = family;
        if (fp64_family) *fp64_family = hprqp::qap_family_of(q, false);
        if (bf16_family) *bf16_family = hprqp::qap_family_of(q, true);
    });
}

int hprqp_qop_info(const hprqp_qop *q, int64_t *n, int64_t *temp_size, const char **name) {
    if (!q) return HPRQP_E_ARG;
    if (n) *n = q->n;
    if (temp_size) {
        if (q->kind == QOP_LASSO)
            *temp_size = q->la.nrow;
        else if (q->kind == QOP_QAP)
            *temp_size = (int64_t)q->nq * q->nq;
        else if (q->kind == QOP_LOWRANK)
            *temp_size = q->lr_k;
        else if (q->kind == QOP_USER && q->vt.temp_size)
            *temp_size = q->vt.temp_size(q->user);
        else
            *temp_size = 0;
    }
    if (name) *name = q->name.c_str();
    return HPRQP_OK;
}

// ---- model -------------------------------------------------------------
int hprqp_model_create(hprqp_ctx *ctx, int32_t n, int32_t m, hprqp_qop *Q, int64_t a_nnz, const int32_t *a_rowptr,
                       const int32_t *a_colidx, const double *a_vals, int index_base, const double *c,
                       const double *AL, const double *AU, const double *l, const double *u, double obj_constant,
                       double lambda, hprqp_model **out) {
    CTX_CHECK(ctx);
    if (!out || !Q || n <= 0 || m < 0 || !c || !l || !u || (m > 0 && (!AL || !AU))) return HPRQP_E_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        HPRQP_REQUIRE(Q->n == n, HPRQP_E_ARG, "Q operator size differs from n");
        hprqp_model *md = new hprqp_model();
        try {
            md->n = n;
            md->m = m;
            md->Q = Q;
            if (m > 0) {
                md->A = hprqp::host_csr_from(m, n, a_nnz, a_rowptr, a_colidx, a_vals, index_base);
            } else {
                md->A.nrow = 0;
                md->A.ncol = n;
                md->A.rp.assign(1, 0);
            }
            md->c.assign(c, c + n);
            md->l.assign(l, l + n);
            md->u.assign(u, u + n);
            md->AL.assign(AL ? AL : c, AL ? AL + m : c);
            md->AU.assign(AU ? AU : c, AU ? AU + m : c);
            if (m == 0) {
                md->AL.clear();
                md->AU.clear();
            }
            // validation (qp_formulation, src/utils.jl:78-146)
            for (int32_t i = 0; i < n; ++i) {
                HPRQP_REQUIRE(std::isfinite(md->c[i]), HPRQP_E_DATA, "c vector contains NaN or Inf values");
                HPRQP_REQUIRE(!std::isnan(md->l[i]) && !std::isnan(md->u[i]), HPRQP_E_DATA,
                              "variable bounds l or u contain NaN values");
                HPRQP_REQUIRE(md->l[i] <= md->u[i], HPRQP_E_DATA, "infeasible variable bounds: l > u");
            }
            for (int32_t i = 0; i < m; ++i) {
                HPRQP_REQUIRE(!std::isnan(md->AL[i]) && !std::isnan(md->AU[i]), HPRQP_E_DATA,
                              "constraint bounds AL or AU contain NaN values");
                HPRQP_REQUIRE(md->AL[i] <= md->AU[i], HPRQP_E_DATA, "infeasible constraint bounds: AL > AU");
            }
            md->obj_constant = obj_constant;
            md->lambda = lambda;
            if (Q->kind == QOP_CSR) {
                // symmetric pattern completion (explicit zeros) and the map
                // from each entry to its transpose partner
                md->Qsym = hprqp::host_csr_symmetric_pattern(Q->hq);
                md->q_tpos = hprqp::host_csr_transpose_positions(md->Qsym);
                std::vector<double> qs(md->Qsym.v.size());
                for (size_t k = 0; k < qs.size(); ++k) qs[k] = (md->Qsym.v[k] + md->Qsym.v[md->q_tpos[k]]) / 2.0;
                if (!hprqp::dict_encode(qs, md->q_dict, md->q_code)) md->q_dict.clear();
            }
            if (hprqp::dict_encode(md->A.v, md->a_dict, md->a_code)) {
                // A^T's codes: the transpose of A with its codes as values
                HostCsr ac = md->A;
                for (size_t k = 0; k < ac.v.size(); ++k) ac.v[k] = (double)md->a_code[k];
                HostCsr tc = hprqp::host_csr_transpose(ac);
                md->at_code.resize(tc.v.size());
                for (size_t k = 0; k < tc.v.size(); ++k) md->at_code[k] = (uint8_t)tc.v[k];
            } else {
                md->a_dict.clear();
                md->a_code.clear();
            }
        } catch (...) {
            delete md;
            throw;
        }
        *out = md;
    });
}

int hprqp_model_destroy(hprqp_ctx *ctx, hprqp_model *model) {
    if (!ctx) return HPRQP_E_ARG;
    delete model;
    return HPRQP_OK;
}

int hprqp_solve(hprqp_ctx *ctx, const hprqp_model *model, const hprqp_params *params, const double *x0,
                const double *y0, double *x_out, double *y_out, double *z_out, double *w_out, double *hist,
                int64_t hist_cap, hprqp_result *result) {
    CTX_CHECK(ctx);
    if (!model || !result) return HPRQP_E_ARG;
    hprqp_params def;
    if (!params) {
        hprqp_params_default(&def);
        params = &def;
    }
    return guarded(ctx, [&] {
        hprqp::solve(ctx, model, params, x0, y0, x_out, y_out, z_out, w_out, hist, hist_cap, result);
    });
}

// ---- step-level interface ----------------------------------------------
#define WS_CHECK(ctx, ws)                                                                            \
    do {                                                                                             \
        CTX_CHECK(ctx);                                                                              \
        if (!(ws)) return HPRQP_E_ARG;                                                               \
        if (hprqp::ws_ctx(ws) != (ctx)) {                                                            \
            (ctx)->err = "workspace belongs to another context";                                     \
            return HPRQP_E_STATE;                                                                    \
        }                                                                                            \
    } while (0)

int hprqp_ws_create(hprqp_ctx *ctx, const hprqp_model *model, const hprqp_params *params, const double *x0,
                    const double *y0, hprqp_ws **out, hprqp_ws_info *info) {
    CTX_CHECK(ctx);
    if (!model || !out) return HPRQP_E_ARG;
    *out = nullptr;
    hprqp_params def;
    if (!params) {
        hprqp_params_default(&def);
        params = &def;
    }
    return guarded(ctx, [&] {
        hprqp_ws *w = hprqp::ws_create(ctx, model, params, x0, y0);
        if (info) hprqp::ws_info(w, info);
        *out = w;
    });
}

int hprqp_ws_destroy(hprqp_ctx *ctx, hprqp_ws *ws) {
    if (!ws) return HPRQP_OK;
    WS_CHECK(ctx, ws);
    return guarded(ctx, [&] {
        HPRQP_HIP(hipStreamSynchronize(ctx->stream));
        hprqp::ws_destroy(ws);
    });
}

static int ws_step_entry(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp, int kind) {
    WS_CHECK(ctx, ws);
    if (!sp) return HPRQP_E_ARG;
    return guarded(ctx, [&] { hprqp::ws_step(ws, kind, sp); });
}

int hprqp_step(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp) { return ws_step_entry(ctx, ws, sp, 0); }
int hprqp_step_zxw1(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp) {
    return ws_step_entry(ctx, ws, sp, 1);
}
int hprqp_step_tempv(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp) {
    return ws_step_entry(ctx, ws, sp, 2);
}
int hprqp_step_y(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp) { return ws_step_entry(ctx, ws, sp, 3); }
int hprqp_step_w2(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp) {
    return ws_step_entry(ctx, ws, sp, 4);
}
int hprqp_step_noc_zxw(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp) {
    return ws_step_entry(ctx, ws, sp, 5);
}
int hprqp_step_noc_tempv(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp) {
    return ws_step_entry(ctx, ws, sp, 6);
}
int hprqp_step_lasso(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp) {
    return ws_step_entry(ctx, ws, sp, 7);
}
int hprqp_step_lp_zx(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp) {
    return ws_step_entry(ctx, ws, sp, 8);
}
int hprqp_step_lp_y(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp) {
    return ws_step_entry(ctx, ws, sp, 9);
}

int hprqp_kkt_reduce(hprqp_ctx *ctx, hprqp_ws *ws, int64_t iter, hprqp_kkt_out *out) {
    WS_CHECK(ctx, ws);
    if (!out || iter < 0) return HPRQP_E_ARG;
    return guarded(ctx, [&] { hprqp::ws_kkt(ws, iter, out); });
}

int hprqp_mnorm(hprqp_ctx *ctx, hprqp_ws *ws, const hprqp_step_params *sp, double *out) {
    WS_CHECK(ctx, ws);
    if (!sp || !out) return HPRQP_E_ARG;
    return guarded(ctx, [&] { *out = hprqp::ws_mnorm(ws, sp); });
}

int hprqp_sigma_dots(hprqp_ctx *ctx, hprqp_ws *ws, double *out) {
    WS_CHECK(ctx, ws);
    if (!out) return HPRQP_E_ARG;
    return guarded(ctx, [&] { hprqp::ws_sigma_dots(ws, out); });
}

int hprqp_sigma_diag_estimate(hprqp_ctx *ctx, hprqp_ws *ws, double a, double b, double *est) {
    WS_CHECK(ctx, ws);
    if (!est) return HPRQP_E_ARG;
    return guarded(ctx, [&] { *est = hprqp::ws_sigma_diag(ws, a, b); });
}

int hprqp_update_Q_factors(hprqp_ctx *ctx, hprqp_ws *ws, double sigma) {
    WS_CHECK(ctx, ws);
    return guarded(ctx, [&] { hprqp::ws_q_factors(ws, sigma); });
}

int hprqp_restart(hprqp_ctx *ctx, hprqp_ws *ws) {
    WS_CHECK(ctx, ws);
    return guarded(ctx, [&] { hprqp::ws_restart(ws); });
}

int hprqp_ws_collect(hprqp_ctx *ctx, hprqp_ws *ws, double *x, double *y, double *z, double *w) {
    WS_CHECK(ctx, ws);
    return guarded(ctx, [&] { hprqp::ws_collect(ws, x, y, z, w); });
}

int hprqp_ws_download(hprqp_ctx *ctx, hprqp_ws *ws, int which, double *host, int64_t len) {
    WS_CHECK(ctx, ws);
    return guarded(ctx, [&] { hprqp::ws_download(ws, which, host, len); });
}

int hprqp_ws_upload(hprqp_ctx *ctx, hprqp_ws *ws, int which, const double *host, int64_t len) {
    WS_CHECK(ctx, ws);
    return guarded(ctx, [&] { hprqp::ws_upload(ws, which, host, len); });
}

}  // extern "C"

static void batch_up32(hprqp::DevVec<float> &dst, const double *src, int64_t len) {
    std::vector<float> f(src, src + len);
    dst.alloc(std::max<int64_t>(len, 1));
    if (len) HPRQP_HIP(hipMemcpy(dst.p, f.data(), sizeof(float) * len, hipMemcpyHostToDevice));
}

// what the two batch kinds share: the instances' sizes and offsets, their
// packed CSR A with per-instance transposes, the vectors and their norms,
// uploaded in FP64 and FP32, and the largest-first dispatch order by
// `cost` (per-iteration work of instance b given its nnz(A))
template <class Cost>
static void batch_pack_common(hprqp_ctx *ctx, hprqp_batch *B, int32_t count, const int32_t *n, const int32_t *m,
                              int64_t a_nnz, const int32_t *a_rowptr, const int32_t *a_colidx, const double *a_vals,
                              const double *c, const double *l, const double *u, const double *AL,
                              const double *AU, const double *obj_constant, Cost cost_of) {
    hipStream_t s = ctx->stream;
    B->count = count;
    B->n.assign(n, n + count);
    B->m.assign(m, m + count);
    B->noff.resize(count);
    B->moff.resize(count);
    std::vector<int64_t> arp_off(count), atrp_off(count), nnz_off(count);
    int64_t N = 0, M = 0, NZ = 0, ARP = 0, ATRP = 0;
    for (int32_t b = 0; b < count; ++b) {
        HPRQP_REQUIRE(n[b] > 0 && m[b] >= 0, HPRQP_E_ARG, "instance dimensions out of range (n > 0, m >= 0)");
        B->noff[b] = N;
        B->moff[b] = M;
        arp_off[b] = ARP;
        atrp_off[b] = ATRP;
        nnz_off[b] = NZ;
        int32_t nnz_b = 0;
        if (m[b] > 0) {
            // the instance's whole rowptr is checked before its last
            // entry is taken as the instance's nonzero count
            HPRQP_REQUIRE(a_rowptr != nullptr, HPRQP_E_ARG, "NULL a_rowptr with m > 0");
            const int32_t *rp = a_rowptr + ARP;
            HPRQP_REQUIRE(rp[0] == 0, HPRQP_E_ARG, "rowptr inconsistent with nnz");
            for (int32_t r = 0; r < m[b]; ++r) HPRQP_REQUIRE(rp[r] <= rp[r + 1], HPRQP_E_ARG, "rowptr not monotone");
            nnz_b = rp[m[b]];
        }
        N += n[b];
        M += m[b];
        NZ += nnz_b;
        ARP += m[b] + 1;
        ATRP += n[b] + 1;
        HPRQP_REQUIRE(NZ < INT32_MAX && N < INT32_MAX, HPRQP_E_ARG, "batch exceeds int32 indexing");
    }
    // the instances' nonzeros together fill a_colidx / a_vals exactly
    HPRQP_REQUIRE(NZ == a_nnz, HPRQP_E_ARG, "rowptr inconsistent with nnz");
    HPRQP_REQUIRE(NZ == 0 || (a_colidx && a_vals), HPRQP_E_ARG, "NULL a_colidx or a_vals with nnz > 0");
    HPRQP_REQUIRE(M == 0 || (AL && AU), HPRQP_E_ARG, "NULL AL or AU with m > 0");
    B->Ntot = N;
    B->Mtot = M;
    B->nnzA = NZ;
    // packed CSR with row pointers made global, and the per-instance transposes
    std::vector<int32_t> arp(ARP), aci(NZ), atrp(ATRP), atci(NZ);
    std::vector<double> av(NZ), atv(NZ);
    for (int32_t b = 0; b < count; ++b) {
        if (m[b] > 0) {
            HostCsr h = hprqp::host_csr_from(m[b], n[b], a_rowptr[arp_off[b] + m[b]], a_rowptr + arp_off[b],
                                             a_colidx + nnz_off[b], a_vals + nnz_off[b], 0);
            HostCsr t = hprqp::host_csr_transpose(h);
            for (int32_t r = 0; r <= m[b]; ++r) arp[arp_off[b] + r] = (int32_t)(h.rp[r] + nnz_off[b]);
            for (int32_t j = 0; j <= n[b]; ++j) atrp[atrp_off[b] + j] = (int32_t)(t.rp[j] + nnz_off[b]);
            std::copy(h.ci.begin(), h.ci.end(), aci.begin() + nnz_off[b]);
            std::copy(h.v.begin(), h.v.end(), av.begin() + nnz_off[b]);
            std::copy(t.ci.begin(), t.ci.end(), atci.begin() + nnz_off[b]);
            std::copy(t.v.begin(), t.v.end(), atv.begin() + nnz_off[b]);
        } else {
            arp[arp_off[b]] = (int32_t)nnz_off[b];
            for (int32_t j = 0; j <= n[b]; ++j) atrp[atrp_off[b] + j] = (int32_t)nnz_off[b];
        }
    }
    // per-instance norms of the data as given (src/utils.jl:404-408)
    B->obj_constant.assign(count, 0.0);
    B->norm_b2.assign(count, 0.0);
    B->norm_c2.assign(count, 0.0);
    B->norm_b_inf.assign(count, 0.0);
    B->norm_c_inf.assign(count, 0.0);
    for (int32_t b = 0; b < count; ++b) {
        if (obj_constant) B->obj_constant[b] = obj_constant[b];
        double s2 = 0, si = 0;
        for (int32_t i = 0; i < n[b]; ++i) {
            const int64_t g = B->noff[b] + i;
            HPRQP_REQUIRE(std::isfinite(c[g]), HPRQP_E_DATA, "c vector contains NaN or Inf values");
            HPRQP_REQUIRE(!std::isnan(l[g]) && !std::isnan(u[g]), HPRQP_E_DATA,
                          "variable bounds l or u contain NaN values");
            HPRQP_REQUIRE(!(l[g] > u[g]), HPRQP_E_DATA, "infeasible variable bounds: l > u");
            s2 += c[g] * c[g];
            si = std::max(si, std::fabs(c[g]));
        }
        B->norm_c2[b] = std::sqrt(s2);
        B->norm_c_inf[b] = si;
        s2 = si = 0;
        for (int32_t r = 0; r < m[b]; ++r) {
            const int64_t g = B->moff[b] + r;
            HPRQP_REQUIRE(!std::isnan(AL[g]) && !std::isnan(AU[g]), HPRQP_E_DATA,
                          "constraint bounds AL or AU contain NaN values");
            HPRQP_REQUIRE(!(AL[g] > AU[g]), HPRQP_E_DATA, "infeasible constraint bounds: AL > AU");
            const double a = AL[g] == -INFINITY ? 0.0 : AL[g], bb = AU[g] == INFINITY ? 0.0 : AU[g];
            const double t = std::max(std::fabs(a), std::fabs(bb));
            s2 += t * t;
            si = std::max(si, t);
        }
        B->norm_b2[b] = std::sqrt(s2);
        B->norm_b_inf[b] = si;
    }
    auto up = [&](auto &dst, const auto *src, int64_t len) {
        dst.alloc(std::max<int64_t>(len, 1));
        if (len) HPRQP_HIP(hipMemcpyAsync(dst.p, src, sizeof(*src) * len, hipMemcpyHostToDevice, s));
    };
    up(B->dn, n, count);
    up(B->dm, m, count);
    up(B->dnoff, B->noff.data(), count);
    up(B->dmoff, B->moff.data(), count);
    up(B->darp_off, arp_off.data(), count);
    up(B->datrp_off, atrp_off.data(), count);
    up(B->a_rp, arp.data(), ARP);
    up(B->a_ci, aci.data(), NZ);
    up(B->a_v, av.data(), NZ);
    up(B->at_rp, atrp.data(), ATRP);
    up(B->at_ci, atci.data(), NZ);
    up(B->at_v, atv.data(), NZ);
    up(B->c, c, N);
    up(B->l, l, N);
    up(B->u, u, N);
    up(B->AL, AL, M);
    up(B->AU, AU, M);
    // largest-first dispatch (longest-processing-time order)
    std::vector<int32_t> order(count);
    std::vector<double> cost(count);
    for (int32_t b = 0; b < count; ++b) {
        order[b] = b;
        cost[b] = cost_of(b, (int64_t)arp[arp_off[b] + m[b]] - nnz_off[b]);
    }
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return cost[x] > cost[y]; });
    up(B->dorder, order.data(), count);
    up(B->dc0, B->obj_constant.data(), count);
    up(B->dnbi, B->norm_b_inf.data(), count);
    up(B->dnci, B->norm_c_inf.data(), count);
    HPRQP_HIP(hipStreamSynchronize(s));
    batch_up32(B->a_v32, av.data(), NZ);
    batch_up32(B->at_v32, atv.data(), NZ);
    batch_up32(B->c32, c, N);
    batch_up32(B->l32, l, N);
    batch_up32(B->u32, u, N);
    batch_up32(B->AL32, AL, M);
    batch_up32(B->AU32, AU, M);
}

extern "C" {

int hprqp_batch_create(hprqp_ctx *ctx, int32_t count, const int32_t *n, const int32_t *m, const int32_t *k,
                       const double *U, const double *diag, int64_t a_nnz, const int32_t *a_rowptr,
                       const int32_t *a_colidx,
                       const double *a_vals, const double *c, const double *l, const double *u, const double *AL,
                       const double *AU, const double *obj_constant, hprqp_batch **out) {
    CTX_CHECK(ctx);
    if (!out || count <= 0 || !n || !m || !k || !diag || !c || !l || !u) return HPRQP_E_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        hprqp_batch *B = new hprqp_batch();
        try {
            hipStream_t s = ctx->stream;
            B->qkind = 0;
            B->k.assign(k, k + count);
            B->uoff.resize(count);
            B->auoff.resize(count);
            int64_t UT = 0;
            for (int32_t b = 0; b < count; ++b) {
                HPRQP_REQUIRE(n[b] > 0 && m[b] >= 0 && k[b] >= 0 && k[b] <= 64, HPRQP_E_ARG,
                              "instance dimensions out of range (n > 0, m >= 0, 0 <= k <= 64)");
                B->uoff[b] = UT;
                B->auoff[b] = B->AUtot;
                B->AUtot += (int64_t)m[b] * k[b];
                UT += (int64_t)n[b] * k[b];
            }
            HPRQP_REQUIRE(UT == 0 || U, HPRQP_E_ARG, "NULL U with k > 0");
            for (int64_t i = 0; i < UT; ++i) HPRQP_REQUIRE(std::isfinite(U[i]), HPRQP_E_DATA, "U has NaN/Inf");
            int64_t N = 0;
            for (int32_t b = 0; b < count; ++b) N += n[b];
            for (int64_t i = 0; i < N; ++i) HPRQP_REQUIRE(std::isfinite(diag[i]), HPRQP_E_DATA, "diag has NaN/Inf");
            B->Utot = UT;
            // per-iteration work: about two passes over U plus the A and A^T products
            batch_pack_common(ctx, B, count, n, m, a_nnz, a_rowptr, a_colidx, a_vals, c, l, u, AL, AU, obj_constant,
                              [&](int32_t b, int64_t nnz_b) {
                                  return 4.0 * (double)n[b] * (k[b] + 1) + 2.0 * (double)nnz_b + m[b];
                              });
            auto up = [&](auto &dst, const auto *src, int64_t len) {
                dst.alloc(std::max<int64_t>(len, 1));
                if (len) HPRQP_HIP(hipMemcpyAsync(dst.p, src, sizeof(*src) * len, hipMemcpyHostToDevice, s));
            };
            up(B->dk, k, count);
            up(B->duoff, B->uoff.data(), count);
            up(B->dauoff, B->auoff.data(), count);
            up(B->U, U, UT);
            up(B->d, diag, N);
            HPRQP_HIP(hipStreamSynchronize(s));
            batch_up32(B->U32, U, UT);
            batch_up32(B->d32, diag, N);
        } catch (...) {
            delete B;
            throw;
        }
        *out = B;
    });
}

int hprqp_batch_create_csr(hprqp_ctx *ctx, int32_t count, const int32_t *n, const int32_t *m, int64_t q_nnz,
                           const int32_t *q_rowptr, const int32_t *q_colidx, const double *q_vals, int64_t a_nnz,
                           const int32_t *a_rowptr, const int32_t *a_colidx, const double *a_vals, const double *c,
                           const double *l, const double *u, const double *AL, const double *AU,
                           const double *obj_constant, hprqp_batch **out) {
    CTX_CHECK(ctx);
    if (!out || count <= 0 || !n || !m || !c || !l || !u || !q_rowptr) return HPRQP_E_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        hprqp_batch *B = new hprqp_batch();
        try {
            hipStream_t s = ctx->stream;
            B->qkind = 1;
            // per instance: Q's rowptr (n + 1 entries, from 0) checked whole
            // before its last entry is taken as the instance's nnz(Q)
            std::vector<int64_t> qnz_off(count), qrp_in(count);
            int64_t QNZ = 0, QRP = 0;
            for (int32_t b = 0; b < count; ++b) {
                HPRQP_REQUIRE(n[b] > 0 && m[b] >= 0, HPRQP_E_ARG, "instance dimensions out of range (n > 0, m >= 0)");
                const int32_t *rp = q_rowptr + QRP;
                HPRQP_REQUIRE(rp[0] == 0, HPRQP_E_ARG, "Q rowptr inconsistent with nnz");
                for (int32_t i = 0; i < n[b]; ++i) HPRQP_REQUIRE(rp[i] <= rp[i + 1], HPRQP_E_ARG, "Q rowptr not monotone");
                qrp_in[b] = QRP;
                qnz_off[b] = QNZ;
                QNZ += rp[n[b]];
                QRP += n[b] + 1;
                HPRQP_REQUIRE(QNZ < INT32_MAX, HPRQP_E_ARG, "batch exceeds int32 indexing");
            }
            HPRQP_REQUIRE(QNZ == q_nnz, HPRQP_E_ARG, "Q rowptr inconsistent with nnz");
            HPRQP_REQUIRE(QNZ == 0 || (q_colidx && q_vals), HPRQP_E_ARG, "NULL q_colidx or q_vals with nnz > 0");
            // each Q as the single-instance model keeps it: the pattern made
            // symmetric (explicit zeros) with the map from each entry to its
            // transpose partner, for the post-scaling (Q + Q')/2
            std::vector<int32_t> srp, sci, stp;
            std::vector<double> sv;
            std::vector<int64_t> qrp_off(count);
            int64_t SNZ = 0, noff_run = 0;
            for (int32_t b = 0; b < count; ++b) {
                const int32_t nb = n[b];
                HostCsr q = hprqp::host_csr_from(nb, nb, q_rowptr[qrp_in[b] + nb], q_rowptr + qrp_in[b],
                                                 q_colidx + qnz_off[b], q_vals + qnz_off[b], 0);
                HostCsr sym = hprqp::host_csr_symmetric_pattern(q);
                std::vector<int32_t> tpos = hprqp::host_csr_transpose_positions(sym);
                // the batch runs the standard path: a diagonal Q, an empty Q
                // (LP) or mostly free variables (noC) take other paths
                bool offdiag = false;
                for (int32_t i = 0; i < nb && !offdiag; ++i) {
                    double sum = 0.0, d = 0.0;
                    for (int32_t t = q.rp[i]; t < q.rp[i + 1]; ++t) {
                        sum += std::fabs(q.v[t]);
                        if (q.ci[t] == i) d = q.v[t];
                    }
                    const double ad = std::fabs(d);
                    offdiag = !(sum == ad || std::fabs(sum - ad) <= 1.4901161193847656e-8 * std::max(sum, ad));
                }
                HPRQP_REQUIRE(offdiag, HPRQP_E_UNSUP,
                              "batched CSR instances need a Q with off-diagonal entries (diagonal or empty Q: solve singly)");
                int64_t nfree = 0;
                for (int32_t i = 0; i < nb; ++i)
                    if (l[noff_run + i] == -INFINITY && u[noff_run + i] == INFINITY) nfree++;
                HPRQP_REQUIRE(!((double)nfree > 0.8 * (double)nb), HPRQP_E_UNSUP,
                              "batched CSR instances run the standard path: more than 80% free variables (noC) "
                              "must be solved singly");
                noff_run += nb;
                qrp_off[b] = (int64_t)srp.size();
                for (int32_t i = 0; i <= nb; ++i) srp.push_back((int32_t)(sym.rp[i] + SNZ));
                sci.insert(sci.end(), sym.ci.begin(), sym.ci.end());
                sv.insert(sv.end(), sym.v.begin(), sym.v.end());
                for (int32_t t : tpos) stp.push_back((int32_t)(t + SNZ));
                SNZ += (int64_t)sym.ci.size();
                HPRQP_REQUIRE(SNZ < INT32_MAX, HPRQP_E_ARG, "batch exceeds int32 indexing");
            }
            B->nnzQ = SNZ;
            B->qrp_off = qrp_off;
            // per-iteration work: two Q products plus the A and A^T products
            batch_pack_common(ctx, B, count, n, m, a_nnz, a_rowptr, a_colidx, a_vals, c, l, u, AL, AU, obj_constant,
                              [&](int32_t b, int64_t nnz_b) {
                                  const int64_t qn = srp[qrp_off[b] + n[b]] - srp[qrp_off[b]];
                                  return 2.0 * (double)qn + 2.0 * (double)nnz_b + 4.0 * n[b] + m[b];
                              });
            auto up = [&](auto &dst, const auto *src, int64_t len) {
                dst.alloc(std::max<int64_t>(len, 1));
                if (len) HPRQP_HIP(hipMemcpyAsync(dst.p, src, sizeof(*src) * len, hipMemcpyHostToDevice, s));
            };
            up(B->dqrp_off, qrp_off.data(), count);
            up(B->q_rp, srp.data(), (int64_t)srp.size());
            up(B->q_ci, sci.data(), SNZ);
            up(B->q_tpos, stp.data(), SNZ);
            up(B->q_v, sv.data(), SNZ);
            // unused by this kind, allocated so every pointer is valid
            B->k.assign(count, 0);
            std::vector<int64_t> zero(count, 0);
            up(B->dk, B->k.data(), count);
            up(B->duoff, zero.data(), count);
            up(B->dauoff, zero.data(), count);
            B->U.alloc(1);
            B->d.alloc(1);
            B->U32.alloc(1);
            B->d32.alloc(1);
            HPRQP_HIP(hipStreamSynchronize(s));
        } catch (...) {
            delete B;
            throw;
        }
        *out = B;
    });
}

int hprqp_batch_destroy(hprqp_ctx *ctx, hprqp_batch *batch) {
    CTX_CHECK(ctx);
    return guarded(ctx, [&] {
        HPRQP_HIP(hipStreamSynchronize(ctx->stream));
        delete batch;
    });
}

int hprqp_batch_solve(hprqp_ctx *ctx, hprqp_batch *batch, const hprqp_params *params, const double *x0,
                      const double *y0, double *x_out, double *y_out, int32_t *status_out, int64_t *iter_out,
                      double *obj_out, double *kkt_out, hprqp_batch_result *result) {
    CTX_CHECK(ctx);
    if (!batch || !result) return HPRQP_E_ARG;
    hprqp_params def;
    if (!params) {
        hprqp_params_default(&def);
        params = &def;
    }
    return guarded(ctx, [&] {
        hprqp::batch_solve(ctx, batch, params, x0, y0, nullptr, nullptr, x_out, y_out, status_out, iter_out, obj_out,
                           kkt_out, result);
    });
}

int hprqp_batch_predict(hprqp_ctx *ctx, hprqp_batch *batch, hprqp_mlp *mlp, const float *feat_mean,
                        const float *feat_std, double y_scale, double *x0_out, double *y0_out) {
    CTX_CHECK(ctx);
    if (!batch || !mlp || !feat_mean || !feat_std) return HPRQP_E_ARG;
    return guarded(ctx, [&] {
        hprqp::DevVec<double> x0(std::max<int64_t>(batch->Ntot, 1)), y0(std::max<int64_t>(batch->Mtot, 1));
        hprqp::batch_predict(ctx, batch, mlp, feat_mean, feat_std, y_scale, x0.p, y0.p);
        if (x0_out) HPRQP_HIP(hipMemcpy(x0_out, x0.p, sizeof(double) * batch->Ntot, hipMemcpyDeviceToHost));
        if (y0_out && batch->Mtot)
            HPRQP_HIP(hipMemcpy(y0_out, y0.p, sizeof(double) * batch->Mtot, hipMemcpyDeviceToHost));
    });
}

int hprqp_batch_solve_warm(hprqp_ctx *ctx, hprqp_batch *batch, hprqp_mlp *mlp, const float *feat_mean,
                           const float *feat_std, double y_scale, const hprqp_params *params, double *x_out,
                           double *y_out, int32_t *status_out, int64_t *iter_out, double *obj_out, double *kkt_out,
                           hprqp_batch_result *result) {
    CTX_CHECK(ctx);
    if (!batch || !mlp || !feat_mean || !feat_std || !result) return HPRQP_E_ARG;
    hprqp_params def;
    if (!params) {
        hprqp_params_default(&def);
        params = &def;
    }
    return guarded(ctx, [&] {
        const auto t0 = std::chrono::steady_clock::now();
        hprqp::DevVec<double> x0(std::max<int64_t>(batch->Ntot, 1)), y0(std::max<int64_t>(batch->Mtot, 1));
        hprqp::batch_predict(ctx, batch, mlp, feat_mean, feat_std, y_scale, x0.p, y0.p);
        const double tp = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        hprqp::batch_solve(ctx, batch, params, nullptr, nullptr, x0.p, y0.p, x_out, y_out, status_out, iter_out,
                           obj_out, kkt_out, result);
        result->predict_time = tp;
    });
}

int hprqp_profile_step(hprqp_ctx *ctx, const hprqp_model *model, const hprqp_params *params, int reps,
                       double *ms_out, int *n_out) {
    CTX_CHECK(ctx);
    if (!model || !ms_out || reps <= 0) return HPRQP_E_ARG;
    hprqp_params def;
    if (!params) {
        hprqp_params_default(&def);
        params = &def;
    }
    return guarded(ctx, [&] { hprqp::profile_step(ctx, model, params, reps, ms_out, n_out); });
}

}  // extern "C"
