// user_csr_qop.hip -- a Q-operator plugin for the tests of
// hprqp_qop_create_user: the part a user's own HIP library plays behind the
// plugin interface (AbstractQOperator / Qmap!,
// src/Q_operators/Q_operator_interface.jl:5-92).  It owns a CSR matrix Q on
// the device and, when the solver calls apply, enqueues Qx = Q x on the HIP
// stream the solver hands over, one thread per row, without synchronising.
// Test infrastructure: built by the build step next to this file and loaded
// only by tests.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "hprqp.h"

namespace {

struct UserCsr {
    int32_t n = 0;
    int64_t nnz = 0;
    int32_t *rp = nullptr, *ci = nullptr;
    double *v = nullptr;
    int64_t calls = 0;
};

__global__ __launch_bounds__(256) void k_user_csr(int32_t n, const int32_t *__restrict__ rp,
                                                  const int32_t *__restrict__ ci, const double *__restrict__ v,
                                                  const double *__restrict__ x, double *__restrict__ y) {
    const int32_t i = (int32_t)(blockIdx.x * 256 + threadIdx.x);
    if (i >= n) return;
    double s = 0.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * x[ci[k]];
    y[i] = s;
}

void release(UserCsr *u) {
    if (!u) return;
    if (u->rp) (void)hipFree(u->rp);
    if (u->ci) (void)hipFree(u->ci);
    if (u->v) (void)hipFree(u->v);
    delete u;
}

}  // namespace

extern "C" {

// Q is n x n CSR, 0-based; the structure is checked here so that the kernel's
// reads stay inside the arrays.  Returns 0, or 1 on bad input / HIP failure.
int user_csr_create(int32_t n, int64_t nnz, const int32_t *rp, const int32_t *ci, const double *v, void **out) {
    if (!out || n <= 0 || nnz < 0 || nnz >= INT32_MAX || !rp || (nnz > 0 && (!ci || !v))) return 1;
    *out = nullptr;
    if (rp[0] != 0 || rp[n] != nnz) return 1;
    for (int32_t i = 0; i < n; ++i)
        if (rp[i] > rp[i + 1]) return 1;
    for (int64_t k = 0; k < nnz; ++k)
        if (ci[k] < 0 || ci[k] >= n) return 1;
    UserCsr *u = new UserCsr();
    u->n = n;
    u->nnz = nnz;
    bool ok = hipMalloc((void **)&u->rp, sizeof(int32_t) * (size_t)(n + 1)) == hipSuccess &&
              hipMalloc((void **)&u->ci, sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1)) == hipSuccess &&
              hipMalloc((void **)&u->v, sizeof(double) * (size_t)(nnz > 0 ? nnz : 1)) == hipSuccess;
    ok = ok && hipMemcpy(u->rp, rp, sizeof(int32_t) * (size_t)(n + 1), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && nnz > 0) {
        ok = hipMemcpy(u->ci, ci, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(u->v, v, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
        release(u);
        return 1;
    }
    *out = u;
    return 0;
}

void user_csr_destroy(void *user) { release(static_cast<UserCsr *>(user)); }

// Qx = Q x on `stream`; x and Qx are device arrays of n doubles
int user_csr_apply(void *user, const double *x, double *Qx, void *stream) {
    UserCsr *u = static_cast<UserCsr *>(user);
    if (!u || !x || !Qx) return 1;
    hipLaunchKernelGGL(k_user_csr, dim3((unsigned)((u->n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, u->n,
                       (const int32_t *)u->rp, (const int32_t *)u->ci, (const double *)u->v, x, Qx);
    u->calls++;
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int64_t user_csr_problem_size(void *user) { return static_cast<UserCsr *>(user)->n; }
int64_t user_csr_temp_size(void *) { return 0; }
const char *user_csr_name(void *) { return "UserCSR"; }
int64_t user_csr_calls(void *user) { return static_cast<UserCsr *>(user)->calls; }

// the vtable a native caller registers with hprqp_qop_create_user
void user_csr_fill_vtable(hprqp_qop_vtable *vt) {
    std::memset(vt, 0, sizeof(*vt));
    vt->apply = user_csr_apply;
    vt->problem_size = user_csr_problem_size;
    vt->temp_size = user_csr_temp_size;
    vt->name = user_csr_name;
}

}  // extern "C"
