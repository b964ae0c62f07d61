"""The iteration loop of solve() on the host, over the step-level C ABI.

HPRQP.jl's own loop (src/algorithm.jl:1977-2039) keeps the restart tests
(check_restart, :441-520), the sigma update (update_sigma!, :323-423, with
its golden-section search; a diagonal Q's estimate comes from the library,
hprqp_sigma_diag_estimate, as hprqp_solve takes it) and the termination
tests on the host and
calls the device for each operation of an iteration: the half-steps
(main_update_gpu!, :1250-1279), the residuals (compute_residuals!, :37-133),
the M-norm (compute_M_norm!, :689-739), the device half of update_sigma!
and of do_restart!.  This module is that loop written against libhprqp's
hprqp_ws_* / hprqp_step_* entry points, the way a Julia host would call
them through ccall.  It follows hprqp_solve's cadence exactly, so both give
the same iterates bit for bit; half_steps=True drives the individual
half-step entry points instead of hprqp_step.

Not covered here (hprqp_solve has them): the reduced-precision polish
hand-over, auto-save and the time limit.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional

import numpy as np

from . import _lib
from .solver import STATUS, DeviceModel, HPRQP_parameters, HPRQP_results

_PHI = (math.sqrt(5.0) - 1.0) / 2.0


def golden(ap, bp, cp, dp, lo, hi, tol, maxiter):
    """golden (src/kernels.jl:1449-1486): minimise a x + b / x + c x^2 / (1 + d x)."""
    f = lambda x: ap * x + bp / x + cp * x * x / (1 + dp * x)
    a, b = lo, hi
    c = b - _PHI * (b - a)
    d = a + _PHI * (b - a)
    fc, fd = f(c), f(d)
    for _ in range(maxiter):
        if fd < fc:
            a, c, fc = c, d, fd
            d = a + _PHI * (b - a)
            fd = f(d)
        else:
            b, d, fd = d, c, fc
            c = b - _PHI * (b - a)
            fc = f(c)
        if (b - a) < tol:
            break
    return 0.5 * (a + b)


def print_step(k):
    """The adaptive print cadence (print_frequency = -1)."""
    if k <= 0:
        return 10
    p = 1
    while p <= k // 10:
        p *= 10
    return max(p // 10, 10)


class Restart:
    """HPRQP_restart and check_restart (src/algorithm.jl:441-520)."""

    def __init__(self, sigma0):
        self.restart_flag = 0
        self.first_restart = True
        self.last_gap = self.current_gap = self.save_gap = math.inf
        self.inner = self.sufficient = self.necessary = self.long = self.times = 0
        self.weighted_norm = self.best_gap = math.inf
        self.best_sigma = sigma0

    def check(self, it, check_iter, sigma):
        self.restart_flag = 0
        if self.first_restart:
            if it == check_iter:
                self.first_restart = False
                self.restart_flag = 1
                self.weighted_norm = self.best_gap = self.current_gap
                self.best_sigma = sigma
        elif it % check_iter == 0:
            if self.current_gap < 0:
                self.current_gap = 1e-6
            if self.current_gap <= 0.2 * self.last_gap:
                self.sufficient += 1
                self.restart_flag = 1
            if self.current_gap <= 0.8 * self.last_gap and self.current_gap > 1.00 * self.save_gap:
                self.necessary += 1
                self.restart_flag = 2
            fact = 0.5 if (self.current_gap / self.weighted_norm > 1e-1) else 0.2
            if self.inner >= fact * it:
                self.long += 1
                self.restart_flag = 3
            if self.best_gap > self.current_gap:
                self.best_gap = self.current_gap
                self.best_sigma = sigma
            self.save_gap = self.current_gap


class Workspace:
    """A libhprqp step-level workspace (hprqp_ws) for a DeviceModel."""

    def __init__(self, dm: DeviceModel, params: HPRQP_parameters):
        self.dm, self.ctx = dm, dm.ctx
        self.lib = dm.ctx.lib
        p = params.to_c()
        x0 = None if params.initial_x is None else np.ascontiguousarray(params.initial_x, dtype=np.float64)
        y0 = None if params.initial_y is None else np.ascontiguousarray(params.initial_y, dtype=np.float64)
        h = C.c_void_p()
        self.info = _lib.hprqp_ws_info()
        _lib.check(self.ctx.h, self.lib.hprqp_ws_create(self.ctx.h, dm.h, C.byref(p), _lib.ptr(x0), _lib.ptr(y0),
                                                        C.byref(h), C.byref(self.info)))
        self.h = h
        self.sp = _lib.hprqp_step_params()

    def _params(self, sigma, h1, lamA, lamQ, full):
        sp = self.sp
        sp.sigma, sp.h1, sp.h2 = sigma, h1, 1.0 - h1
        sp.lambda_max_A, sp.lambda_max_Q, sp.full = lamA, lamQ, int(full)
        return C.byref(sp)

    def step(self, name, sigma, h1, lamA, lamQ, full):
        fn = getattr(self.lib, "hprqp_step" if name == "main" else "hprqp_step_" + name)
        _lib.check(self.ctx.h, fn(self.ctx.h, self.h, self._params(sigma, h1, lamA, lamQ, full)))

    def kkt(self, it):
        o = _lib.hprqp_kkt_out()
        _lib.check(self.ctx.h, self.lib.hprqp_kkt_reduce(self.ctx.h, self.h, it, C.byref(o)))
        return o

    def mnorm(self, sigma, lamA, lamQ):
        v = C.c_double()
        _lib.check(self.ctx.h, self.lib.hprqp_mnorm(self.ctx.h, self.h, self._params(sigma, 0.5, lamA, lamQ, 0),
                                                    C.byref(v)))
        return v.value

    def sigma_dots(self):
        out = np.zeros(6)
        _lib.check(self.ctx.h, self.lib.hprqp_sigma_dots(self.ctx.h, self.h, _lib.ptr(out)))
        return out

    def sigma_diag_estimate(self, a, b):
        """hprqp_sigma_diag_estimate: the diagonal-Q sigma estimate over the
        last sigma_dots' A^T dy and Q A^T dy."""
        out = C.c_double()
        _lib.check(self.ctx.h, self.lib.hprqp_sigma_diag_estimate(self.ctx.h, self.h, float(a), float(b),
                                                                  C.byref(out)))
        return out.value

    def update_Q_factors(self, sigma):
        _lib.check(self.ctx.h, self.lib.hprqp_update_Q_factors(self.ctx.h, self.h, sigma))

    def restart(self):
        _lib.check(self.ctx.h, self.lib.hprqp_restart(self.ctx.h, self.h))

    def download(self, which, length):
        out = np.zeros(length)
        _lib.check(self.ctx.h, self.lib.hprqp_ws_download(self.ctx.h, self.h, which, _lib.ptr(out), length))
        return out

    def upload(self, which, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        _lib.check(self.ctx.h, self.lib.hprqp_ws_upload(self.ctx.h, self.h, which, _lib.ptr(v), v.size))

    def collect(self):
        n, m = self.info.n, self.info.m
        x, y, z, w = np.zeros(n), np.zeros(m), np.zeros(n), np.zeros(n)
        _lib.check(self.ctx.h, self.lib.hprqp_ws_collect(self.ctx.h, self.h, _lib.ptr(x), _lib.ptr(y), _lib.ptr(z),
                                                         _lib.ptr(w)))
        return x, y, z, w

    def close(self):
        if self.h:
            self.lib.hprqp_ws_destroy(self.ctx.h, self.h)
            self.h = None


# the half-steps main_update runs along each path (src/algorithm.jl:1250-1279)
HALF_STEPS = {
    _lib.PATH_STANDARD: ("zxw1", "tempv", "y", "w2"),
    _lib.PATH_NOC: ("noc_zxw", "noc_tempv", "y"),
    _lib.PATH_LASSO: ("lasso",),
    _lib.PATH_LP: ("lp_zx", "lp_y"),
}


def _update_sigma(ws, r, res, sigma, lamA, lamQ, params):
    """update_sigma! (src/algorithm.jl:323-423): device dots, then the
    golden-section estimate, the log blend with the best sigma and the
    kappa correction on the host."""
    if params.sigma_fixed or r.restart_flag < 1:
        return sigma
    info = ws.info
    n, m, diag = info.n, info.m, bool(info.Q_is_diag)
    dxdx, dQwATdy, dQwdQw, dwdQw, ATdyQATdy, dydy = (float(v) for v in ws.sigma_dots())
    a, b, c, d = 0.0, dxdx, 0.0, 0.0
    if m > 0:
        a = lamA * dydy - 2 * dQwATdy
    if diag:
        a += dQwdQw
    else:
        a += lamQ * dwdQw
        if m > 0:
            c, d = ATdyQATdy, lamQ
    a, b = max(a, 1e-12), max(b, 1e-12)
    if diag:
        # the library's estimate, as hprqp_solve takes it (the golden search
        # on host copies for small n, the device's search from n = 2,048)
        est = ws.sigma_diag_estimate(a, b) if m > 0 else math.sqrt(b / a)
    else:
        est = golden(a, b, c, d, 1e-12, 1e12, 1e-13, 200) if m > 0 else math.sqrt(b / a)
    fact = math.exp(-0.05 * (r.current_gap / r.best_gap))
    temp_1 = max(min(res.err_Rd, res.err_Rp), min(res.rel_gap, r.current_gap))
    cand = math.exp(fact * math.log(est) + (1 - fact) * math.log(r.best_sigma))
    clamp = lambda x, lo, hi: hi if x > hi else (lo if x < lo else x)
    if temp_1 > 9e-10:
        kappa = 1.0
    elif temp_1 > 5e-10:
        kappa = clamp(math.sqrt(res.err_Rd / res.err_Rp), 1e-2, 100.0)
    else:
        kappa = clamp(res.err_Rd / res.err_Rp, 1e-2, 100.0)
    new = kappa * cand
    if diag and abs(sigma - new) > 1e-15:
        ws.update_Q_factors(new)
    return new


def solve_by_steps(model, params: Optional[HPRQP_parameters] = None, half_steps: bool = False,
                   device_model: Optional[DeviceModel] = None) -> HPRQP_results:
    """solve()'s loop (src/algorithm.jl:1977-2039) on the host over the step
    ABI.  Returns the same HPRQP_results as optimize().  device_model, when
    given, is used instead of uploading `model` again; it then shares the
    row-product shapes the autotuner cached in it with its own solves."""
    params = params or HPRQP_parameters()
    dm = device_model or DeviceModel(model, params.device_number)
    ws = Workspace(dm, params)
    try:
        info = ws.info
        lamA, lamQ, sigma = info.lambda_max_A, info.lambda_max_Q, info.sigma0
        r = Restart(sigma)
        check_iter, pf, max_iter = int(params.check_iter), int(params.print_frequency), int(params.max_iter)
        steps = HALF_STEPS[info.path]
        res, status, it = None, None, 0
        while it <= max_iter:
            ps = print_step(it) if pf == -1 else pf
            print_yes = (it % ps == 0) or it == max_iter
            if (it % check_iter == 0) or print_yes or res is None:
                res = ws.kkt(it)
            if res.kkt < params.stoptol:
                status = 0
            elif it == max_iter:
                status = 1
            r.check(it, check_iter, sigma)
            sigma = _update_sigma(ws, r, res, sigma, lamA, lamQ, params)
            if r.restart_flag > 0:  # do_restart! (src/algorithm.jl:522-555)
                ws.restart()
                r.last_gap, r.save_gap = r.current_gap, math.inf
                r.times += 1
                r.inner = 0
            if status is not None:
                break
            nxt = it + 1
            ps2 = print_step(nxt) if pf == -1 else pf
            full = (nxt % check_iter == 0) or (r.restart_flag > 0) or (nxt % ps2 == 0)
            h1 = 1.0 / (r.inner + 2.0)  # perform_iteration_step! (src/algorithm.jl:1881-1900)
            if half_steps:
                for name in steps:
                    if name == "tempv" and info.m == 0:
                        continue
                    ws.step(name, sigma, h1, lamA, lamQ, full)
            else:
                ws.step("main", sigma, h1, lamA, lamQ, full)
            if r.restart_flag > 0:
                r.last_gap = ws.mnorm(sigma, lamA, lamQ)
            if (it + 1) % check_iter == 0:
                r.current_gap = ws.mnorm(sigma, lamA, lamQ)
            r.inner += 1
            it += 1
        x, y, z, w = ws.collect()
        return HPRQP_results(iter=it, primal_obj=res.primal_obj, residuals=res.kkt, gap=res.rel_gap, x=x, y=y, z=z,
                             w=w, status=STATUS[status], dual_obj=res.dual_obj, err_Rp=res.err_Rp,
                             err_Rd=res.err_Rd, lambda_max_A=lamA, lambda_max_Q=lamQ, sigma=sigma,
                             sigma0=info.sigma0, restarts=r.times, Q_is_diag=bool(info.Q_is_diag),
                             noC=bool(info.noC), power_iter_A=info.power_iter_A, power_iter_Q=info.power_iter_Q)
    finally:
        ws.close()
        if device_model is None:
            dm.close()
