"""optimize(): HPRQP.jl's solve entry, running on the GPU through libhprqp.

Mirrors HPRQP_parameters / HPRQP_results (src/structs.jl:195-309) and
optimize (src/algorithm.jl:1511-1554).  There is no CPU path: use_gpu=False
or an invalid device raises instead of falling back.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import scipy.sparse as sp

from . import _lib
from .model import (AbstractQOperator, DiagonalPlusLowRank_Q_operator, LASSO_Q_operator, QAP_Q_operator,
                    QP_info_cpu)

STATUS = {0: "OPTIMAL", 1: "MAX_ITER", 2: "TIME_LIMIT"}
PROBLEM_TYPES = {"QP": 0, "LASSO": 1, "QAP": 2}
PRECISIONS = {"fp64": 0, "mixed": 1, "fp32": 2}


@dataclass
class HPRQP_parameters:
    stoptol: float = 1e-6
    sigma: float = -1.0
    max_iter: int = 2**31 - 1
    sigma_fixed: bool = False
    time_limit: float = 3600.0
    eig_factor: float = 1.05
    check_iter: int = 100
    warm_up: bool = False
    spmv_mode_Q: str = "auto"
    spmv_mode_A: str = "auto"
    print_frequency: int = -1
    device_number: int = 0
    use_Ruiz_scaling: bool = True
    use_bc_scaling: bool = False
    use_l2_scaling: bool = False
    use_Pock_Chambolle_scaling: bool = True
    problem_type: str = "QP"
    lam: float = 0.0
    initial_x: Optional[np.ndarray] = None
    initial_y: Optional[np.ndarray] = None
    auto_save: bool = False
    save_filename: str = "hprqp_autosave.h5"
    verbose: bool = True
    use_gpu: bool = True
    # MI355X-specific
    precision: str = "fp64"      # "fp64" (plain), "mixed" or "fp32"
    spmv_lanes: int = 0          # 0 = automatic
    time_kernels: bool = False   # HIP-event timing of the first fused step kernel
    polish_stall_checks: int = 0  # checks without KKT halving before FP64 polish (0: 100 mixed, 40 fp32, 400 fp32 offset form)
    polish_switch: float = 5.0  # mixed / plain fp32: FP64 polish once KKT < polish_switch * stoptol
    qap_recompute_Qw: bool = False  # QAP with doubly stochastic rows: True = two applies per iteration as the reference
    qap_family: int = 0  # QAP n >= 256: 0 autotune, 2 blocked, 3 LDS 64x64, 4 LDS 64x128 (split-BF16)

    def to_c(self) -> _lib.hprqp_params:
        p = _lib.hprqp_params()
        _lib.load().hprqp_params_default(C.byref(p))
        p.stoptol = float(self.stoptol)
        p.sigma = float(self.sigma)
        p.max_iter = int(self.max_iter)
        p.sigma_fixed = int(bool(self.sigma_fixed))
        if self.problem_type not in PROBLEM_TYPES:
            raise ValueError(f"unknown problem_type {self.problem_type}")
        p.problem_type = PROBLEM_TYPES[self.problem_type]
        p.time_limit = float(self.time_limit)
        p.eig_factor = float(self.eig_factor)
        p.check_iter = int(self.check_iter)
        if self.print_frequency != -1 and self.print_frequency <= 0:
            raise ValueError(f"Invalid print_frequency: {self.print_frequency}. It should be a positive "
                             "integer or -1 for automatic printing.")
        p.print_frequency = int(self.print_frequency)
        p.use_Ruiz_scaling = int(bool(self.use_Ruiz_scaling))
        p.use_Pock_Chambolle_scaling = int(bool(self.use_Pock_Chambolle_scaling))
        p.use_bc_scaling = int(bool(self.use_bc_scaling))
        if self.precision not in PRECISIONS:
            raise ValueError(f"unknown precision {self.precision}")
        p.precision = PRECISIONS[self.precision]
        p.verbose = int(bool(self.verbose))
        p.spmv_lanes = int(self.spmv_lanes)
        p.time_kernels = int(bool(self.time_kernels))
        p.polish_stall_checks = int(self.polish_stall_checks)
        p.polish_switch = float(self.polish_switch)
        p.auto_save = int(bool(self.auto_save))
        p.qap_recompute_Qw = int(bool(self.qap_recompute_Qw))
        if self.qap_family not in (0, 2, 3, 4):
            raise ValueError(f"qap_family must be 0, 2, 3 or 4, got {self.qap_family}")
        p.qap_family = int(self.qap_family)
        return p


def autosave_path(filename: str) -> str:
    """Where auto_save writes: a NumPy .npz archive (no HDF5 library is
    available here), at save_filename with ".npz" appended unless present."""
    return filename if filename.endswith(".npz") else filename + ".npz"


def load_autosave(filename: str, which: str = "best"):
    """(x, y) of a saved state ("best" or "current"), in the original space,
    for HPRQP_parameters.initial_x / initial_y -- resuming as the reference
    does from its saved file.  y is None when the problem has no rows."""
    if which not in ("best", "current"):
        raise ValueError("which must be 'best' or 'current'")
    with np.load(autosave_path(filename), allow_pickle=False) as f:
        x = f[f"{which}/x_org"].copy()
        y = f[f"{which}/y_org"].copy() if f"{which}/y_org" in f.files else None
    return x, y


def _autosave_writer(params: HPRQP_parameters):
    """The auto-save callback: writes the reference's groups (current/,
    best/, parameters/; save_state_to_hdf5!, src/algorithm.jl:183-301) as
    "group/name" entries of an .npz archive, replacing the file each time."""
    import datetime

    path = autosave_path(params.save_filename)

    def write(st_p, _user):
        try:
            st = st_p.contents
            n, m = st.n, st.m

            def vec(ptr, k):
                return np.ctypeslib.as_array(ptr, shape=(k,)).copy() if k > 0 else np.zeros(0)

            d = {
                "current/iteration": np.int64(st.iteration),
                "current/time_elapsed": np.float64(st.time_elapsed),
                "current/timestamp": np.str_(datetime.datetime.now().isoformat()),
                "current/x_org": vec(st.x, n), "current/w_org": vec(st.w, n), "current/z_org": vec(st.z, n),
                "current/sigma": np.float64(st.sigma),
                "current/err_Rp": np.float64(st.err_Rp), "current/err_Rd": np.float64(st.err_Rd),
                "current/primal_obj": np.float64(st.primal_obj), "current/dual_obj": np.float64(st.dual_obj),
                "current/rel_gap": np.float64(st.rel_gap), "current/KKTx_and_gap": np.float64(st.kkt),
                "best/x_org": vec(st.best_x, n), "best/w_org": vec(st.best_w, n), "best/z_org": vec(st.best_z, n),
                "best/sigma": np.float64(st.best_sigma), "best/iteration": np.int64(st.best_iteration),
                "best/err_Rp": np.float64(st.best_err_Rp), "best/err_Rd": np.float64(st.best_err_Rd),
                "best/primal_obj": np.float64(st.best_primal_obj), "best/dual_obj": np.float64(st.best_dual_obj),
                "best/rel_gap": np.float64(st.best_rel_gap),
                "best/KKTx_and_gap": np.float64(max(st.best_err_Rp, st.best_err_Rd, st.best_rel_gap)),
            }
            if m > 0:
                d["current/y_org"] = vec(st.y, m)
                d["best/y_org"] = vec(st.best_y, m)
            for k in ("stoptol", "sigma", "max_iter", "sigma_fixed", "time_limit", "eig_factor", "check_iter",
                      "warm_up", "spmv_mode_Q", "spmv_mode_A", "print_frequency", "device_number",
                      "use_Ruiz_scaling", "use_bc_scaling", "use_l2_scaling", "use_Pock_Chambolle_scaling",
                      "problem_type", "auto_save"):
                d["parameters/" + k] = np.asarray(getattr(params, k))
            d["parameters/lambda"] = np.float64(params.lam)
            if params.initial_x is not None:
                d["parameters/initial_x"] = np.asarray(params.initial_x, dtype=np.float64)
            if params.initial_y is not None:
                d["parameters/initial_y"] = np.asarray(params.initial_y, dtype=np.float64)
            with open(path, "wb") as f:
                np.savez(f, **d)
            if params.verbose:
                print(f"State saved to {path} at iteration {st.iteration}")
            return 0
        except Exception as e:  # an exception must not cross the C ABI
            print(f"auto-save failed: {e}")
            return 1

    return _lib.SAVE_FN(write)


@dataclass
class HPRQP_results:
    iter: int = 0
    iter_4: int = 0
    iter_6: int = 0
    time: float = 0.0
    time_4: float = 0.0
    time_6: float = 0.0
    power_time: float = 0.0
    primal_obj: float = 0.0
    residuals: float = 0.0
    gap: float = 0.0
    x: np.ndarray = field(default_factory=lambda: np.zeros(0))
    y: np.ndarray = field(default_factory=lambda: np.zeros(0))
    z: np.ndarray = field(default_factory=lambda: np.zeros(0))
    w: np.ndarray = field(default_factory=lambda: np.zeros(0))
    status: str = ""
    # diagnostics beyond the reference struct
    dual_obj: float = 0.0
    err_Rp: float = 0.0
    err_Rd: float = 0.0
    lambda_max_A: float = 0.0
    lambda_max_Q: float = 0.0
    sigma: float = 0.0
    sigma0: float = 0.0
    restarts: int = 0
    Q_is_diag: bool = False
    noC: bool = False
    setup_time: float = 0.0
    polish_iter: int = 0
    power_iter_A: int = 0
    power_iter_Q: int = 0
    kernel_ms: float = 0.0
    kernel_calls: int = 0
    lanes: tuple = (0, 0, 0)  # row-product shapes used for Q, A, A^T
    qap_family: int = 0  # QAP operator: kernel family of the iteration's passes (qap_family codes)
    history: np.ndarray = field(default_factory=lambda: np.zeros((0, 8)))


def _csr_arrays(M):
    M = sp.csr_matrix(M, dtype=np.float64)
    M.sort_indices()
    M.sum_duplicates()
    return (np.ascontiguousarray(M.indptr, dtype=np.int32), np.ascontiguousarray(M.indices, dtype=np.int32),
            np.ascontiguousarray(M.data, dtype=np.float64), M)


class DeviceQOperator:
    """A Q operator resident on the GPU (to_gpu of the reference)."""

    def __init__(self, ctx: _lib.Context, Q, n: int):
        self.ctx = ctx
        lib = ctx.lib
        h = C.c_void_p()
        keep = []
        if isinstance(Q, LASSO_Q_operator):
            rp, ci, v, M = _csr_arrays(Q.A)
            keep += [rp, ci, v]
            _lib.check(ctx.h, lib.hprqp_qop_create_lasso(ctx.h, M.shape[0], M.shape[1], M.nnz, _lib.ptr(rp),
                                                         _lib.ptr(ci), _lib.ptr(v), 0, C.byref(h)))
        elif isinstance(Q, QAP_Q_operator):
            mats = [np.asfortranarray(M).ravel(order="F") for M in (Q.A, Q.B, Q.S, Q.T)]
            mats = [np.ascontiguousarray(M) for M in mats]
            keep += mats
            _lib.check(ctx.h, lib.hprqp_qop_create_qap(ctx.h, Q.n, *(_lib.ptr(M) for M in mats), C.byref(h)))
        elif isinstance(Q, DiagonalPlusLowRank_Q_operator):
            U = np.ascontiguousarray(Q.U.ravel(order="F"))
            d = np.ascontiguousarray(Q.diag)
            keep += [U, d]
            _lib.check(ctx.h, lib.hprqp_qop_create_lowrank(ctx.h, Q.diag.size, Q.U.shape[1], _lib.ptr(U),
                                                           _lib.ptr(d), C.byref(h)))
        elif isinstance(Q, AbstractQOperator):
            keep += self._bind_user(ctx, Q, n, h)
        else:
            rp, ci, v, M = _csr_arrays(Q)
            if M.shape != (n, n):
                raise ValueError("Q must be n x n")
            keep += [rp, ci, v]
            _lib.check(ctx.h, lib.hprqp_qop_create_csr(ctx.h, n, M.nnz, _lib.ptr(rp), _lib.ptr(ci), _lib.ptr(v), 0,
                                                       C.byref(h)))
        self.h = h
        self.n = n
        self._keep = keep

    @staticmethod
    def _bind_user(ctx, Q, n, h):
        """A user AbstractQOperator through hprqp_qop_create_user: its
        Qmap_device, get_problem_size, get_temp_size and name become the
        vtable's C callbacks.  The callbacks (and the name buffer) are kept
        alive with the device operator; an exception in Qmap_device becomes
        a non-zero return, which the solver reports as a failed apply."""
        if type(Q).Qmap_device is AbstractQOperator.Qmap_device:
            raise NotImplementedError(f"{type(Q).__name__} has no Qmap_device: a user Q operator must enqueue "
                                      "Qx = Q x on the device (there is no CPU path)")
        size = int(Q.get_problem_size())
        if size != n:
            raise ValueError(f"operator problem size {size} differs from n = {n}")
        name_buf = C.create_string_buffer(str(getattr(Q, "name", type(Q).__name__)).encode())

        def apply(_user, x, qx, stream):
            try:
                Q.Qmap_device(x or 0, qx or 0, stream or 0)
                return 0
            except Exception as e:  # nothing may cross the C ABI
                print(f"user Q operator {type(Q).__name__}: Qmap_device failed: {e}")
                return 1

        vt = _lib.hprqp_qop_vtable()
        vt.apply = _lib.QOP_APPLY_FN(apply)
        vt.problem_size = _lib.QOP_SIZE_FN(lambda _u: size)
        vt.temp_size = _lib.QOP_SIZE_FN(lambda _u: int(Q.get_temp_size()))
        vt.name = _lib.QOP_NAME_FN(lambda _u: C.addressof(name_buf))
        _lib.check(ctx.h, ctx.lib.hprqp_qop_create_user(ctx.h, C.byref(vt), None, C.byref(h)))
        return [vt, name_buf, Q]

    def Qmap(self, x: np.ndarray, precision: str = "fp64") -> np.ndarray:
        """Qmap!(x, Qx, Q) through the device.  precision "fp64" (the
        reference's arithmetic) or "mixed" (the product the mixed-precision
        iteration uses: FP32 matrix data; split-BF16 x3 MFMA for QAP)."""
        xb = _lib.DeviceBuffer.from_array(self.ctx, x)
        qb = _lib.DeviceBuffer(self.ctx, self.n)
        if precision == "fp64":
            _lib.check(self.ctx.h, self.ctx.lib.hprqp_qop_apply(self.ctx.h, self.h, xb.h, qb.h))
        elif precision == "mixed":
            _lib.check(self.ctx.h, self.ctx.lib.hprqp_qop_apply_prec(self.ctx.h, self.h, xb.h, qb.h, 1))
        else:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        return qb.download()

    def set_qap_family(self, family: int):
        """QAP operator: the kernel family of its applies outside a solve
        (0 = size rule, 2 blocked, 3 LDS 64x64, 4 LDS 64x128 split-BF16);
        returns the (FP64, split-BF16) families the applies then run."""
        f64 = C.c_int32()
        f16 = C.c_int32()
        _lib.check(self.ctx.h, self.ctx.lib.hprqp_qop_set_qap_family(self.ctx.h, self.h, int(family), C.byref(f64),
                                                                     C.byref(f16)))
        return f64.value, f16.value

    def info(self):
        n = C.c_int64()
        t = C.c_int64()
        nm = C.c_char_p()
        _lib.check(self.ctx.h, self.ctx.lib.hprqp_qop_info(self.h, C.byref(n), C.byref(t), C.byref(nm)))
        return n.value, t.value, nm.value.decode()

    def close(self):
        if self.h:
            self.ctx.lib.hprqp_qop_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_gpu(Q, n: Optional[int] = None, device: int = 0) -> DeviceQOperator:
    if n is None:
        n = Q.get_problem_size() if isinstance(Q, AbstractQOperator) else Q.shape[0]
    return DeviceQOperator(_lib.context(device), Q, n)


class DeviceModel:
    """A formulated model uploaded once; solve() may be called repeatedly."""

    def __init__(self, model: QP_info_cpu, device: int = 0):
        self.ctx = _lib.context(device)
        self.model = model
        n, m = model.n, model.m
        self.qop = DeviceQOperator(self.ctx, model.Q, n)
        A = sp.csr_matrix(model.A, dtype=np.float64)
        rp, ci, v, A = _csr_arrays(A)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (model.c, model.AL, model.AU, model.l, model.u)]
        h = C.c_void_p()
        _lib.check(self.ctx.h, self.ctx.lib.hprqp_model_create(
            self.ctx.h, n, m, self.qop.h, A.nnz, _lib.ptr(rp), _lib.ptr(ci), _lib.ptr(v), 0,
            *(_lib.ptr(a) for a in arrs), float(model.obj_constant), float(model.lam), C.byref(h)))
        self.h = h
        self.n, self.m = n, m

    def solve(self, params: HPRQP_parameters, hist_cap: int = 100000) -> HPRQP_results:
        p = params.to_c()
        cb = None
        if params.auto_save:
            cb = _autosave_writer(params)  # kept alive for the call
            p.save_cb = C.cast(cb, C.c_void_p)
        n, m = self.n, self.m
        x0 = None if params.initial_x is None else np.ascontiguousarray(params.initial_x, dtype=np.float64)
        y0 = None if params.initial_y is None else np.ascontiguousarray(params.initial_y, dtype=np.float64)
        if x0 is not None and x0.size != n:
            raise ValueError(f"initial_x has length {x0.size}, expected {n}")
        if y0 is not None and y0.size != m:
            raise ValueError(f"initial_y has length {y0.size}, expected {m}")
        x = np.zeros(n)
        y = np.zeros(m)
        z = np.zeros(n)
        w = np.zeros(n)
        hist = np.zeros((hist_cap, 8))
        r = _lib.hprqp_result()
        _lib.check(self.ctx.h, self.ctx.lib.hprqp_solve(
            self.ctx.h, self.h, C.byref(p), _lib.ptr(x0), _lib.ptr(y0), _lib.ptr(x), _lib.ptr(y), _lib.ptr(z),
            _lib.ptr(w), _lib.ptr(hist), hist_cap, C.byref(r)))
        return HPRQP_results(
            iter=r.iter, iter_4=r.iter_4, iter_6=r.iter_6, time=r.time, time_4=r.time_4, time_6=r.time_6,
            power_time=r.power_time, primal_obj=r.primal_obj, residuals=r.residuals, gap=r.gap, x=x, y=y, z=z,
            w=w, status=STATUS.get(r.status, str(r.status)), dual_obj=r.dual_obj, err_Rp=r.err_Rp,
            err_Rd=r.err_Rd, lambda_max_A=r.lambda_max_A, lambda_max_Q=r.lambda_max_Q, sigma=r.sigma,
            sigma0=r.sigma0, restarts=r.restarts, Q_is_diag=bool(r.Q_is_diag), noC=bool(r.noC),
            setup_time=r.setup_time, polish_iter=r.polish_iter, power_iter_A=r.power_iter_A,
            power_iter_Q=r.power_iter_Q, kernel_ms=r.kernel_ms, kernel_calls=r.kernel_calls,
            lanes=(r.lanes_Q, r.lanes_A, r.lanes_AT), qap_family=r.qap_family,
            history=hist[: r.n_history].copy())

    def profile(self, params: HPRQP_parameters, reps: int = 50):
        """Average ms of the fused kernels (hprqp_profile_step)."""
        p = params.to_c()
        ms = np.zeros(8)
        cnt = C.c_int()
        _lib.check(self.ctx.h, self.ctx.lib.hprqp_profile_step(self.ctx.h, self.h, C.byref(p), int(reps),
                                                               _lib.ptr(ms), C.byref(cnt)))
        keys = ["zxw1", "tempv", "y", "w2", "iteration", "residuals"]
        return {k: float(ms[i]) for i, k in enumerate(keys)}

    def close(self):
        if self.h:
            self.ctx.lib.hprqp_model_destroy(self.ctx.h, self.h)
            self.h = None
        self.qop.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def optimize(model: QP_info_cpu, params: Optional[HPRQP_parameters] = None) -> HPRQP_results:
    """optimize(model, params) (src/algorithm.jl:1511-1554)."""
    params = params or HPRQP_parameters()
    if not params.use_gpu:
        raise NotImplementedError("this implementation runs on the GPU only (use_gpu=False is not available)")
    dm = DeviceModel(model, params.device_number)
    try:
        if params.warm_up:
            import copy

            wp = copy.copy(params)
            wp.max_iter = 200
            wp.verbose = False
            dm.solve(wp, hist_cap=16)
        return dm.solve(params)
    finally:
        dm.close()
