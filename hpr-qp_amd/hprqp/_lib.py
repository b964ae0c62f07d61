"""ctypes binding of libhprqp (include/hprqp.h).

The shared library is built in-tree next to this file by
``hprqp._build.build_library()``.  There is no CPU fallback: importing the
binding without the built library, or calling it on a machine without a
GPU, raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libhprqp.so"

HPRQP_OK = 0
ERRORS = {
    -1: "invalid argument",
    -2: "HIP runtime error",
    -3: "allocation failed",
    -4: "invalid state",
    -5: "unsupported",
    -6: "invalid problem data",
}


class hprqp_params(C.Structure):
    _fields_ = [
        ("stoptol", C.c_double),
        ("sigma", C.c_double),
        ("max_iter", C.c_int64),
        ("sigma_fixed", C.c_int32),
        ("problem_type", C.c_int32),
        ("time_limit", C.c_double),
        ("eig_factor", C.c_double),
        ("check_iter", C.c_int64),
        ("print_frequency", C.c_int64),
        ("use_Ruiz_scaling", C.c_int32),
        ("use_Pock_Chambolle_scaling", C.c_int32),
        ("use_bc_scaling", C.c_int32),
        ("precision", C.c_int32),
        ("verbose", C.c_int32),
        ("spmv_lanes", C.c_int32),
        ("time_kernels", C.c_int32),
        ("polish_stall_checks", C.c_int32),
        ("polish_switch", C.c_double),
        ("auto_save", C.c_int32),
        ("qap_recompute_Qw", C.c_int32),
        ("qap_family", C.c_int32),
        ("save_cb", C.c_void_p),
        ("save_user", C.c_void_p),
    ]


_D = C.POINTER(C.c_double)


class hprqp_save_state(C.Structure):
    """State the solver hands to the auto-save callback (include/hprqp.h)."""
    _fields_ = [
        ("n", C.c_int32),
        ("m", C.c_int32),
        ("iteration", C.c_int64),
        ("time_elapsed", C.c_double),
        ("sigma", C.c_double),
        ("err_Rp", C.c_double),
        ("err_Rd", C.c_double),
        ("primal_obj", C.c_double),
        ("dual_obj", C.c_double),
        ("rel_gap", C.c_double),
        ("kkt", C.c_double),
        ("x", _D), ("y", _D), ("z", _D), ("w", _D),
        ("best_iteration", C.c_int64),
        ("best_sigma", C.c_double),
        ("best_err_Rp", C.c_double),
        ("best_err_Rd", C.c_double),
        ("best_primal_obj", C.c_double),
        ("best_dual_obj", C.c_double),
        ("best_rel_gap", C.c_double),
        ("best_x", _D), ("best_y", _D), ("best_z", _D), ("best_w", _D),
    ]


SAVE_FN = C.CFUNCTYPE(C.c_int, C.POINTER(hprqp_save_state), C.c_void_p)


class hprqp_result(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("restarts", C.c_int32),
        ("iter", C.c_int64),
        ("iter_4", C.c_int64),
        ("iter_6", C.c_int64),
        ("time", C.c_double),
        ("time_4", C.c_double),
        ("time_6", C.c_double),
        ("power_time", C.c_double),
        ("setup_time", C.c_double),
        ("primal_obj", C.c_double),
        ("dual_obj", C.c_double),
        ("residuals", C.c_double),
        ("gap", C.c_double),
        ("err_Rp", C.c_double),
        ("err_Rd", C.c_double),
        ("lambda_max_A", C.c_double),
        ("lambda_max_Q", C.c_double),
        ("sigma", C.c_double),
        ("sigma0", C.c_double),
        ("Q_is_diag", C.c_int32),
        ("noC", C.c_int32),
        ("n_history", C.c_int64),
        ("polish_iter", C.c_int64),
        ("power_iter_A", C.c_int64),
        ("power_iter_Q", C.c_int64),
        ("kernel_ms", C.c_double),
        ("kernel_calls", C.c_int64),
        ("lanes_Q", C.c_int32),
        ("lanes_A", C.c_int32),
        ("lanes_AT", C.c_int32),
        ("qap_family", C.c_int32),
    ]


class hprqp_batch_result(C.Structure):
    _fields_ = [
        ("time", C.c_double),
        ("power_time", C.c_double),
        ("solved", C.c_int64),
        ("intervals", C.c_int64),
        ("max_iter_reached", C.c_int64),
        ("total_iter", C.c_int64),
        ("predict_time", C.c_double),
        ("interval_time", C.c_double),
    ]


QOP_APPLY_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)
QOP_SIZE_FN = C.CFUNCTYPE(C.c_int64, C.c_void_p)
# const char *(*name)(void *): declared as returning a plain pointer so that a
# Python callback can hand back the address of a buffer it keeps alive
QOP_NAME_FN = C.CFUNCTYPE(C.c_void_p, C.c_void_p)


class hprqp_ws_info(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("m", C.c_int32),
        ("Q_is_diag", C.c_int32),
        ("noC", C.c_int32),
        ("path", C.c_int32),
        ("precision", C.c_int32),
        ("lambda_max_A", C.c_double),
        ("lambda_max_Q", C.c_double),
        ("sigma0", C.c_double),
        ("b_scale", C.c_double),
        ("c_scale", C.c_double),
        ("norm_b", C.c_double),
        ("norm_c", C.c_double),
        ("power_iter_A", C.c_int64),
        ("power_iter_Q", C.c_int64),
        ("setup_time", C.c_double),
        ("power_time", C.c_double),
    ]


class hprqp_step_params(C.Structure):
    _fields_ = [
        ("sigma", C.c_double),
        ("h1", C.c_double),
        ("h2", C.c_double),
        ("lambda_max_A", C.c_double),
        ("lambda_max_Q", C.c_double),
        ("full", C.c_int32),
        ("reserved", C.c_int32),
    ]


class hprqp_kkt_out(C.Structure):
    _fields_ = [
        ("err_Rp", C.c_double),
        ("err_Rd", C.c_double),
        ("rel_gap", C.c_double),
        ("primal_obj", C.c_double),
        ("dual_obj", C.c_double),
        ("kkt", C.c_double),
    ]


PATH_STANDARD, PATH_NOC, PATH_LASSO, PATH_LP = 0, 1, 2, 3
(VEC_X_BAR, VEC_Y_BAR, VEC_Z_BAR, VEC_W_BAR, VEC_ATDY, VEC_QATDY, VEC_DIAG_Q, VEC_X, VEC_W, VEC_Y, VEC_ATY, VEC_QW,
 VEC_X_HAT, VEC_DX, VEC_DW, VEC_DY, VEC_S, VEC_ATY_BAR, VEC_TEMPV, VEC_QW_BAR, VEC_LAST_X, VEC_LAST_Y, VEC_LAST_W,
 VEC_LAST_ATY, VEC_LAST_QW, VEC_FACT1, VEC_FACT2, VEC_FACT) = range(28)


class hprqp_qop_vtable(C.Structure):
    _fields_ = [
        ("apply", QOP_APPLY_FN),
        ("problem_size", QOP_SIZE_FN),
        ("temp_size", QOP_SIZE_FN),
        ("name", QOP_NAME_FN),
    ]


_P = C.c_void_p
_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

# (name, restype, argtypes)
SIGNATURES = [
    ("hprqp_abi_version", C.c_int, []),
    ("hprqp_params_default", None, [C.POINTER(hprqp_params)]),
    ("hprqp_ctx_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("hprqp_ctx_destroy", C.c_int, [_P]),
    ("hprqp_last_error", C.c_char_p, [_P]),
    ("hprqp_ctx_sync", C.c_int, [_P]),
    ("hprqp_buf_alloc", C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    ("hprqp_buf_free", C.c_int, [_P, _P]),
    ("hprqp_buf_upload", C.c_int, [_P, _P, _P, C.c_size_t]),
    ("hprqp_buf_download", C.c_int, [_P, _P, _P, C.c_size_t]),
    ("hprqp_buf_device_ptr", _P, [_P]),
    ("hprqp_buf_size", C.c_size_t, [_P]),
    ("hprqp_csr_create", C.c_int, [_P, C.c_int32, C.c_int32, C.c_int64, _P, _P, _P, C.c_int, C.c_int,
                                   C.POINTER(_P)]),
    ("hprqp_csr_destroy", C.c_int, [_P, _P]),
    ("hprqp_csr_spmv", C.c_int, [_P, _P, C.c_int, _P, _P]),
    ("hprqp_qop_create_csr", C.c_int, [_P, C.c_int32, C.c_int64, _P, _P, _P, C.c_int, C.POINTER(_P)]),
    ("hprqp_qop_create_lasso", C.c_int, [_P, C.c_int32, C.c_int32, C.c_int64, _P, _P, _P, C.c_int,
                                         C.POINTER(_P)]),
    ("hprqp_qop_create_qap", C.c_int, [_P, C.c_int32, _P, _P, _P, _P, C.POINTER(_P)]),
    ("hprqp_qop_create_lowrank", C.c_int, [_P, C.c_int32, C.c_int32, _P, _P, C.POINTER(_P)]),
    ("hprqp_qop_create_user", C.c_int, [_P, C.POINTER(hprqp_qop_vtable), _P, C.POINTER(_P)]),
    ("hprqp_qop_destroy", C.c_int, [_P, _P]),
    ("hprqp_qop_apply", C.c_int, [_P, _P, _P, _P]),
    ("hprqp_qop_set_qap_family", C.c_int, [_P, _P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("hprqp_qop_apply_prec", C.c_int, [_P, _P, _P, _P, C.c_int]),
    ("hprqp_qop_apply_noalloc", C.c_int, [_P, _P, _P, _P]),
    ("hprqp_qop_info", C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_char_p)]),
    ("hprqp_model_create", C.c_int, [_P, C.c_int32, C.c_int32, _P, C.c_int64, _P, _P, _P, C.c_int,
                                     _P, _P, _P, _P, _P, C.c_double, C.c_double, C.POINTER(_P)]),
    ("hprqp_model_destroy", C.c_int, [_P, _P]),
    ("hprqp_solve", C.c_int, [_P, _P, C.POINTER(hprqp_params), _P, _P, _P, _P, _P, _P, _P, C.c_int64,
                              C.POINTER(hprqp_result)]),
    ("hprqp_profile_step", C.c_int, [_P, _P, C.POINTER(hprqp_params), C.c_int, _P, C.POINTER(C.c_int)]),
    ("hprqp_ws_create", C.c_int, [_P, _P, C.POINTER(hprqp_params), _P, _P, C.POINTER(_P),
                                  C.POINTER(hprqp_ws_info)]),
    ("hprqp_ws_destroy", C.c_int, [_P, _P]),
    ("hprqp_step", C.c_int, [_P, _P, C.POINTER(hprqp_step_params)]),
    ("hprqp_step_zxw1", C.c_int, [_P, _P, C.POINTER(hprqp_step_params)]),
    ("hprqp_step_tempv", C.c_int, [_P, _P, C.POINTER(hprqp_step_params)]),
    ("hprqp_step_y", C.c_int, [_P, _P, C.POINTER(hprqp_step_params)]),
    ("hprqp_step_w2", C.c_int, [_P, _P, C.POINTER(hprqp_step_params)]),
    ("hprqp_step_noc_zxw", C.c_int, [_P, _P, C.POINTER(hprqp_step_params)]),
    ("hprqp_step_noc_tempv", C.c_int, [_P, _P, C.POINTER(hprqp_step_params)]),
    ("hprqp_step_lasso", C.c_int, [_P, _P, C.POINTER(hprqp_step_params)]),
    ("hprqp_step_lp_zx", C.c_int, [_P, _P, C.POINTER(hprqp_step_params)]),
    ("hprqp_step_lp_y", C.c_int, [_P, _P, C.POINTER(hprqp_step_params)]),
    ("hprqp_kkt_reduce", C.c_int, [_P, _P, C.c_int64, C.POINTER(hprqp_kkt_out)]),
    ("hprqp_mnorm", C.c_int, [_P, _P, C.POINTER(hprqp_step_params), C.POINTER(C.c_double)]),
    ("hprqp_sigma_dots", C.c_int, [_P, _P, _P]),
    ("hprqp_sigma_diag_estimate", C.c_int, [_P, _P, C.c_double, C.c_double, C.POINTER(C.c_double)]),
    ("hprqp_update_Q_factors", C.c_int, [_P, _P, C.c_double]),
    ("hprqp_restart", C.c_int, [_P, _P]),
    ("hprqp_ws_collect", C.c_int, [_P, _P, _P, _P, _P, _P]),
    ("hprqp_ws_download", C.c_int, [_P, _P, C.c_int, _P, C.c_int64]),
    ("hprqp_ws_upload", C.c_int, [_P, _P, C.c_int, _P, C.c_int64]),
    ("hprqp_batch_create", C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, C.c_int64, _P, _P, _P, _P, _P, _P, _P,
                                     _P, _P, C.POINTER(_P)]),
    ("hprqp_batch_create_csr", C.c_int, [_P, C.c_int32, _P, _P, C.c_int64, _P, _P, _P, C.c_int64, _P, _P, _P, _P,
                                         _P, _P, _P, _P, _P, C.POINTER(_P)]),
    ("hprqp_batch_destroy", C.c_int, [_P, _P]),
    ("hprqp_mlp_create", C.c_int, [_P, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(_P)]),
    ("hprqp_mlp_destroy", C.c_int, [_P, _P]),
    ("hprqp_mlp_param_count", C.c_int64, [_P]),
    ("hprqp_mlp_get_params", C.c_int, [_P, _P, _P]),
    ("hprqp_mlp_set_params", C.c_int, [_P, _P, _P]),
    ("hprqp_mlp_forward", C.c_int, [_P, _P, C.c_int64, _P, _P]),
    ("hprqp_mlp_loss_grad", C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, _P, _P, C.c_float,
                                      C.POINTER(C.c_double), _P]),
    ("hprqp_mlp_train", C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, _P, _P, C.c_float, C.c_int32, C.c_int32,
                                  C.c_float, C.c_uint64, _P]),
    ("hprqp_batch_solve", C.c_int, [_P, _P, C.POINTER(hprqp_params), _P, _P, _P, _P, _P, _P, _P, _P,
                                    C.POINTER(hprqp_batch_result)]),
    ("hprqp_batch_predict", C.c_int, [_P, _P, _P, _P, _P, C.c_double, _P, _P]),
    ("hprqp_batch_solve_warm", C.c_int, [_P, _P, _P, _P, _P, C.c_double, C.POINTER(hprqp_params), _P, _P, _P, _P,
                                         _P, _P, C.POINTER(hprqp_batch_result)]),
]

_lib = None


def load() -> C.CDLL:
    """Load libhprqp.so (built in-tree); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"libhprqp.so not found at {LIB_PATH}; build it with __graft_entry__.build() "
            "(hipcc --offload-arch=gfx950). There is no CPU fallback.")
    lib = C.CDLL(str(LIB_PATH))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """Host pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


class HPRQPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, 'error')} ({code}): {msg}")
        self.code = code


def check(ctx, rc):
    if rc != HPRQP_OK:
        lib = load()
        msg = lib.hprqp_last_error(ctx).decode() if ctx else ""
        raise HPRQPError(rc, msg)


class Context:
    """One HIP device + stream (hprqp_ctx)."""

    def __init__(self, device: int = 0):
        lib = load()
        h = C.c_void_p()
        rc = lib.hprqp_ctx_create(int(device), C.byref(h))
        if rc != HPRQP_OK:
            raise HPRQPError(rc, f"cannot create a context on device {device}")
        self.h = h
        self.device = device
        self.lib = lib

    def close(self):
        if self.h:
            self.lib.hprqp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts = {}


def context(device: int = 0) -> Context:
    if device not in _contexts:
        _contexts[device] = Context(device)
    return _contexts[device]


class DeviceBuffer:
    """FP64 device vector (hprqp_buf)."""

    def __init__(self, ctx: Context, n: int):
        self.ctx = ctx
        self.n = int(n)
        h = C.c_void_p()
        check(ctx.h, ctx.lib.hprqp_buf_alloc(ctx.h, 8 * max(self.n, 1), C.byref(h)))
        self.h = h

    @classmethod
    def from_array(cls, ctx, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = cls(ctx, a.size)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        check(self.ctx.h, self.ctx.lib.hprqp_buf_upload(self.ctx.h, self.h, ptr(a), 8 * self.n))

    def download(self):
        out = np.empty(self.n, dtype=np.float64)
        check(self.ctx.h, self.ctx.lib.hprqp_buf_download(self.ctx.h, self.h, ptr(out), 8 * self.n))
        return out

    def free(self):
        if self.h:
            self.ctx.lib.hprqp_buf_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def exported_symbols():
    return [s[0] for s in SIGNATURES]


def lib_env_summary():
    return {"path": str(LIB_PATH), "exists": LIB_PATH.exists(), "pid": os.getpid()}
