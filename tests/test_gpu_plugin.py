"""The user Q-operator plugin (AbstractQOperator / Qmap!,
src/Q_operators/Q_operator_interface.jl:5-92) through hprqp_qop_create_user.

tests/plugin/libuser_csr_qop.so stands for a user's HIP library: it owns a
CSR matrix Q and enqueues Qx = Q x on the stream the solver hands its apply
callback.  It is registered two ways: natively (a C vtable of the library's
own functions) and from Python (an AbstractQOperator whose Qmap_device calls
the library, bound through ctypes callbacks).  An operator Q skips scaling
(src/utils.jl:364-406), so both solves are checked against the oracle run on
the same Q as a sparse matrix with scaling off, which is the same algorithm."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from test_gpu_parity import _check

pytestmark = pytest.mark.gpu

PLUGIN = Path(__file__).resolve().parent / "plugin" / "libuser_csr_qop.so"


def _plugin_lib():
    if not PLUGIN.exists():
        pytest.fail(f"{PLUGIN} is not built (run __graft_entry__.build())")
    lib = C.CDLL(str(PLUGIN))
    lib.user_csr_create.argtypes = [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.POINTER(C.c_void_p)]
    lib.user_csr_create.restype = C.c_int
    lib.user_csr_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.user_csr_apply.restype = C.c_int
    lib.user_csr_calls.argtypes = [C.c_void_p]
    lib.user_csr_calls.restype = C.c_int64
    lib.user_csr_destroy.argtypes = [C.c_void_p]
    lib.user_csr_fill_vtable.argtypes = [C.c_void_p]
    return lib


def _make_operator_class():
    from hprqp.model import AbstractQOperator

    class PluginCsrOperator(AbstractQOperator):
        """A user operator written in Python over the plugin library."""

        name = "PluginCSR"

        def __init__(self, Q, fail=False):
            self.lib = _plugin_lib()
            self.Q = sp.csr_matrix(Q, dtype=np.float64)
            self.Q.sort_indices()
            self._arrs = (self.Q.indptr.astype(np.int32), self.Q.indices.astype(np.int32), self.Q.data.copy())
            h = C.c_void_p()
            rp, ci, v = self._arrs
            assert self.lib.user_csr_create(self.Q.shape[0], self.Q.nnz, rp.ctypes.data, ci.ctypes.data,
                                            v.ctypes.data, C.byref(h)) == 0
            self.h = h
            self.fail = fail

        def get_problem_size(self):
            return self.Q.shape[0]

        def Qmap(self, x):
            return self.Q @ x

        def Qmap_device(self, x_ptr, Qx_ptr, stream):
            if self.fail:
                raise RuntimeError("planted failure of the user apply")
            if self.lib.user_csr_apply(self.h, x_ptr, Qx_ptr, stream) != 0:
                raise RuntimeError("user_csr_apply failed")

        def calls(self):
            return self.lib.user_csr_calls(self.h)

        def close(self):
            if self.h:
                self.lib.user_csr_destroy(self.h)
                self.h = None

    return PluginCsrOperator


def _base_model():
    from hprqp import instances

    return instances.random_sparse_qp(1500, 700, seed=5)


def _oracle_unscaled(model):
    from oracle import oracle

    return oracle.solve(model, use_Ruiz_scaling=False, use_Pock_Chambolle_scaling=False)


def _native_solve(base, params):
    """hprqp_qop_create_user with the plugin library's own C vtable, then
    hprqp_model_create / hprqp_solve through the C ABI."""
    from hprqp import _lib
    from hprqp.solver import STATUS, HPRQP_results, _csr_arrays

    ctx = _lib.context(0)
    plug = _plugin_lib()
    Qc = sp.csr_matrix(base.Q)
    Qc.sort_indices()
    rp, ci, v = Qc.indptr.astype(np.int32), Qc.indices.astype(np.int32), Qc.data.copy()
    user = C.c_void_p()
    assert plug.user_csr_create(base.n, Qc.nnz, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, C.byref(user)) == 0
    vt = _lib.hprqp_qop_vtable()
    plug.user_csr_fill_vtable(C.addressof(vt))
    q = C.c_void_p()
    _lib.check(ctx.h, ctx.lib.hprqp_qop_create_user(ctx.h, C.byref(vt), user, C.byref(q)))
    n_, t_, nm = C.c_int64(), C.c_int64(), C.c_char_p()
    _lib.check(ctx.h, ctx.lib.hprqp_qop_info(q, C.byref(n_), C.byref(t_), C.byref(nm)))
    assert (n_.value, t_.value, nm.value) == (base.n, 0, b"UserCSR")
    # Qmap! itself through the plugin
    x = np.random.default_rng(0).standard_normal(base.n)
    xb, qb = _lib.DeviceBuffer.from_array(ctx, x), _lib.DeviceBuffer(ctx, base.n)
    _lib.check(ctx.h, ctx.lib.hprqp_qop_apply(ctx.h, q, xb.h, qb.h))
    np.testing.assert_allclose(qb.download(), Qc @ x, rtol=1e-13, atol=1e-13 * np.abs(Qc @ x).max())
    arp, aci, av, A = _csr_arrays(base.A)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (base.c, base.AL, base.AU, base.l, base.u)]
    mdl = C.c_void_p()
    _lib.check(ctx.h, ctx.lib.hprqp_model_create(ctx.h, base.n, base.m, q, A.nnz, _lib.ptr(arp), _lib.ptr(aci),
                                                 _lib.ptr(av), 0, *(_lib.ptr(a) for a in arrs), 0.0, 0.0,
                                                 C.byref(mdl)))
    p = params.to_c()
    xo, yo, zo, wo = np.zeros(base.n), np.zeros(base.m), np.zeros(base.n), np.zeros(base.n)
    r = _lib.hprqp_result()
    _lib.check(ctx.h, ctx.lib.hprqp_solve(ctx.h, mdl, C.byref(p), None, None, _lib.ptr(xo), _lib.ptr(yo),
                                          _lib.ptr(zo), _lib.ptr(wo), None, 0, C.byref(r)))
    calls = plug.user_csr_calls(user)
    ctx.lib.hprqp_model_destroy(ctx.h, mdl)
    ctx.lib.hprqp_qop_destroy(ctx.h, q)
    plug.user_csr_destroy(user)
    return HPRQP_results(iter=r.iter, primal_obj=r.primal_obj, residuals=r.residuals, x=xo, y=yo, z=zo, w=wo,
                         status=STATUS[r.status], lambda_max_A=r.lambda_max_A, lambda_max_Q=r.lambda_max_Q,
                         restarts=r.restarts, Q_is_diag=bool(r.Q_is_diag), noC=bool(r.noC),
                         power_iter_A=r.power_iter_A, power_iter_Q=r.power_iter_Q), calls


def test_native_vtable_plugin_solves_like_the_oracle():
    from hprqp import HPRQP_parameters

    base = _base_model()
    res, calls = _native_solve(base, HPRQP_parameters(verbose=False))
    ref = _oracle_unscaled(base)
    print(f"plugin (native vtable): {res.status} iter={res.iter} obj={res.primal_obj:.12e} apply calls={calls}; "
          f"oracle iter={ref.iter} obj={ref.primal_obj:.12e}")
    _check(res, ref)
    # the solver's Q products ran in the plugin: two per iteration plus the
    # power method, the residual checks and the M-norm
    assert calls >= 2 * res.iter + res.power_iter_Q


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_python_operator_bound_by_callbacks(precision):
    """The same operator as a Python AbstractQOperator (QP_info_cpu with Q an
    operator, as the reference's custom operators enter): optimize() binds
    its Qmap_device through ctypes; FP64 gives the native-vtable solve
    exactly (the same kernels on the same stream), mixed reaches 1e-6 KKT
    with the objective within 1e-6."""
    from hprqp import HPRQP_parameters, optimize
    from hprqp.model import QP_info_cpu

    base = _base_model()
    Op = _make_operator_class()
    op = Op(base.Q)
    model = QP_info_cpu(Q=op, c=base.c, A=base.A, AL=base.AL, AU=base.AU, l=base.l, u=base.u)
    try:
        res = optimize(model, HPRQP_parameters(verbose=False, precision=precision))
        calls = op.calls()
    finally:
        op.close()
    native, _ = _native_solve(base, HPRQP_parameters(verbose=False))
    print(f"python operator {precision}: {res.status} iter={res.iter} obj={res.primal_obj:.12e} "
          f"apply calls={calls}; native iter={native.iter}")
    assert calls > 0
    if precision == "fp64":
        assert res.iter == native.iter and res.restarts == native.restarts
        np.testing.assert_array_equal(res.x, native.x)
        np.testing.assert_array_equal(res.y, native.y)
    else:
        assert res.status == "OPTIMAL" and res.residuals < 1e-6
        assert abs(res.primal_obj - native.primal_obj) / (1 + abs(native.primal_obj)) < 1e-6


def test_failing_user_apply_is_an_error_and_the_context_survives():
    """A Qmap_device that raises returns non-zero through the callback: the
    solve fails with HPRQP_E_ARG and the solver's message, nothing crosses
    the ABI, and the context solves again afterwards."""
    from hprqp import HPRQP_parameters, optimize
    from hprqp._lib import HPRQPError
    from hprqp.model import QP_info_cpu

    base = _base_model()
    Op = _make_operator_class()
    bad = Op(base.Q, fail=True)
    model = QP_info_cpu(Q=bad, c=base.c, A=base.A, AL=base.AL, AU=base.AU, l=base.l, u=base.u)
    try:
        with pytest.raises(HPRQPError) as ei:
            optimize(model, HPRQP_parameters(verbose=False))
    finally:
        bad.close()
    assert ei.value.code == -1 and "user Q operator apply failed" in str(ei.value)
    res = optimize(base, HPRQP_parameters(verbose=False))
    assert res.status == "OPTIMAL"


def test_operator_without_device_apply_is_refused():
    """No CPU path: an AbstractQOperator without Qmap_device is refused."""
    from hprqp import HPRQP_parameters, optimize
    from hprqp.model import AbstractQOperator, QP_info_cpu

    class CpuOnly(AbstractQOperator):
        def __init__(self, Q):
            self.Q = Q

        def get_problem_size(self):
            return self.Q.shape[0]

        def Qmap(self, x):
            return self.Q @ x

    base = _base_model()
    model = QP_info_cpu(Q=CpuOnly(base.Q), c=base.c, A=base.A, AL=base.AL, AU=base.AU, l=base.l, u=base.u)
    with pytest.raises(NotImplementedError):
        optimize(model, HPRQP_parameters(verbose=False))
