"""The C ABI's own input checks (what a Julia ccall caller meets when it
skips the host package's checks): hprqp_model_create rejects bad problem
data with the reference's messages (qp_formulation, src/utils.jl:78-146) and
bad CSR structure with HPRQP_E_ARG, before anything reaches the device; the
context stays usable afterwards.  hprqp_batch_create checks each instance's
row pointers before it sums them, and the sum against the total nnz it is
given, so no instance can claim entries past the packed arrays."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

HPRQP_E_ARG = -1
HPRQP_E_DATA = -6


def _valid():
    n, m = 4, 2
    A = sp.csr_matrix(np.array([[1.0, 1.0, 0.0, 0.0], [0.0, 1.0, -1.0, 2.0]]))
    return dict(n=n, m=m, rp=A.indptr.astype(np.int32), ci=A.indices.astype(np.int32), v=A.data.copy(),
                nnz=A.nnz, c=np.array([1.0, -1.0, 0.5, 0.0]), AL=np.array([0.0, -1.0]), AU=np.array([1.0, 1.0]),
                l=np.full(n, -2.0), u=np.full(n, 2.0))


def _create(d):
    from hprqp import _lib
    from hprqp.solver import DeviceQOperator

    ctx = _lib.context(0)
    qop = DeviceQOperator(ctx, sp.identity(d["n"], format="csr"), d["n"])
    h = C.c_void_p()
    arrs = [np.ascontiguousarray(d[k], dtype=np.float64) for k in ("c", "AL", "AU", "l", "u")]
    rc = ctx.lib.hprqp_model_create(ctx.h, d["n"], d["m"], qop.h, d["nnz"], _lib.ptr(d["rp"]), _lib.ptr(d["ci"]),
                                    _lib.ptr(d["v"]), 0, *(_lib.ptr(a) for a in arrs), 0.0, 0.0, C.byref(h))
    msg = ctx.lib.hprqp_last_error(ctx.h).decode() if rc else ""
    if rc == 0:
        ctx.lib.hprqp_model_destroy(ctx.h, h)
    else:
        assert not h.value, "a failed create must not hand out a model"
    qop.close()
    return rc, msg


def _bad(key, idx, val):
    d = _valid()
    d[key] = d[key].copy()
    d[key][idx] = val
    return d


@pytest.mark.parametrize("case,code,text", [
    (("c", 1, np.nan), HPRQP_E_DATA, "c vector contains NaN or Inf"),
    (("c", 0, np.inf), HPRQP_E_DATA, "c vector contains NaN or Inf"),
    (("l", 2, np.nan), HPRQP_E_DATA, "variable bounds l or u contain NaN"),
    (("l", 3, 3.0), HPRQP_E_DATA, "infeasible variable bounds: l > u"),
    (("AL", 0, np.nan), HPRQP_E_DATA, "constraint bounds AL or AU contain NaN"),
    (("AL", 1, 5.0), HPRQP_E_DATA, "infeasible constraint bounds: AL > AU"),
    (("v", 2, np.inf), HPRQP_E_DATA, "matrix contains NaN or Inf"),
    (("ci", 3, 4), HPRQP_E_ARG, "column index out of range"),
    (("ci", 0, -1), HPRQP_E_ARG, "column index out of range"),
    (("rp", 2, 4), HPRQP_E_ARG, "rowptr inconsistent with nnz"),
    # row 0 would claim entries 0..5 of a 5-entry array: rejected before any is read
    (("rp", 1, 6), HPRQP_E_ARG, "rowptr not monotone"),
    (("ci", 1, 0), HPRQP_E_ARG, "duplicate entry in CSR row"),
], ids=["c_nan", "c_inf", "l_nan", "l_gt_u", "AL_nan", "AL_gt_AU", "A_inf", "col_past_n", "col_negative",
        "rowptr_end", "rowptr_order", "duplicate"])
def test_model_create_rejects(case, code, text):
    rc, msg = _create(_bad(*case))
    print(f"{case}: rc={rc} msg={msg!r}")
    assert rc == code
    assert text in msg


def test_context_usable_after_rejections():
    from hprqp import HPRQP_parameters, build_from_QAbc, optimize

    assert _create(_bad("c", 1, np.nan))[0] == HPRQP_E_DATA
    assert _create(_valid()) == (0, "")
    d = _valid()
    A = sp.csr_matrix((d["v"], d["ci"], d["rp"]), shape=(d["m"], d["n"]))
    model = build_from_QAbc(sp.identity(d["n"], format="csr"), d["c"], A, d["AL"], d["AU"], d["l"], d["u"],
                            verbose=False)
    res = optimize(model, HPRQP_parameters(verbose=False))
    assert res.status == "OPTIMAL" and res.residuals < 1e-6


def test_products_reject_short_buffers_and_missing_transpose():
    """hprqp_csr_spmv and hprqp_qop_apply check buffer sizes (and that a
    transpose was built) on the host before launching anything."""
    from hprqp import _lib
    from hprqp.solver import DeviceQOperator

    HPRQP_E_STATE = -4
    ctx = _lib.context(0)
    M = sp.csr_matrix(np.arange(1.0, 16.0).reshape(3, 5))
    rp, ci, v = M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.copy()
    h = C.c_void_p()
    _lib.check(ctx.h, ctx.lib.hprqp_csr_create(ctx.h, 3, 5, M.nnz, _lib.ptr(rp), _lib.ptr(ci), _lib.ptr(v), 0, 0,
                                               C.byref(h)))
    x = np.arange(5.0)
    cases = [(0, 4, 3, HPRQP_E_ARG, "x buffer too small"), (0, 5, 2, HPRQP_E_ARG, "y buffer too small"),
             (1, 3, 5, HPRQP_E_STATE, "without its transpose")]
    for transpose, nx, ny, code, text in cases:
        xb, yb = _lib.DeviceBuffer(ctx, nx), _lib.DeviceBuffer(ctx, ny)
        rc = ctx.lib.hprqp_csr_spmv(ctx.h, h, transpose, xb.h, yb.h)
        assert rc == code and text in ctx.lib.hprqp_last_error(ctx.h).decode()
    xb, yb = _lib.DeviceBuffer.from_array(ctx, x), _lib.DeviceBuffer(ctx, 3)
    _lib.check(ctx.h, ctx.lib.hprqp_csr_spmv(ctx.h, h, 0, xb.h, yb.h))
    np.testing.assert_allclose(yb.download(), M @ x, rtol=1e-15)
    ctx.lib.hprqp_csr_destroy(ctx.h, h)

    qop = DeviceQOperator(ctx, sp.identity(6, format="csr"), 6)
    short, full = _lib.DeviceBuffer(ctx, 5), _lib.DeviceBuffer(ctx, 6)
    for a, b in ((short, full), (full, short)):
        rc = ctx.lib.hprqp_qop_apply(ctx.h, qop.h, a.h, b.h)
        assert rc == HPRQP_E_ARG and "smaller than the operator size" in ctx.lib.hprqp_last_error(ctx.h).decode()
    qop.close()


def _batch_valid():
    """Two instances packed as hprqp_batch_create takes them: (n, m, k) =
    (3, 1, 2) and (2, 2, 1); row pointers local to each instance."""
    return dict(n=np.array([3, 2], np.int32), m=np.array([1, 2], np.int32), k=np.array([2, 1], np.int32),
                U=np.arange(1.0, 9.0) * 0.1, d=np.full(5, 0.5), rp=np.array([0, 3, 0, 1, 2], np.int32),
                ci=np.array([0, 1, 2, 0, 1], np.int32), av=np.ones(5), c=np.linspace(-1.0, 1.0, 5),
                l=np.zeros(5), u=np.ones(5), AL=np.array([1.0, 0.0, 0.5]), AU=np.array([1.0, 1.0, 0.5]),
                c0=np.zeros(2), nnz=5)


def _batch_create(d):
    from hprqp import _lib

    ctx = _lib.context(0)
    P = lambda key: None if d[key] is None else _lib.ptr(np.ascontiguousarray(d[key]))
    h = C.c_void_p()
    rc = ctx.lib.hprqp_batch_create(ctx.h, 2, P("n"), P("m"), P("k"), P("U"), P("d"), d["nnz"], P("rp"), P("ci"),
                                    P("av"), P("c"), P("l"), P("u"), P("AL"), P("AU"), P("c0"), C.byref(h))
    msg = ctx.lib.hprqp_last_error(ctx.h).decode() if rc else ""
    if rc == 0:
        ctx.lib.hprqp_batch_destroy(ctx.h, h)
    else:
        assert not h.value, "a failed create must not hand out a batch"
    return rc, msg


def _batch_bad(key, idx, val):
    d = _batch_valid()
    if idx is None:
        d[key] = val
    else:
        d[key] = d[key].copy()
        d[key][idx] = val
    return d


@pytest.mark.parametrize("case,code,text", [
    # instance 0's last row pointer claims 9 nonzeros where a_colidx holds 5
    (("rp", 1, 9), HPRQP_E_ARG, "rowptr inconsistent with nnz"),
    (("rp", 2, 1), HPRQP_E_ARG, "rowptr inconsistent with nnz"),
    (("nnz", None, 6), HPRQP_E_ARG, "rowptr inconsistent with nnz"),
    (("rp", 3, 3), HPRQP_E_ARG, "rowptr not monotone"),
    (("rp", None, None), HPRQP_E_ARG, "NULL a_rowptr"),
    (("ci", None, None), HPRQP_E_ARG, "NULL a_colidx or a_vals"),
    (("AL", None, None), HPRQP_E_ARG, "NULL AL or AU"),
    (("U", None, None), HPRQP_E_ARG, "NULL U"),
    (("U", 4, np.inf), HPRQP_E_DATA, "U has NaN/Inf"),
    (("ci", 4, 2), HPRQP_E_ARG, "column index out of range"),
    (("l", 1, np.nan), HPRQP_E_DATA, "variable bounds l or u contain NaN"),
    (("u", 4, np.nan), HPRQP_E_DATA, "variable bounds l or u contain NaN"),
    (("l", 0, 2.0), HPRQP_E_DATA, "infeasible variable bounds: l > u"),
    (("AU", 2, np.nan), HPRQP_E_DATA, "constraint bounds AL or AU contain NaN"),
    (("AL", 1, 3.0), HPRQP_E_DATA, "infeasible constraint bounds: AL > AU"),
], ids=["rowptr_last_oversized", "rowptr_start", "nnz_total", "rowptr_order", "rowptr_null", "colidx_null",
        "AL_null", "U_null", "U_inf", "col_past_n", "l_nan", "u_nan", "l_gt_u", "AU_nan", "AL_gt_AU"])
def test_batch_create_rejects(case, code, text):
    rc, msg = _batch_create(_batch_bad(*case))
    print(f"{case}: rc={rc} msg={msg!r}")
    assert rc == code
    assert text in msg


def test_batch_create_accepts_the_valid_batch():
    assert _batch_create(_batch_valid()) == (0, "")
