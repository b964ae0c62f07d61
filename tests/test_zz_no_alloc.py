"""hprqp_qop_apply_noalloc: the FP64 Qmap! with the operator's temporaries in
the library's static device array instead of memory it allocates.  Checked
through the C ABI against hprqp_qop_apply on the same input (the same
kernels and row-product shapes, so the same bits), on calls in a row where
each operator kind is applied at a large size and then at a smaller one
(the second call reuses the array the first one filled); an operator whose
temporaries exceed the array is refused before any launch; and the calls
leave the device's free memory unchanged."""
import numpy as np
import pytest

from conftest import fixture_path

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 1 << 20


def _apply(dev, x, noalloc):
    from hprqp import _lib

    ctx = dev.ctx
    xb = _lib.DeviceBuffer.from_array(ctx, x)
    qb = _lib.DeviceBuffer.from_array(ctx, np.full(dev.n, np.nan))
    fn = ctx.lib.hprqp_qop_apply_noalloc if noalloc else ctx.lib.hprqp_qop_apply
    _lib.check(ctx.h, fn(ctx.h, dev.h, xb.h, qb.h))
    return qb.download()


def _operators():
    """(name, Q) pairs, each kind large then small."""
    from hprqp import build_from_mat, instances
    from hprqp.model import DiagonalPlusLowRank_Q_operator

    rng = np.random.default_rng(7)
    return [
        ("qap128_flow_distance", instances.qap(128, seed=0)[0].Q),
        ("qap64_esc64a_full_S_T", build_from_mat(fixture_path("esc64a.mat"), "QAP", verbose=False).Q),
        ("lasso_500x20000", instances.lasso(500, 20_000, mean_nnz=300, seed=2).Q),
        ("lasso_60x900", instances.lasso(60, 900, mean_nnz=40, seed=7).Q),
        ("lowrank_n5000_k40", DiagonalPlusLowRank_Q_operator(rng.uniform(0.1, 1.0, 5000),
                                                             rng.normal(size=(5000, 40)) * 0.1)),
        ("lowrank_n300_k5", DiagonalPlusLowRank_Q_operator(rng.uniform(0.1, 1.0, 300), rng.normal(size=(300, 5)))),
        ("csr_random_n10000", instances.random_sparse_qp(10_000, 5_000, seed=0).Q),
        ("csr_random_n500", instances.random_sparse_qp(500, 100, seed=3).Q),
    ]


def test_noalloc_apply_matches_qop_apply():
    from hprqp import to_gpu

    rng = np.random.default_rng(0)
    for name, Q in _operators():
        n = Q.get_problem_size() if hasattr(Q, "get_problem_size") else Q.shape[0]
        dev = to_gpu(Q, n)
        try:
            x = rng.standard_normal(n)
            want = _apply(dev, x, noalloc=False)
            got = _apply(dev, x, noalloc=True)
            print(f"{name}: n={n} max|Qx|={np.max(np.abs(want)):.3e} identical={np.array_equal(got, want)}",
                  flush=True)
            assert np.all(np.isfinite(got))
            np.testing.assert_array_equal(got, want)
        finally:
            dev.close()


def test_noalloc_refuses_an_operator_beyond_the_array():
    """A QAP operator with nq = 800 needs 2 nq^2 = 1,280,000 doubles of
    temporaries, more than the array's 1,048,576: HPRQP_E_ARG, nothing is
    launched (the output keeps its contents), and the next call works."""
    from hprqp import _lib, instances, to_gpu

    big = instances.qap(800, seed=1)[0].Q
    assert 2 * 800 * 800 > SCRATCH_DOUBLES
    dev = to_gpu(big, 800 * 800)
    ctx = dev.ctx
    try:
        xb = _lib.DeviceBuffer.from_array(ctx, np.ones(dev.n))
        sentinel = np.full(dev.n, 7.0)
        qb = _lib.DeviceBuffer.from_array(ctx, sentinel)
        rc = ctx.lib.hprqp_qop_apply_noalloc(ctx.h, dev.h, xb.h, qb.h)
        msg = ctx.lib.hprqp_last_error(ctx.h).decode()
        print(f"nq=800: rc={rc} msg={msg!r}")
        assert rc == -1 and "exceed the static scratch" in msg
        np.testing.assert_array_equal(qb.download(), sentinel)
    finally:
        dev.close()
    small = instances.qap(40, seed=2)[0].Q
    dev = to_gpu(small, 1600)
    try:
        x = np.random.default_rng(1).standard_normal(1600)
        np.testing.assert_array_equal(_apply(dev, x, noalloc=True), _apply(dev, x, noalloc=False))
    finally:
        dev.close()


def test_noalloc_calls_leave_device_memory_unchanged():
    """After a first call, further calls change nothing in the device's free
    memory (the buffers used here are created before the measurement)."""
    import ctypes as C

    from hprqp import _lib, instances, to_gpu

    hip = C.CDLL("libamdhip64.so")  # the HIP runtime libhprqp links
    hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    hip.hipMemGetInfo.restype = C.c_int

    def free_bytes():
        f, t = C.c_size_t(), C.c_size_t()
        assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    Q = instances.lasso(200, 4000, mean_nnz=60, seed=1).Q
    dev = to_gpu(Q, 4000)
    ctx = dev.ctx
    try:
        xb = _lib.DeviceBuffer.from_array(ctx, np.ones(4000))
        qb = _lib.DeviceBuffer(ctx, 4000)
        _lib.check(ctx.h, ctx.lib.hprqp_qop_apply_noalloc(ctx.h, dev.h, xb.h, qb.h))
        _lib.check(ctx.h, ctx.lib.hprqp_ctx_sync(ctx.h))
        free0 = free_bytes()
        for _ in range(50):
            _lib.check(ctx.h, ctx.lib.hprqp_qop_apply_noalloc(ctx.h, dev.h, xb.h, qb.h))
        _lib.check(ctx.h, ctx.lib.hprqp_ctx_sync(ctx.h))
        free1 = free_bytes()
        print(f"free device memory before {free0} after {free1}")
        assert free1 == free0
    finally:
        dev.close()
