"""The diagonal form: a matrix whose nonzeros lie on at most 8 diagonals j =
floor(i ncol / nrow) + offset is kept as one array per diagonal, indexed by
row (k_dia), forced with spmv_lanes = -2000 (Q must have one; A and A^T take
theirs where they have one) or taken by the autotuner where it times faster.
On an even grid Q (the 5-point Laplacian), A (the pair operator: 8 offsets
from column 2i) and A^T (5 offsets from column floor(i / 2)) all have one.

* Bitwise: the four fused half-steps write bitwise the same vectors with Q,
  A and A^T in the diagonal form as in the CSR-stream form of 256 rows per
  block, from the same uploaded state (both add a row's products in
  ascending column order; the padding adds exact zeros).
* Solves with the form forced match the oracle (iterations within 10 %,
  objective within 1e-6 relative, the power method's stop iterations within
  one) and report the form for exactly the matrices that have one: even and
  odd grids, a banded QP (offsets 0, +-1, +-36, +-37), a 3-D 7-point
  Laplacian and a tridiagonal Q with a restriction operator; the reduced
  precisions reach 1e-6.
* A Q on more than 8 diagonals is refused when the form is forced and keeps
  its CSR form under the autotuner."""
import numpy as np
import pytest
import scipy.sparse as sp

from test_gpu_parity import _check, _solve_both

pytestmark = pytest.mark.gpu

VEC = {"x_bar": 0, "y_bar": 1, "z_bar": 2, "w_bar": 3, "ATdy": 4, "x": 7, "w": 8, "y": 9, "ATy": 10, "Qw": 11,
       "x_hat": 12, "dx": 13, "dw": 14, "dy": 15, "s": 16, "ATy_bar": 17, "tempv": 18, "Qw_bar": 19,
       "last_x": 20, "last_y": 21, "last_w": 22, "last_ATy": 23}
M_VECS = {"y", "y_bar", "dy", "s", "last_y"}


def _banded_qp(n=3000, m=700, seed=5):
    from hprqp import build_from_QAbc

    rng = np.random.default_rng(seed)
    B = sp.diags([rng.uniform(0.5, 1.5, n), rng.uniform(-0.5, 0.5, n - 1), rng.uniform(-0.3, 0.3, n - 37)],
                 [0, 1, 37], format="csr")
    Q = (B.T @ B + 1e-2 * sp.eye(n)).tocsr()  # offsets 0, +-1, +-36, +-37
    A = sp.random(m, n, density=4.0 / n, random_state=seed, format="csr") + sp.eye(m, n)
    b = A @ rng.uniform(-1, 1, n)
    return build_from_QAbc(Q, rng.standard_normal(n), A.tocsr(), b - 0.5, b + 0.5, np.full(n, -2.0),
                           np.full(n, 2.0), verbose=False)


def _grid3d_qp(N=24, seed=6):
    """7-point Laplacian on an N^3 grid (7 diagonals) with a random sparse A."""
    from hprqp import build_from_QAbc

    rng = np.random.default_rng(seed)
    n = N ** 3
    L1 = sp.diags([np.full(N - 1, -1.0), np.full(N, 2.0), np.full(N - 1, -1.0)], [-1, 0, 1])
    Id = sp.eye(N)
    Q = (sp.kron(sp.kron(L1, Id), Id) + sp.kron(sp.kron(Id, L1), Id) + sp.kron(sp.kron(Id, Id), L1)
         + 1e-2 * sp.eye(n)).tocsr()
    m = n // 3
    A = (sp.random(m, n, density=4.0 / n, random_state=seed, format="csr") + sp.eye(m, n)).tocsr()
    b = A @ rng.uniform(0, 1, n)
    return build_from_QAbc(Q, rng.standard_normal(n), A, b, b + 0.1, np.zeros(n), np.full(n, 2.0), verbose=False)


def _restriction_qp(n=40000, seed=7):
    """Tridiagonal Q and a restriction operator A (row i: 1, 2, 1 at columns
    2i-1, 2i, 2i+1): A has 3 diagonals from column 2i, A^T 2 from floor(i/2)."""
    from hprqp import build_from_QAbc

    rng = np.random.default_rng(seed)
    m = n // 2
    Q = (sp.diags([np.full(n - 1, -1.0), np.full(n, 2.05), np.full(n - 1, -1.0)], [-1, 0, 1])).tocsr()
    r = np.arange(m)
    rows, cols = np.concatenate([r, r, r]), np.concatenate([2 * r - 1, 2 * r, 2 * r + 1])
    vals = np.concatenate([np.ones(m), 2 * np.ones(m), np.ones(m)])
    keep = (cols >= 0) & (cols < n)
    A = sp.csr_matrix((vals[keep], (rows[keep], cols[keep])), shape=(m, n))
    b = A @ rng.uniform(-1, 1, n)
    return build_from_QAbc(Q, rng.standard_normal(n), A, b - 0.2, b + 0.2, np.full(n, -3.0), np.full(n, 3.0),
                           verbose=False)


def test_half_steps_bitwise_equal_to_csr_stream():
    from hprqp import HPRQP_parameters, instances
    from hprqp.solver import DeviceModel
    from hprqp.steploop import Workspace

    model = instances.grid_qp(120, seed=3)
    n, m = model.n, model.m
    dm = DeviceModel(model)
    ws = [Workspace(dm, HPRQP_parameters(verbose=False, spmv_lanes=lanes)) for lanes in (-256, -2000)]
    try:
        rng = np.random.default_rng(8)
        state = {k: rng.uniform(-1.0, 1.0, m if k in M_VECS else n)
                 for k in ("x", "w", "ATy", "last_x", "last_w", "last_ATy", "w_bar", "y", "last_y")}
        for w in ws:
            for k, v in state.items():
                w.upload(VEC[k], v)
        for name, outs in (("zxw1", ("x", "x_hat", "w_bar", "Qw", "dx", "x_bar", "z_bar")),
                           ("tempv", ("tempv",)),
                           ("y", ("y", "y_bar", "s", "dy")),
                           ("w2", ("w", "ATy", "w_bar", "dw", "ATdy", "ATy_bar"))):
            for w in ws:
                w.step(name, 0.7, 0.25, 2.0, 8.0, True)
            for k in outs:
                a, b = (w.download(VEC[k], m if k in M_VECS else n) for w in ws)
                print(f"{name}: {k} max |diff| {np.max(np.abs(a - b)):.3e}")
                assert np.array_equal(a, b), (name, k)
            # the product entered: Q w is not zero, tempv = x_hat + sigma (Q w - Q w_bar) is not x_hat
            if name == "zxw1":
                assert np.any(ws[0].download(VEC["Qw"], n) != 0.0)
            elif name == "tempv":
                assert np.any(ws[0].download(VEC["tempv"], n) != ws[0].download(VEC["x_hat"], n))
    finally:
        for w in ws:
            w.close()
        dm.close()


# the shapes each case must report: whose patterns have a diagonal form
SHAPES = {"grid": (True, True, True), "grid_odd": (True, True, False), "banded": (True, False, False),
          "grid3d": (True, False, False), "restriction": (True, True, True)}


@pytest.mark.parametrize("case", list(SHAPES))
def test_forced_diagonal_form_matches_oracle(case):
    from hprqp import instances

    model = {"grid": lambda: instances.grid_qp(90, seed=4), "grid_odd": lambda: instances.grid_qp(91, seed=4),
             "banded": _banded_qp, "grid3d": _grid3d_qp, "restriction": _restriction_qp}[case]()
    res, ref = _solve_both(model, spmv_lanes=-2000)
    print(f"{case}: shapes {res.lanes}")
    # an odd grid's A^T needs 6 diagonals at 1.52 entries per nonzero (over the padding limit)
    assert tuple(lanes == -2000 for lanes in res.lanes) == SHAPES[case]
    _check(res, ref)


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_forced_diagonal_form_reduced_precision(precision):
    from hprqp import HPRQP_parameters, instances, optimize
    from oracle import oracle

    model = instances.grid_qp(90, seed=4)
    res = optimize(model, HPRQP_parameters(verbose=False, precision=precision, spmv_lanes=-2000))
    ref = oracle.solve(model)
    rel = abs(res.primal_obj - ref.primal_obj) / (1.0 + abs(ref.primal_obj))
    print(f"{precision}: {res.status} iter={res.iter} polish={res.polish_iter} kkt={res.residuals:.2e} "
          f"rel={rel:.2e} (oracle iter={ref.iter}) shapes {res.lanes}")
    assert res.lanes[0] == -2000
    assert res.status == "OPTIMAL" and res.residuals < 1e-6 and rel < 1e-6


def test_more_than_eight_diagonals():
    from hprqp import HPRQP_parameters, instances, optimize

    model = instances.random_sparse_qp(2000, 800, seed=2)
    with pytest.raises(RuntimeError, match="more than 8 diagonals"):
        optimize(model, HPRQP_parameters(verbose=False, spmv_lanes=-2000))
    res = optimize(model, HPRQP_parameters(verbose=False))
    assert res.status == "OPTIMAL" and res.lanes[0] != -2000


def test_autotuned_grid():
    """grid 300 x 300 (450k nonzeros in Q): the autotuner times the diagonal
    form against Q's CSR shape; whichever it keeps, the solve matches the
    oracle."""
    from hprqp import instances

    model = instances.grid_qp(300, seed=1)
    res, ref = _solve_both(model)
    print(f"autotuned shapes {res.lanes}")
    _check(res, ref)
