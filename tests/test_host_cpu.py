"""Host logic and the C ABI, without a GPU.

* input validation mirrors the reference's "Input Validation" testset
  (test/runtests.jl:18-125): ErrorException -> ValueError here;
* the MPS reader against the reference fixtures;
* the instance generators' documented shapes;
* libhprqp.so loads and exports every symbol declared in include/hprqp.h
  (no compute calls: there is no GPU in this container).
"""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import ROOT, fixture_path


def test_non_square_Q_rejected():
    from hprqp import build_from_QAbc

    with pytest.raises(ValueError, match="square"):
        build_from_QAbc(np.ones((2, 3)), np.zeros(2), np.ones((1, 2)), [0.0], [1.0], np.zeros(2), np.ones(2),
                        verbose=False)


@pytest.mark.parametrize("which", ["c", "l", "u", "A", "AL", "AU"])
def test_dimension_mismatch_rejected(which):
    from hprqp import build_from_QAbc

    args = dict(Q=np.eye(3), c=np.zeros(3), A=np.ones((2, 3)), AL=np.zeros(2), AU=np.ones(2), l=np.zeros(3),
                u=np.ones(3))
    bad = {"c": np.zeros(2), "l": np.zeros(4), "u": np.ones(2), "A": np.ones((2, 4)), "AL": np.zeros(3),
           "AU": np.ones(1)}
    args[which] = bad[which]
    with pytest.raises(ValueError, match="Dimension mismatch"):
        build_from_QAbc(**args, verbose=False)


def test_infeasible_bounds_rejected():
    from hprqp import build_from_QAbc

    with pytest.raises(ValueError, match="l > u"):
        build_from_QAbc(np.eye(2), np.zeros(2), np.ones((1, 2)), [0.0], [1.0], np.array([1.0, 0.0]),
                        np.array([0.0, 1.0]), verbose=False)
    with pytest.raises(ValueError, match="AL > AU"):
        build_from_QAbc(np.eye(2), np.zeros(2), np.ones((1, 2)), [2.0], [1.0], np.zeros(2), np.ones(2),
                        verbose=False)


def test_nan_rejected():
    from hprqp import build_from_QAbc

    with pytest.raises(ValueError, match="NaN"):
        build_from_QAbc(np.eye(2), np.array([np.nan, 0.0]), np.ones((1, 2)), [0.0], [1.0], np.zeros(2),
                        np.ones(2), verbose=False)
    Q = np.eye(2)
    Q[0, 0] = np.inf
    with pytest.raises(ValueError, match="NaN or Inf"):
        build_from_QAbc(Q, np.zeros(2), np.ones((1, 2)), [0.0], [1.0], np.zeros(2), np.ones(2), verbose=False)


def test_asymmetric_Q_is_symmetrised():
    from hprqp import build_from_QAbc

    Q = np.array([[2.0, 1.0], [0.0, 2.0]])
    with pytest.warns(UserWarning, match="not symmetric"):
        model = build_from_QAbc(Q, np.zeros(2), np.ones((1, 2)), [0.0], [1.0], np.zeros(2), np.ones(2),
                                verbose=False)
    np.testing.assert_allclose(model.Q.toarray(), [[2.0, 0.5], [0.5, 2.0]])


def test_empty_and_free_rows_deleted():
    """src/utils.jl:153-163: rows of A that are all zero, or with AL=-inf and
    AU=+inf, are removed."""
    from hprqp import build_from_QAbc

    A = np.array([[1.0, 1.0], [0.0, 0.0], [1.0, -1.0]])
    model = build_from_QAbc(np.eye(2), np.zeros(2), A, [0.0, 0.0, -np.inf], [1.0, 1.0, np.inf], np.zeros(2),
                            np.ones(2), verbose=False)
    assert model.m == 1
    assert list(model.deleted_rows) == [1, 2]


def test_mps_reader_model_mps():
    from hprqp import build_from_mps

    model = build_from_mps(fixture_path("model.mps"), verbose=False)
    np.testing.assert_allclose(model.Q.toarray(), [[2.0, 0.0], [0.0, 2.0]])
    np.testing.assert_allclose(model.c, [-3.0, -5.0])
    np.testing.assert_allclose(model.A.toarray(), [[1.0, 2.0], [3.0, 1.0]])
    np.testing.assert_allclose(model.AU, [10.0, 12.0])
    assert np.all(np.isneginf(model.AL))
    np.testing.assert_allclose(model.l, [0.0, 0.0])
    assert np.all(np.isposinf(model.u))


def test_mps_reader_aug2d_shapes():
    """AUG2D: 20,200 free variables, 10,000 equalities, diagonal Q with
    19,800 QUADOBJ entries, objective constant +9900 (RHS OBJ -9900)."""
    from hprqp import build_from_mps

    model = build_from_mps(fixture_path("AUG2D.mps"), verbose=False)
    assert (model.n, model.m) == (20200, 10000)
    assert model.obj_constant == 9900.0
    assert np.all(np.isneginf(model.l)) and np.all(np.isposinf(model.u))
    assert np.all(model.AL == model.AU)
    Q = model.Q.tocoo()
    assert np.all(Q.row == Q.col) and Q.nnz == 19800


def test_mps_wrong_extension():
    from hprqp import build_from_mps

    with pytest.raises(ValueError, match="mps"):
        build_from_mps("x.lp")


def test_qap_builder_constraints():
    from hprqp import build_from_ABST

    n = 4
    rng = np.random.default_rng(0)
    model = build_from_ABST(*(rng.standard_normal((n, n)) for _ in range(4)), verbose=False)
    X = rng.uniform(size=(n, n))
    x = X.ravel(order="F")
    Ax = model.A @ x
    np.testing.assert_allclose(Ax[:n], X.sum(axis=1))   # row sums
    np.testing.assert_allclose(Ax[n:], X.sum(axis=0))   # column sums
    assert model.m == 2 * n and model.n == n * n


def test_lasso_builder():
    from hprqp import build_from_Ab_lambda

    A = sp.random(20, 50, density=0.2, random_state=1, format="csr")
    b = np.arange(20.0)
    model = build_from_Ab_lambda(A, b, 0.5, verbose=False)
    np.testing.assert_allclose(model.c, -(A.T @ b))
    assert model.obj_constant == 0.5 * float(b @ b) and model.m == 0 and model.lam == 0.5
    with pytest.raises(AssertionError):
        build_from_Ab_lambda(A, b, 0.0, verbose=False)


def test_grid_generator_shapes():
    from hprqp import instances

    N = 30
    model = instances.grid_qp(N, seed=0)
    assert model.n == N * N and model.m == N * (N // 2)
    assert model.Q.nnz == 5 * N * N - 4 * N
    rows = np.diff(sp.csr_matrix(model.A).indptr)
    assert rows.min() >= 5 and rows.max() <= 8
    np.testing.assert_allclose(np.abs(model.A.data), 1.0)
    assert np.all(model.AL == model.AU)


def test_random_qp_generator_shapes():
    from hprqp import instances

    model = instances.random_sparse_qp(500, 200, seed=0)
    rows = np.diff(sp.csr_matrix(model.A).indptr)
    assert np.all(rows == 5)
    assert np.sum(model.AL == model.AU) == 60
    assert np.all(model.l == -10) and np.all(model.u == 10)


HEADER = ROOT / "include" / "hprqp.h"


def _declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ \*]*?\b(hprqp_[a-z0-9_]+)\s*\(", text,
                                 re.M)))


def test_header_declares_expected_entry_points():
    syms = _declared_symbols()
    for s in ("hprqp_ctx_create", "hprqp_solve", "hprqp_qop_apply", "hprqp_csr_spmv", "hprqp_model_create"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from hprqp import _lib

    if not _lib.LIB_PATH.exists():
        pytest.skip("libhprqp.so not built (run __graft_entry__.build())")
    lib = C.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in _declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes binding covers every declared entry point
    assert set(_declared_symbols()) <= set(_lib.exported_symbols())
    lib.hprqp_abi_version.restype = C.c_int
    assert lib.hprqp_abi_version() == 6


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of hprqp_params, hprqp_result and hprqp_save_state
    have the sizes and field offsets the C header gives (a C program built
    with gcc against include/hprqp.h prints them)."""
    import shutil
    import subprocess

    from hprqp import _lib

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"hprqp_params": _lib.hprqp_params, "hprqp_result": _lib.hprqp_result,
               "hprqp_save_state": _lib.hprqp_save_state, "hprqp_ws_info": _lib.hprqp_ws_info,
               "hprqp_step_params": _lib.hprqp_step_params, "hprqp_kkt_out": _lib.hprqp_kkt_out,
               "hprqp_batch_result": _lib.hprqp_batch_result}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "hprqp.h"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'printf("{name} sizeof %zu\\n", sizeof({name}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{name} {f} %zu\\n", offsetof({name}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for name, cls in structs.items():
        assert got[(name, "sizeof")] == C.sizeof(cls), name
        for f, _ in cls._fields_:
            assert got[(name, f)] == getattr(cls, f).offset, (name, f)


def test_params_default_matches_reference_defaults():
    """HPRQP_parameters() defaults (src/structs.jl:274)."""
    from hprqp import HPRQP_parameters, _lib

    if not _lib.LIB_PATH.exists():
        pytest.skip("libhprqp.so not built")
    p = HPRQP_parameters().to_c()
    assert p.stoptol == 1e-6 and p.sigma == -1 and p.max_iter == 2**31 - 1
    assert p.time_limit == 3600.0 and p.eig_factor == 1.05 and p.check_iter == 100
    assert p.print_frequency == -1 and p.use_Ruiz_scaling == 1 and p.use_Pock_Chambolle_scaling == 1
    assert p.use_bc_scaling == 0 and p.problem_type == 0


def test_no_gpu_raises_instead_of_falling_back():
    """The product path has no CPU fallback: creating a context without a GPU
    fails loudly."""
    from hprqp import _lib

    if not _lib.LIB_PATH.exists():
        pytest.skip("libhprqp.so not built")
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    with pytest.raises(_lib.HPRQPError):
        _lib.Context(0)


def test_mps_ranges_bound_types_and_objective_constant():
    """tests/golden/bound_edges.mps (written for this test): every RANGES sign per row type, every bound type (UP below zero on a
    default lower bound, LO, FX, FR, MI, PL after UP, BV, default), the
    objective-row RHS as a constant, and the lower-triangle QUADOBJ, with
    the expected arrays written out by hand."""
    from hprqp.mps import read_mps

    Q, c, A, lcon, ucon, lvar, uvar, c0 = read_mps(fixture_path("bound_edges.mps"))
    inf = np.inf
    np.testing.assert_array_equal(c, [1, -1, 0, 0.5, 0, 2, 0, 1])
    assert c0 == -7.0
    np.testing.assert_array_equal(A.toarray(), [[1, 1, 0, 0, 0, 1, 0, 0],
                                                [1, 0, 2, 0, 0, 0, 1, 0],
                                                [0, 1, 0, -1, 0, 0, 0, 1],
                                                [0, 0, 1, 0, 1, 0, 0, 0]])
    # E with R > 0: [rhs, rhs+|R|]; L: [rhs-|R|, rhs]; G: [rhs, rhs+|R|]; E with R < 0: [rhs-|R|, rhs]
    np.testing.assert_array_equal(lcon, [4, 2, 1, -2])
    np.testing.assert_array_equal(ucon, [6, 5, 2.5, 0])
    np.testing.assert_array_equal(lvar, [-inf, -3, 2.5, -inf, -inf, 0, 0, 0])
    np.testing.assert_array_equal(uvar, [-1, inf, 2.5, inf, inf, inf, 1, inf])
    Qd = np.zeros((8, 8))
    Qd[:2, :2] = [[2, 1], [1, 2]]
    np.testing.assert_array_equal(Q.toarray(), Qd)
