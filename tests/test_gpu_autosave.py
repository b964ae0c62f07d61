"""auto_save (src/algorithm.jl:97-112, 183-301): the best iterate by
max(err_Rp, err_Rd, rel_gap) over the residual checks and the current one,
written with the parameters at print steps and at termination.  The archive
is NumPy .npz with the reference's "group/name" keys (no HDF5 library here)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _load(path):
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autosave_current_and_best(tmp_path, precision):
    from hprqp import HPRQP_parameters, instances, optimize

    model = instances.random_sparse_qp(400, 200, seed=0)
    fn = str(tmp_path / "state")
    res = optimize(model, HPRQP_parameters(verbose=False, auto_save=True, save_filename=fn, precision=precision))
    d = _load(fn + ".npz")
    print(f"{precision}: iter={res.iter} saved current {int(d['current/iteration'])} best "
          f"{int(d['best/iteration'])} kkt {float(d['best/KKTx_and_gap']):.2e}")
    assert res.status == "OPTIMAL"
    # the last write is at termination: the current iterate is the result
    assert int(d["current/iteration"]) == res.iter
    np.testing.assert_array_equal(d["current/x_org"], res.x)
    np.testing.assert_array_equal(d["current/y_org"], res.y)
    np.testing.assert_array_equal(d["current/w_org"], res.w)
    assert abs(float(d["current/primal_obj"]) - res.primal_obj) <= 1e-12 * (1 + abs(res.primal_obj))
    # the best iterate is never worse than the current one and comes no later
    assert float(d["best/KKTx_and_gap"]) <= float(d["current/KKTx_and_gap"])
    assert int(d["best/iteration"]) <= res.iter
    if int(d["best/iteration"]) == res.iter:
        np.testing.assert_array_equal(d["best/x_org"], res.x)
    assert d["best/x_org"].shape == (model.n,) and d["best/y_org"].shape == (model.m,)
    assert float(d["parameters/stoptol"]) == 1e-6 and str(d["parameters/problem_type"]) == "QP"
    assert bool(d["parameters/auto_save"])


def test_autosave_unscaled_z_and_failed_write(tmp_path):
    """With scaling off (col_norm = 1) the saved z is the returned z.  A write
    that fails (unwritable save_filename) is reported and the solve goes on,
    as the reference catches a failed save (src/algorithm.jl:2005-2012): the
    result is the same OPTIMAL solution as without auto_save."""
    from hprqp import HPRQP_parameters, instances, optimize

    model = instances.random_sparse_qp(300, 150, seed=1)
    fn = str(tmp_path / "plain.npz")
    p = HPRQP_parameters(verbose=False, auto_save=True, save_filename=fn, use_Ruiz_scaling=False,
                         use_Pock_Chambolle_scaling=False)
    res = optimize(model, p)
    d = _load(fn)
    np.testing.assert_array_equal(d["current/z_org"], res.z)
    plain = optimize(model, HPRQP_parameters(verbose=False))
    bad = HPRQP_parameters(verbose=True, auto_save=True, save_filename=str(tmp_path / "missing" / "x.npz"))
    res_bad = optimize(model, bad)
    assert res_bad.status == "OPTIMAL" and res_bad.residuals < 1e-6
    assert res_bad.iter == plain.iter
    np.testing.assert_array_equal(res_bad.x, plain.x)
    assert not (tmp_path / "missing").exists()


def test_resume_from_the_saved_best_iterate(tmp_path):
    """The saved best iterate as a warm start (initial_x / initial_y in the
    original space) resumes at the optimum: the restarted solve needs far
    fewer iterations and reaches the same objective."""
    from hprqp import HPRQP_parameters, instances, load_autosave, optimize

    model = instances.random_sparse_qp(400, 200, seed=0)
    fn = str(tmp_path / "resume")
    first = optimize(model, HPRQP_parameters(verbose=False, auto_save=True, save_filename=fn))
    x0, y0 = load_autosave(fn)
    again = optimize(model, HPRQP_parameters(verbose=False, initial_x=x0, initial_y=y0))
    rel = abs(again.primal_obj - first.primal_obj) / (1.0 + abs(first.primal_obj))
    print(f"first iter={first.iter} resumed iter={again.iter} rel obj {rel:.2e}")
    assert again.status == "OPTIMAL" and again.iter < first.iter // 4 and rel < 1e-6
