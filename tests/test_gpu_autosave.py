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


def test_saved_best_state_recomputed_independently(tmp_path):
    """The saved best iterate checked against quantities formed without the
    solver: its iteration is the first minimiser of max(err_Rp, err_Rd,
    rel_gap) over the run's residual history (the reference replaces the
    best only on a strict decrease, src/algorithm.jl:97-112), its recorded
    residuals are that history row, and the primal objective and the primal
    and dual residuals recomputed by scipy from the saved (x, y, z) in the
    original space (scaling off, so the saved z is the result's z) equal the
    recorded ones.  The run is stopped at a residual check whose merit is
    above an earlier one (found from a first run's history), so the best
    iterate is not the current one."""
    import scipy.sparse as sp

    from hprqp import HPRQP_parameters, instances, optimize

    model = instances.random_sparse_qp(400, 200, seed=0)  # its merit rises once, at iteration 510
    kw = dict(verbose=False, use_Ruiz_scaling=False, use_Pock_Chambolle_scaling=False)
    h0 = optimize(model, HPRQP_parameters(**kw)).history
    m0 = np.maximum(np.maximum(h0[:, 1], h0[:, 2]), h0[:, 3])
    # a check whose merit rose above an earlier check's after the start
    above = [r for r in range(2, len(h0)) if m0[r] > np.min(m0[1:r]) and np.argmin(m0[:r]) > 0]
    assert above, "the merit never rose"
    stop = int(h0[above[-1], 0])
    fn = str(tmp_path / "best.npz")
    res = optimize(model, HPRQP_parameters(auto_save=True, save_filename=fn, max_iter=stop, **kw))
    assert res.iter == stop
    d = _load(fn)
    h = res.history
    merit = np.maximum(np.maximum(h[:, 1], h[:, 2]), h[:, 3])
    k = int(np.argmin(merit))
    print(f"status {res.status} iter {res.iter}; {len(h)} residual checks; best at {int(h[k, 0])} "
          f"(merit {merit[k]:.3e}); saved best {int(d['best/iteration'])}")
    assert 0 < int(d["best/iteration"]) == int(h[k, 0]) < stop
    assert float(d["best/err_Rp"]) == h[k, 1] and float(d["best/err_Rd"]) == h[k, 2]
    assert float(d["best/rel_gap"]) == h[k, 3] and float(d["best/primal_obj"]) == h[k, 4]
    pobj, dobj = float(d["best/primal_obj"]), float(d["best/dual_obj"])
    gap = abs(pobj - dobj) / (1.0 + max(abs(pobj), abs(dobj)))
    assert abs(gap - float(d["best/rel_gap"])) <= 1e-12 * (1.0 + gap)
    # recomputed from the saved vectors
    x, y, z = d["best/x_org"], d["best/y_org"], d["best/z_org"]
    Q, A = sp.csr_matrix(model.Q), sp.csr_matrix(model.A)
    Qx, Ax, ATy = Q @ x, A @ x, A.T @ y
    p_re = float(model.c @ x + 0.5 * x @ Qx + model.obj_constant)
    rp = np.max(np.abs(np.clip(Ax, model.AL, model.AU) - Ax)) / (1.0 + max(np.max(np.abs(np.where(
        np.isfinite(model.AL), model.AL, 0.0))), np.max(np.abs(np.where(np.isfinite(model.AU), model.AU, 0.0))),
        np.max(np.abs(Ax))))
    rd = np.max(np.abs(Qx + model.c - ATy - z)) / (1.0 + max(np.max(np.abs(model.c)), np.max(np.abs(ATy)),
                                                          np.max(np.abs(Qx))))
    print(f"recomputed: pobj {p_re:.12e} vs {pobj:.12e}, err_Rp {rp:.6e} vs {float(d['best/err_Rp']):.6e}, "
          f"err_Rd {rd:.6e} vs {float(d['best/err_Rd']):.6e}")
    assert abs(p_re - pobj) <= 1e-10 * (1.0 + abs(pobj))
    assert abs(rp - float(d["best/err_Rp"])) <= 1e-8 * float(d["best/err_Rp"]) + 1e-15
    assert abs(rd - float(d["best/err_Rd"])) <= 1e-8 * float(d["best/err_Rd"]) + 1e-15
