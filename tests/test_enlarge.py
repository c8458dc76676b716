"""CPU checks of the basis-enlargement loop and its candidate generator,
of the simultaneous optimiser and of the BBOP runner (no GPU).

GenerateTrialParam (src/workproc.f90:1375-1487) is pinned by the
distributions its text fixes: uniform [-0.5, 0.5) parameters for an empty
basis; method 1 with probability p1, every parameter scaled by an
independent 1 + s1 N(0,1); method 2 with one common factor outside
(0.8, 1.2); consecutive prototypes from a uniform k; and, with fewer
functions than requested, the reference's method-2 factor equal to the
uniform number that picked the prototype, so (k - 1)/K <= factor < k/K.

BasisEnlI's loop (:3734-4379) runs on a stand-in context with a separable
energy (tests/standin_ctx.py): the basis grows to Kstop in steps of Kstep,
the kept candidate is the best trial, the failure limits and the retry on
a failed overlap or coefficient check behave as in the reference, and the
batched screening of trials keeps the same candidate as the exact loop.
"""
from __future__ import annotations

import numpy as np
import pytest

from ecg_nonbo.enlarge import BasisEnlError, basis_enl, generate_trial_param
from ecg_nonbo.optimise import BasisOrder, full_opt1
from standin_ctx import SeparableContext


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


# ----------------------------------------------------------- the generator
def test_generator_empty_basis_uniform():
    x, k, m = generate_trial_param(_rng(1), np.zeros((0, 6)), 4000, 0.5, 0.1, 0.3)
    assert (k, m) == (0, 0) and x.shape == (4000, 6)
    assert x.min() >= -0.5 and x.max() < 0.5
    # mean 0 and variance 1/12 of U[-0.5, 0.5) (24000 draws: 5 sigma bounds)
    assert abs(x.mean()) < 5 * np.sqrt(1 / 12 / x.size)
    assert abs(x.var() - 1 / 12) < 5 * np.sqrt(1 / 180 / x.size)


def test_generator_methods_and_prototypes():
    rng = _rng(2)
    K, npt, nfun, p1, s1, s2 = 30, 6, 3, 0.3, 0.05, 0.4
    params = 0.5 + rng.random((K, npt))
    n1 = 0
    ratios1 = []
    N = 3000
    for _ in range(N):
        x, k, m = generate_trial_param(rng, params, nfun, p1, s1, s2)
        assert 1 <= k <= K - nfun + 1
        ratio = x / params[k - 1:k - 1 + nfun]
        if m == 1:
            n1 += 1
            ratios1.append(ratio.ravel() - 1.0)
        else:
            assert m == 2
            # one common factor, outside (0.8, 1.2)
            np.testing.assert_allclose(ratio, ratio[0, 0], rtol=1e-14)
            assert not (0.8 < abs(ratio[0, 0]) < 1.2)
    # method 1 with probability p1 (binomial, 5 sigma)
    assert abs(n1 / N - p1) < 5 * np.sqrt(p1 * (1 - p1) / N)
    r = np.concatenate(ratios1)
    assert abs(r.mean()) < 5 * s1 / np.sqrt(r.size)
    assert abs(r.std() / s1 - 1.0) < 0.05


def test_generator_fewer_functions_than_requested():
    rng = _rng(3)
    K, npt = 4, 6
    params = 1.0 + rng.random((K, npt))
    seen = set()
    for _ in range(400):
        x, k, m = generate_trial_param(rng, params, 6, 0.0, 0.1, 0.5)   # p1 = 0: always method 2
        assert m == 2 and x.shape == (6, npt)
        # the last trial function's factor is the uniform number that picked its prototype
        f = x[-1] / params[k - 1]
        np.testing.assert_allclose(f, f[0], rtol=1e-14)
        assert (k - 1) / K <= f[0] < k / K
        # every trial function is a multiple of one of the functions
        for row in x:
            q = row / params
            assert np.any(np.all(np.abs(q - q[:, :1]) <= 1e-14 * np.abs(q[:, :1]) + 1e-300, axis=1))
        seen.add(k)
    assert seen == {1, 2, 3, 4}


def test_generator_method2_factor_loop_bounded():
    # a factor spread that can never leave (0.8, 1.2) is reported instead of looping for ever
    with pytest.raises(BasisEnlError):
        generate_trial_param(_rng(4), np.ones((3, 2)), 1, 0.0, 0.1, 0.0)


# ------------------------------------------------------------- BasisEnl
def _grow(ctx, K0, k_stop, k_step, seed=5, **kw):
    rng = _rng(seed)
    t = ctx.t
    p0 = t + 0.8 * rng.normal(size=(K0, t.size))
    order = BasisOrder.identity(p0)
    hist = [{"energy": 0.0, "cycles_done": 2, "init_func_at_last_step": 0, "num_energy_evals_full_opt": 0}
            for _ in range(K0)]
    recs = basis_enl(ctx, order, k_stop, k_step, kw.pop("n_trials", 12), kw.pop("max_energy_eval", 15), -1.0,
                     rng, history=hist, max_step=0.5, **kw)
    return order, hist, recs


@pytest.mark.parametrize("K0,k_stop,k_step", [(3, 10, 1), (4, 11, 3), (0, 5, 2)])
def test_basis_enl_grows_and_keeps_best(K0, k_stop, k_step):
    ctx = SeparableContext(np.array([0.3, -0.2, 0.5, 0.1]))
    order, hist, recs = _grow(ctx, K0, k_stop, k_step)
    assert len(order.func_num) == k_stop
    np.testing.assert_array_equal(order.func_num, np.arange(1, k_stop + 1))
    assert len(hist) == k_stop and all(h["cycles_done"] == 0 for h in hist[K0:])
    sizes = [r.nfo for r in recs]
    assert sum(sizes) == k_stop - K0 and all(s == k_step for s in sizes[:-1])
    for r in recs:
        assert r.energy_after <= r.energy_trials < r.energy_before
        assert r.attempts == 1 and r.checks_passed
    # energies of the new functions' history entries are the step energies
    k = K0
    for r in recs:
        for j in range(k, k + r.nfo):
            assert hist[j]["energy"] == r.energy_after
        k += r.nfo
    # final energy is that of the basis
    assert recs[-1].energy_after == pytest.approx(float(np.sum(ctx._phi(order.params))), abs=1e-14)


def test_basis_enl_kept_trial_is_the_best_one():
    ctx = SeparableContext(np.array([0.3, -0.2]))
    order, _, recs = _grow(ctx, 3, 4, 1, n_trials=20, optimise=False)
    # energies of the trial calls at K = 4: the first 20 after the initial energy
    trial_E = [e for kind, K, e in ctx.log if kind == "E" and K == 4][:20]
    assert recs[0].energy_trials == min(trial_E)
    assert recs[0].energy_after == recs[0].energy_trials     # no optimisation: the kept trial itself


def test_basis_enl_screen_matches_exact():
    t = np.array([0.3, -0.2, 0.5])
    a = SeparableContext(t)
    b = SeparableContext(t)
    oa, _, ra = _grow(a, 5, 9, 1, seed=7)
    ob, _, rb = _grow(b, 5, 9, 1, seed=7, trials="screen")
    assert b.screened == 4 * 12
    np.testing.assert_allclose(ob.params, oa.params, rtol=0, atol=1e-14)
    assert [r.prototype for r in ra] == [r.prototype for r in rb]


def test_basis_enl_trial_failure_limit():
    ctx = SeparableContext(np.array([0.3, -0.2]))
    ctx.fail_calls = {1, 2}         # 2 of 10 trials fail: above 15 %
    with pytest.raises(BasisEnlError, match="limit of eigenvalue problem failures"):
        _grow(ctx, 3, 4, 1, n_trials=10)
    ctx = SeparableContext(np.array([0.3, -0.2]))
    ctx.fail_calls = {1}            # 1 of 10: allowed
    _, _, recs = _grow(ctx, 3, 4, 1, n_trials=10)
    assert recs[0].trial_failures == 1


def test_basis_enl_no_improvement_stops():
    ctx = SeparableContext(np.array([0.0, 0.0]))
    ctx._phi = lambda p: np.zeros(np.shape(p)[:-1])     # no trial can lower the energy
    with pytest.raises(BasisEnlError, match="did not result in any energy improvement"):
        _grow(ctx, 2, 3, 1)


def test_basis_enl_retries_after_bad_coefficients():
    # As in the reference, a retry's trials must beat the energy of the
    # rejected attempt (Glob_CurrEnergy is not reset, src/workproc.f90:4047-4060);
    # an energy that drifts down with every call makes that certain here.
    t = np.array([0.3, -0.2])
    ctx = SeparableContext(t)
    ctx.drift = 1.0
    # coefficients too large at the first check (the first attempt's), fine after
    checks = []

    def coeff(ncalls):
        checks.append(ncalls)
        return 50.0 if len(checks) == 1 else 1.0

    ctx.coeff = coeff
    order, _, recs = _grow(ctx, 3, 4, 1, n_trials=6, max_energy_eval=6, lin_coeff_threshold=10.0)
    assert recs[0].attempts == 2 and recs[0].checks_passed
    # with the limit reached the last attempt is kept
    ctx = SeparableContext(t, coeff=50.0)
    ctx.drift = 1.0
    order, _, recs = _grow(ctx, 3, 4, 1, n_trials=6, max_energy_eval=4, lin_coeff_threshold=10.0, bad_lim=3)
    assert recs[0].attempts == 3 and not recs[0].checks_passed and len(order.func_num) == 4


def test_basis_enl_retry_must_beat_rejected_attempt():
    # without the drift the optimised (rejected) function cannot be beaten by
    # new trials around the old functions: the reference stops there
    ctx = SeparableContext(np.array([0.3, -0.2]), coeff=50.0)
    with pytest.raises(BasisEnlError, match="did not result in any energy improvement"):
        _grow(ctx, 3, 4, 1, n_trials=6, max_energy_eval=30, lin_coeff_threshold=10.0)


def test_basis_enl_overlap_check():
    t = np.array([0.3, -0.2])
    ctx = SeparableContext(t)
    ctx.drift = 1.0      # lets the retry's trials beat the rejected attempt (see above)
    # all functions at t and trials within about 1 % of them: every new function overlaps (S^2 ~ 1)
    order = BasisOrder.identity(np.tile(t, (3, 1)))
    recs = basis_enl(ctx, order, 4, 1, 8, 0, -1.0, _rng(9), generator_param=(1.0, 0.01, 0.3),
                     overlap_threshold=0.9, bad_lim=2, optimise=False)
    assert recs[0].attempts == 2 and not recs[0].checks_passed
    # far-apart functions pass at the first attempt
    ctx = SeparableContext(t)
    order = BasisOrder.identity(t + np.array([[0.0, 0.0], [3.0, 0.0], [0.0, 3.0]]))
    recs = basis_enl(ctx, order, 4, 1, 8, 0, -1.0, _rng(9), generator_param=(0.0, 0.01, 1.5),
                     overlap_threshold=0.9, bad_lim=2, optimise=False)
    P, _ = ctx.overlap_penalty(3, 1.0, 0.81)
    assert recs[0].checks_passed == (P <= 0.0)


# ------------------------------------------------------------- FullOpt1
def test_full_opt1_subrange_order_and_penalty():
    t = np.array([0.3, -0.2, 0.1])
    rng = _rng(11)
    K = 9
    p0 = t + rng.normal(size=(K, 3))
    ctx = SeparableContext(t)
    order = BasisOrder.identity(p0)
    h = {"energy": 0.0, "cycles_done": 1, "init_func_at_last_step": 0, "num_energy_evals_full_opt": 7}
    res = full_opt1(ctx, order, 3, 6, 200, -1.0, overlap_threshold=0.5, max_overlap_penalty=1.0, history=h,
                    max_step=0.5)
    assert 1 <= res.code <= 10
    # input order back; functions outside 3..6 untouched
    np.testing.assert_array_equal(order.func_num, np.arange(1, K + 1))
    np.testing.assert_array_equal(order.params[:2], p0[:2])
    np.testing.assert_array_equal(order.params[6:], p0[6:])
    assert res.evaluations <= 200 and h["num_energy_evals_full_opt"] == 7 + res.evaluations
    # the penalty keeps the optimised functions apart: no pair among them is above threshold 0.5
    q = order.params[2:6]
    d2 = np.sum((q[:, None] - q[None]) ** 2, axis=-1)
    s2 = np.exp(-d2[np.triu_indices(4, 1)])
    assert np.all(s2 <= 0.25 + 0.05)
    assert res.energy_after == pytest.approx(h["energy"]) and res.energy_after < res.energy_before
    assert ctx.pen[0] is False              # switched off on return


def test_full_opt1_without_penalty_reaches_optimum():
    t = np.array([0.3, -0.2])
    rng = _rng(12)
    p0 = t + 0.5 * rng.normal(size=(5, 2))
    ctx = SeparableContext(t)
    order = BasisOrder.identity(p0)
    res = full_opt1(ctx, order, 1, 5, 400, -1.0, max_step=0.5)
    np.testing.assert_allclose(order.params, np.tile(t, (5, 1)), atol=1e-5)
    assert res.penalty == 0.0 and res.energy_after == pytest.approx(-5.0, abs=1e-9)


# ------------------------------------------------------------- BBOP runner
def _fixture(golden_dir, K=None):
    import os
    from ecg_nonbo import read_inout
    f = read_inout(os.path.join(golden_dir, "hdplus", "inout_1_500.txt"))
    if K is not None:
        f.params = f.params[:K].copy()
        f.history = f.history[:K]
        f.history_energy = f.history_energy[:K]
        f.basis_size = K
    return f


def test_run_bbop_completed_file_is_a_no_op(golden_dir):
    # every step of the reference's finished HD+ file fails its guard: the
    # last BASIS_ENL stops at K = 500, the last OPT_CYCLE has its 10 cycles
    from ecg_nonbo.bbop import run_bbop
    f = _fixture(golden_dir)
    p0 = f.params.copy()
    ctx = SeparableContext(np.mean(f.params, axis=0))
    reps = run_bbop(ctx, f, _rng(1))
    assert [r.status for r in reps] == ["not run"] * 10
    assert ctx.calls == 0
    np.testing.assert_array_equal(f.params, p0)


def test_run_bbop_steps(golden_dir, tmp_path):
    from ecg_nonbo import BbopStep, read_inout
    from ecg_nonbo.bbop import run_bbop
    f = _fixture(golden_dir, K=20)
    f.current_energy = -0.5
    ctx = SeparableContext(np.mean(f.params, axis=0))
    save = str(tmp_path / "saved.txt")
    f.bbop = [
        BbopStep("BASIS_ENL", "I", [1, 24, 2, 10, 5], [-1.0, -1.0]),
        BbopStep("OPT_CYCLE", "I", [24, 21, 24, 1, 1, 1, 4], [-1.0, -1.0]),
        BbopStep("OPT_CYCLE", "I", [23, 1, 23, 1, 1, 1, 4], [-1.0, -1.0]),        # wrong K: not run
        BbopStep("FULL_OPT1", "G", [24, 22, 24, 6, 0, 0], [1.0, 1.0], ["none"]),    # no penalty
        BbopStep("ELIM_LCFN", "G", [24], [0.5], ["elim.txt"]),
        BbopStep("SAVE_FILE", "", [24], [], [save]),
    ]
    reps = run_bbop(ctx, f, _rng(2), path=str(tmp_path / "inout.txt"))
    assert [r.status for r in reps] == ["done", "done", "not run", "done", "skipped", "done"]
    assert [r.basis_size for r in reps] == [24] * 6
    # the energy is that of the final basis; it never rose
    E = [r.energy for r in reps]
    assert all(b <= a + 1e-15 for a, b in zip(E, E[1:]))
    assert E[-1] == pytest.approx(float(np.sum(ctx._phi(f.params))), abs=1e-12)
    assert f.history[23].cycles_done == 1 and f.history[23].num_energy_evals_full_opt > 0
    for path in (save, str(tmp_path / "inout.txt")):
        g = read_inout(path)
        assert g.basis_size == 24 and len(g.history) == 24
        np.testing.assert_allclose(g.params, f.params, rtol=1e-15)
        assert g.current_energy == pytest.approx(f.current_energy, rel=1e-15)
