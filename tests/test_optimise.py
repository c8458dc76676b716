"""CPU checks of the cyclic optimiser's bookkeeping (no GPU).

OptCycleI (src/workproc.f90:4902-5423) rotates the functions so that the
one being optimised sits at the end of the basis: after ReverseFuncOrder
and PermuteFunctions, each step applies PermuteFunctions and
PermuteFunctions2 with NumOfRowsToPermForUnitShift rows.  A stand-in
context with a known separable energy (E = sum_k |p_k - t_k|^2 over the
functions' original numbers, exact gradient) checks that every cycle
optimises FuncBegin..FuncEnd in order, that each step changes only rows
from the PermuteFunctions2 block on, that the basis ends in its input
order, that a run resumes from the history fields, and that the minimiser
(the restated DRMNG, tested against the reference's own in test_drmng.py)
reaches the targets.
"""
from __future__ import annotations

import numpy as np
import pytest

from ecg_nonbo.optimise import BasisOrder, num_rows_to_perm_for_unit_shift, opt_cycle


def test_num_rows_to_perm_for_unit_shift():
    # the table in the reference's comment, src/workproc.f90:3040-3061
    want = [1, 2, 1, 4, 1, 2, 1, 8, 1, 2, 1, 4, 1, 2, 1, 16, 1]
    assert [num_rows_to_perm_for_unit_shift(j) for j in range(1, 18)] == want


def test_permutations():
    o = BasisOrder.identity(np.arange(10, dtype=float).reshape(10, 1))
    o.permute_functions(3, 5)
    assert list(o.func_num) == [1, 2, 6, 7, 8, 9, 10, 3, 4, 5]
    assert list(o.params[:, 0]) == [0, 1, 5, 6, 7, 8, 9, 2, 3, 4]
    o.permute_functions2(2, 3, 6)
    assert list(o.func_num) == [1, 7, 8, 9, 2, 6, 10, 3, 4, 5]
    o.reverse(4, 9)
    assert list(o.func_num) == [1, 7, 8, 4, 3, 10, 6, 2, 9, 5]
    o.sort_decreasing(2, 10)
    assert list(o.func_num) == [1, 10, 9, 8, 7, 6, 5, 4, 3, 2]


class QuadraticContext:
    """Stand-in for EcgContext: E = sum_k |p_k - t_{num(k)}|^2 (fixed shift
    ignored), gradient of the last K - nfru functions; remembers the first
    changed row of each call."""

    def __init__(self, targets, order):
        self.t = targets
        self.order = order
        self.prev = None
        self.first_changed = []

    def set_params(self, p):
        p = np.array(p)
        if self.prev is not None:
            diff = np.nonzero(np.any(p != self.prev, axis=1))[0]
            self.first_changed.append(int(diff[0]) if diff.size else len(p))
        self.prev = p
        self.p = p

    def _E(self):
        t = self.t[self.order.func_num - 1]
        return float(np.sum((self.p - t) ** 2))

    def energy(self, which, sigma, tol, maxiter, want_vector=False):
        info = np.zeros(8)
        info[7] = -1
        return 0, self._E(), (np.ones(len(self.p)) if want_vector else None), info

    def energy_grad(self, nfru, which, sigma, tol, maxiter):
        t = self.t[self.order.func_num - 1]
        g = 2.0 * (self.p - t)[nfru:].reshape(-1)
        info = np.zeros(8)
        info[7] = -1
        return 0, self._E(), g, info

    def overlap_penalty(self, nfru, b, t2):
        return 0.0, None


@pytest.mark.parametrize("K,fb,fe,ncyc", [(12, 1, 12, 2), (20, 5, 17, 1), (9, 3, 9, 3)])
def test_opt_cycle_order_and_convergence(K, fb, fe, ncyc):
    rng = np.random.default_rng(K)
    p0 = rng.normal(size=(K, 3))
    targets = p0 + 0.1 * rng.normal(size=(K, 3))
    order = BasisOrder.identity(p0)
    ctx = QuadraticContext(targets, order)
    seen = []

    def on_step(rec, o):
        seen.append(rec.func)
        # the function being optimised sits at the end of the basis
        assert o.func_num[-1] == rec.func

    recs = opt_cycle(ctx, order, fb, fe, ncyc, max_energy_eval=40, sigma=0.0, on_step=on_step, max_step=1.0)
    assert seen == list(range(fb, fe + 1)) * ncyc
    # input order restored; untouched functions unchanged; optimised ones at their targets
    np.testing.assert_array_equal(order.func_num, np.arange(1, K + 1))
    np.testing.assert_array_equal(order.params[:fb - 1], p0[:fb - 1])
    np.testing.assert_array_equal(order.params[fe:], p0[fe:])
    np.testing.assert_allclose(order.params[fb - 1:fe], targets[fb - 1:fe], atol=1e-6)
    assert all(r.accepted for r in recs)
    assert all(r.energy_after <= r.energy_before for r in recs)


def test_opt_cycle_resumes_from_history():
    K, fb, fe = 10, 1, 10
    rng = np.random.default_rng(3)
    p0 = rng.normal(size=(K, 2))
    targets = p0 + 0.05 * rng.normal(size=(K, 2))
    order = BasisOrder.identity(p0)
    ctx = QuadraticContext(targets, order)
    seen = []
    hist = {"cycles_done": 0, "init_func_at_last_step": 6}   # functions 1..6 done in this cycle
    opt_cycle(ctx, order, fb, fe, 1, 30, 0.0, history=hist, on_step=lambda r, o: seen.append(r.func), max_step=1.0)
    assert seen == [7, 8, 9, 10]
    assert hist["cycles_done"] == 1 and hist["init_func_at_last_step"] == 0
    np.testing.assert_array_equal(order.func_num, np.arange(1, K + 1))
