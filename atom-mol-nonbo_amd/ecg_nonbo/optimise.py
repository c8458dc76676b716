"""Cyclic optimisation of basis functions one (or a few) at a time, the
loop of the reference's OptCycleI (src/workproc.f90:4902-5423), over the
device energy path.

What follows the reference: the function order bookkeeping
(ReverseFuncOrder, PermuteFunctions, PermuteFunctions2,
SortBasisFuncAndMatElem's parameter part, NumOfRowsToPermForUnitShift;
src/workproc.f90:2414-2487, 2731-2760, 2857-2982, 3040-3078), which brings
each function to the end of the basis with few rows changed, the energy
calls at a fixed shift (EnergyIA(Nmin, K) / EnergyIB, which the device
context turns into reassembly and refactorisation from the first changed
row), the acceptance tests on overlaps and linear coefficients, and the
history fields (CyclesDone, InitFuncAtLastStep) with which the reference
resumes a cycle.

What does not: the minimiser.  The reference drives PORT's DRMNG (a
reverse-communication secant method with a trust region); here a compact
BFGS with a bounded first step and a backtracking line search stands in,
so optimisation trajectories are not expected to match the reference's
(they depend on the minimiser's internals), only its contract: at most
MaxEnergyEval energy evaluations per step, the best point found is kept.
"""
from __future__ import annotations

import dataclasses
from typing import Callable, List, Optional

import numpy as np

__all__ = ["BasisOrder", "num_rows_to_perm_for_unit_shift", "bfgs_minimise", "opt_cycle", "OptStepRecord"]


def num_rows_to_perm_for_unit_shift(j: int) -> int:
    """1, 2, 1, 4, 1, 2, 1, 8, ...: the largest power of two dividing j
    (NumOfRowsToPermForUnitShift, src/workproc.f90:3040-3078)."""
    k = 1
    while j % k == 0:
        k *= 2
    return k // 2


@dataclasses.dataclass
class BasisOrder:
    """Glob_NonlinParam (here rows, K x npt) and Glob_FuncNum (1-based
    numbers of the functions in their original order); all positions below
    are 1-based and inclusive, as in the reference."""
    params: np.ndarray
    func_num: np.ndarray

    @classmethod
    def identity(cls, params):
        return cls(np.array(params, dtype=np.float64, copy=True), np.arange(1, len(params) + 1))

    def permute_functions(self, fb: int, fe: int) -> None:
        """Functions fb..fe go to the end of the basis, the ones after them
        move up (PermuteFunctions, src/workproc.f90:2414-2447)."""
        K = len(self.func_num)
        idx = np.r_[np.arange(0, fb - 1), np.arange(fe, K), np.arange(fb - 1, fe)]
        self.params = self.params[idx].copy()
        self.func_num = self.func_num[idx].copy()

    def permute_functions2(self, fb1: int, fe1: int, fe2: int) -> None:
        """The block fb1..fe1 exchanges places with fe1+1..fe2
        (PermuteFunctions2, src/workproc.f90:2451-2484)."""
        K = len(self.func_num)
        idx = np.r_[np.arange(0, fb1 - 1), np.arange(fe1, fe2), np.arange(fb1 - 1, fe1), np.arange(fe2, K)]
        self.params = self.params[idx].copy()
        self.func_num = self.func_num[idx].copy()

    def reverse(self, fb: int, fe: int) -> None:
        """ReverseFuncOrder (src/workproc.f90:2731-2760)."""
        self.params[fb - 1:fe] = self.params[fb - 1:fe][::-1].copy()
        self.func_num[fb - 1:fe] = self.func_num[fb - 1:fe][::-1].copy()

    def sort_decreasing(self, fb: int, fe: int) -> None:
        """Functions fb..fe in decreasing Glob_FuncNum order
        (SortBasisFuncAndMatElem's parameter part, src/workproc.f90:2857-2885);
        the numbers must be consecutive."""
        seg = slice(fb - 1, fe)
        order = np.argsort(-self.func_num[seg], kind="stable")
        self.params[seg] = self.params[seg][order].copy()
        self.func_num[seg] = self.func_num[seg][order].copy()


def bfgs_minimise(fg: Callable[[np.ndarray], tuple], f_only: Callable[[np.ndarray], float], x0: np.ndarray,
                  max_evals: int, max_step: float, scale: np.ndarray):
    """Minimise with at most max_evals energy evaluations (gradient calls
    count as evaluations, as the reference counts EnergyIA and EnergyIB
    calls against its limit).  Variables are scaled by `scale` (the role of
    DRMNG's D vector); the first step's scaled length is at most max_step.
    Returns (x_best, f_best, evaluations)."""
    x = np.array(x0, dtype=np.float64)
    f, g = fg(x)
    nev = 1
    best = (x.copy(), f)
    n = x.size
    Hinv = np.eye(n)
    gs = g / scale
    first = True
    while nev < max_evals:
        p = -(Hinv @ gs)
        if first:
            nrm = np.linalg.norm(p)
            if nrm > max_step:
                p *= max_step / nrm
            first = False
        slope = gs @ p
        if slope >= 0.0:   # not a descent direction: restart from steepest descent
            Hinv = np.eye(n)
            p = -gs * min(1.0, max_step / max(np.linalg.norm(gs), 1e-300))
            slope = gs @ p
            if slope >= 0.0:
                break
        alpha = 1.0
        accepted = False
        while nev < max_evals:
            xn = x + alpha * p / scale
            fn = f_only(xn)
            nev += 1
            if np.isfinite(fn) and fn < best[1]:
                best = (xn.copy(), fn)
            if np.isfinite(fn) and fn <= f + 1e-4 * alpha * slope:
                accepted = True
                break
            alpha *= 0.25
            if alpha < 1e-8:
                break
        if not accepted or nev >= max_evals:
            break
        fn, gn = fg(xn)
        nev += 1
        if fn < best[1]:
            best = (xn.copy(), fn)
        gsn = gn / scale
        s = alpha * p
        yv = gsn - gs
        sy = s @ yv
        if sy > 1e-300:
            rho = 1.0 / sy
            V = np.eye(n) - rho * np.outer(s, yv)
            Hinv = V @ Hinv @ V.T + rho * np.outer(s, s)
        x, f, gs = xn, fn, gsn
        if np.linalg.norm(gs) * np.linalg.norm(s) < 1e-15 * max(1.0, abs(f)):
            break
    return best[0], best[1], nev


@dataclasses.dataclass
class OptStepRecord:
    func: int                 # function number (Glob_FuncNum) optimised
    energy_before: float
    energy_after: float
    evaluations: int
    accepted: bool
    tail_rows: List[int]      # rows refactored by the energy calls of the step (-1: full)


def opt_cycle(ctx, order: BasisOrder, func_begin: int, func_end: int, num_cycles: int, max_energy_eval: int,
              sigma: float, overlap_threshold: float = -1.0, lin_coeff_threshold: float = -1.0,
              num_func_to_opt: int = 1, num_func_to_shift: int = 1, history: Optional[dict] = None,
              tol: float = 1e-12, max_step: float = 0.1,
              on_step: Optional[Callable[[OptStepRecord, BasisOrder], None]] = None) -> List[OptStepRecord]:
    """OptCycleI (src/workproc.f90:4902-5423) on a device context: cycles of
    optimising functions func_begin..func_end (their numbering in `order`'s
    original basis), num_func_to_opt at a time, each step moving to the next
    ones by num_func_to_shift.  `history` holds the reference's
    Glob_History(K) fields 'cycles_done' and 'init_func_at_last_step' (and
    receives 'energy'); a run can resume from them.  sigma: the shift
    (Glob_ApproxEnergy).  Returns one record per step; `order` ends in the
    input order of the functions (sorted back, as the reference leaves it)."""
    K = len(order.func_num)
    npt = order.params.shape[1]
    if history is None:
        history = {"cycles_done": 0, "init_func_at_last_step": 0}
    nfs = num_func_to_shift
    fbn = func_begin + K - func_end
    records: List[OptStepRecord] = []
    if history["cycles_done"] >= num_cycles and history["init_func_at_last_step"] >= func_end:
        return records
    order.reverse(func_begin, func_end)
    order.permute_functions(func_begin, func_end)
    if history["init_func_at_last_step"] < func_begin:
        history["init_func_at_last_step"] = func_begin - nfs
    if history["init_func_at_last_step"] >= func_end:
        history["init_func_at_last_step"] = func_begin - nfs
        history["cycles_done"] += 1
    ip = history["init_func_at_last_step"] - func_begin + nfs
    if ip > 0:
        order.permute_functions(fbn, K - ip)

    def energy(which_grad: Optional[int]):
        ctx.set_params(order.params)
        if which_grad is None:
            rc, E, _, info = ctx.energy(1, sigma, tol, 30)
            return rc, E, None, info
        return ctx.energy_grad(which_grad, 1, sigma, tol, 30)

    rc, E_cur, _, info = energy(None)
    if rc not in (0, 2):
        raise RuntimeError("OptCycle: initial energy failed with status %d" % rc)
    for cycle in range(history["cycles_done"] + 1, num_cycles + 1):
        cur_begin = history["init_func_at_last_step"] + nfs
        q = func_end - history["init_func_at_last_step"] - nfs + 1
        opt_iter = 0
        for cur in range(cur_begin, func_end + 1, nfs):
            nfo = min(func_end - cur + 1, num_func_to_opt)
            nfru = K - nfo
            opt_iter += 1
            tails = []
            if opt_iter != 1:
                ns = min(nfs, nfo)
                order.permute_functions(K - nfo - ns + 1, K - ns)
                nr = nfs * num_rows_to_perm_for_unit_shift(opt_iter - 1)
                m = K - nfo
                j = m - nr
                i = (j - nr + 1) if q - nfo >= nr * 2 else (K - q + 1)
                order.permute_functions2(i, j, m)
                rc, E_cur, _, info = energy(None)
                tails.append(int(info[7]))
                if rc not in (0, 2):
                    raise RuntimeError("OptCycle: energy after the permutation failed with status %d" % rc)
            x_init = order.params[nfru:].reshape(-1).copy()
            e_before = E_cur
            scale = np.full(x_init.size, 1.0)

            def set_x(x):
                order.params[nfru:] = x.reshape(nfo, npt)

            def f_only(x):
                set_x(x)
                rc_, e_, _, inf_ = energy(None)
                tails.append(int(inf_[7]))
                return e_ if rc_ in (0, 2) else np.inf

            def fg(x):
                set_x(x)
                rc_, e_, g_, inf_ = energy(nfru)
                tails.append(int(inf_[7]))
                return (e_, g_) if rc_ in (0, 2) else (np.inf, np.zeros_like(x))

            x_best, e_best, nev = bfgs_minimise(fg, f_only, x_init, max_energy_eval, max_step, scale)
            set_x(x_best)
            ctx.set_params(order.params)
            rc, E_new, c, info = ctx.energy(1, sigma, tol, 30, want_vector=True)
            tails.append(int(info[7]))
            ok = rc in (0, 2) and E_new <= e_before
            if ok and overlap_threshold > 0.0:
                pen, _ = ctx.overlap_penalty(nfru, 1.0, overlap_threshold ** 2)
                ok = pen <= 0.0
            if ok and lin_coeff_threshold > 0.0:
                ok = bool(np.all(np.abs(c) <= lin_coeff_threshold))
            if not ok:
                set_x(x_init)
                rc, E_new, _, info = energy(None)
                tails.append(int(info[7]))
            E_cur = E_new
            last = cur > func_end - nfs
            history["energy"] = E_cur
            if last:
                history["init_func_at_last_step"] = 0
                history["cycles_done"] += 1
            else:
                history["init_func_at_last_step"] = cur
            rec = OptStepRecord(func=int(order.func_num[K - 1]), energy_before=float(e_before),
                                energy_after=float(E_cur), evaluations=nev, accepted=bool(ok), tail_rows=tails)
            records.append(rec)
            if on_step:
                on_step(rec, order)
        if cycle != num_cycles:
            history["init_func_at_last_step"] = func_begin - nfs
            order.sort_decreasing(fbn, K)
            rc, E_cur, _, info = energy(None)
    order.sort_decreasing(func_begin, K)
    order.reverse(func_begin, K)
    return records
