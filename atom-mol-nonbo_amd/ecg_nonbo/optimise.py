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
row), the minimiser and its driver loop (PORT's DRMNG restated in
csrc/drmng.cpp with the reference's settings: scale max(1 / (K^2 sqrt K),
1e4 eps) on every parameter, first scaled step 0.001, at most
MaxEnergyEval energy evaluations, the best point kept; `drmng.minimise`),
the acceptance tests on overlaps and linear coefficients, and the history
fields (CyclesDone, InitFuncAtLastStep) with which the reference resumes a
cycle.  FullOpt1 (`full_opt1`) drives the same minimiser over a block of
functions at once.
"""
from __future__ import annotations

import dataclasses
from typing import Callable, List, Optional

import numpy as np

from . import drmng

__all__ = ["BasisOrder", "num_rows_to_perm_for_unit_shift", "opt_cycle", "OptStepRecord", "EnergyCalls",
           "full_opt1", "FullOptResult", "FAILED_STATUSES"]

# Statuses with which an energy call counts as a failure (the reference's
# ErrorCode /= 0): singular H - sigma S, inverse iteration not converged,
# root not isolated, non-finite pair element.  Other negative statuses are
# library errors and raise.
FAILED_STATUSES = (1, 2, 3, -4)


class EnergyCalls:
    """The energy routines an optimiser loop calls, bound to a context and a
    shift: EnergyIA / EnergyIAM (`energy`) and EnergyIB (`energy_grad`),
    src/workproc.f90:1954-2154 for method 'I' (sigma finite, the fixed
    Glob_ApproxEnergy) and EnergyGA / EnergyGB (:1697-1950) for method 'G'
    (sigma NaN: the device isolates root `which` by inertia counts).  Each
    call returns its status instead of raising for the statuses in
    FAILED_STATUSES, as the reference's routines return ErrorCode."""

    def __init__(self, ctx, sigma: float, which: int = 1, tol: float = 1e-12, maxiter: int = 30):
        self.ctx = ctx
        self.sigma = float(sigma)
        self.which = int(which)
        self.tol = float(tol)
        self.maxiter = int(maxiter)
        self.tails: List[int] = []

    def _call(self, fn, *args, **kw):
        try:
            return fn(*args, **kw)
        except RuntimeError as e:
            code = getattr(e, "code", None)
            if code in FAILED_STATUSES:
                return code, np.inf, None, np.full(8, np.nan)
            raise

    def energy(self, params, want_vector: bool = False):
        """(status, E, c or None, info) of the basis `params` (K x npt)."""
        self.ctx.set_params(params)
        rc, E, c, info = self._call(self.ctx.energy, self.which, self.sigma, self.tol, self.maxiter,
                                    want_vector=want_vector)
        self.tails.append(int(info[7]) if np.isfinite(info[7]) else -1)
        return rc, E, c, info

    def energy_grad(self, params, nfru: int):
        """(status, E, gradient of functions nfru+1..K, info)."""
        self.ctx.set_params(params)
        rc, E, g, info = self._call(self.ctx.energy_grad, nfru, self.which, self.sigma, self.tol, self.maxiter)
        self.tails.append(int(info[7]) if np.isfinite(info[7]) else -1)
        return rc, E, g, info



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


@dataclasses.dataclass
class OptStepRecord:
    func: int                 # function number (Glob_FuncNum) optimised
    energy_before: float
    energy_after: float
    evaluations: int          # energy evaluations (NumOfEnergyEval)
    accepted: bool
    tail_rows: List[int]      # rows refactored by the energy calls of the step (-1: full)
    grad_evaluations: int = 0   # gradient evaluations (NumOfGradEval)


def opt_cycle(ctx, order: BasisOrder, func_begin: int, func_end: int, num_cycles: int, max_energy_eval: int,
              sigma: float, overlap_threshold: float = -1.0, lin_coeff_threshold: float = -1.0,
              num_func_to_opt: int = 1, num_func_to_shift: int = 1, history: Optional[dict] = None,
              tol: float = 1e-12, max_step: float = drmng.MAX_SCALED_FIRST_STEP, which: int = 1, maxiter: int = 30,
              on_step: Optional[Callable[[OptStepRecord, BasisOrder], None]] = None) -> List[OptStepRecord]:
    """OptCycleI (src/workproc.f90:4902-5423) on a device context: cycles of
    optimising functions func_begin..func_end (their numbering in `order`'s
    original basis), num_func_to_opt at a time, each step moving to the next
    ones by num_func_to_shift.  `history` holds the reference's
    Glob_History(K) fields 'cycles_done' and 'init_func_at_last_step' (and
    receives 'energy'); a run can resume from them.  sigma: the shift
    (Glob_ApproxEnergy) for OptCycleI; NaN gives OptCycleG
    (src/workproc.f90:4383-4898, the same loop over EnergyGA / EnergyGB:
    root `which` isolated on every call).  Returns one record per step;
    `order` ends in the input order of the functions (sorted back, as the
    reference leaves it).  max_step: the minimiser's bound on its first
    scaled step (V(35); the reference's Glob_MaxScStepAllowedInOpt by
    default).  An energy call that fails inside the minimiser leaves the
    last energy in place (up to Glob_MaxEnergyFailsAllowed of them, then
    RuntimeError); failures of the calls outside it raise, as the reference
    stops there."""
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

    ev = EnergyCalls(ctx, sigma, which, tol, maxiter)

    def energy(which_grad: Optional[int]):
        if which_grad is None:
            return ev.energy(order.params)
        return ev.energy_grad(order.params, which_grad)

    rc, E_cur, _, info = energy(None)
    if rc != 0:
        raise RuntimeError("OptCycle: initial energy failed with status %d" % rc)
    for cycle in range(history["cycles_done"] + 1, num_cycles + 1):
        cur_begin = history["init_func_at_last_step"] + nfs
        q = func_end - history["init_func_at_last_step"] - nfs + 1
        opt_iter = 0
        for cur in range(cur_begin, func_end + 1, nfs):
            nfo = min(func_end - cur + 1, num_func_to_opt)
            nfru = K - nfo
            opt_iter += 1
            t0 = len(ev.tails)
            if opt_iter != 1:
                ns = min(nfs, nfo)
                order.permute_functions(K - nfo - ns + 1, K - ns)
                nr = nfs * num_rows_to_perm_for_unit_shift(opt_iter - 1)
                m = K - nfo
                j = m - nr
                i = (j - nr + 1) if q - nfo >= nr * 2 else (K - q + 1)
                order.permute_functions2(i, j, m)
            x_init = order.params[nfru:].reshape(-1).copy()
            e_before = E_cur

            def set_x(x):
                order.params[nfru:] = x.reshape(nfo, npt)

            def f_only(x):
                set_x(x)
                rc_, e_, _, _ = energy(None)
                return rc_ == 0, e_

            def fg(x):
                set_x(x)
                rc_, e_, g_, _ = energy(nfru)
                return rc_ == 0, e_, g_

            # src/workproc.f90:4648-4731: the workspace from its initial
            # copy, every scale max(1 / (K^2 sqrt K), 1e4 eps), the loop
            # ended by the energy count
            d = np.full(x_init.size, drmng.scale_for_basis(K))
            res = drmng.minimise(f_only, fg, x_init, d, max_energy_eval, E_cur, lmax0=max_step)
            set_x(res.x_best)
            rc, E_new, c, info = ev.energy(order.params, want_vector=True)
            ok = rc == 0
            if ok and overlap_threshold > 0.0:
                pen, _ = ctx.overlap_penalty(nfru, 1.0, overlap_threshold ** 2)
                ok = pen <= 0.0
            if ok and lin_coeff_threshold > 0.0:
                ok = bool(np.all(np.abs(c) <= lin_coeff_threshold))
            if not ok:
                set_x(x_init)
                rc, E_new, _, info = energy(None)
                if rc != 0:
                    raise RuntimeError("OptCycle: energy at the step's starting point failed with status %d" % rc)
            E_cur = E_new
            last = cur > func_end - nfs
            history["energy"] = E_cur
            if last:
                history["init_func_at_last_step"] = 0
                history["cycles_done"] += 1
            else:
                history["init_func_at_last_step"] = cur
            rec = OptStepRecord(func=int(order.func_num[K - 1]), energy_before=float(e_before),
                                energy_after=float(E_cur), evaluations=res.energy_evals, accepted=bool(ok),
                                tail_rows=list(ev.tails[t0:]), grad_evaluations=res.grad_evals)
            records.append(rec)
            if on_step:
                on_step(rec, order)
        if cycle != num_cycles:
            history["init_func_at_last_step"] = func_begin - nfs
            order.sort_decreasing(fbn, K)
    order.sort_decreasing(func_begin, K)
    order.reverse(func_begin, K)
    return records


@dataclasses.dataclass
class FullOptResult:
    energy_before: float      # E + P at the starting point (P = 0 without the penalty)
    energy_after: float       # energy without the penalty at the end (Glob_CurrEnergy)
    penalty: float            # Glob_TotalOverlapPenalty at the end
    evaluations: int          # energy evaluations (the reference's NumOfEnergyEval)
    grad_evaluations: int     # gradient evaluations (NumOfGradEval)
    code: int                 # the minimiser's IV(1) when the loop ended


def full_opt1(ctx, order: BasisOrder, init_func: int, final_func: int, max_energy_eval: int, sigma: float,
              overlap_threshold: float = 1.0, max_overlap_penalty: float = 1.0, history: Optional[dict] = None,
              which: int = 1, tol: float = 1e-12, maxiter: int = 30,
              max_step: float = drmng.MAX_SCALED_FIRST_STEP) -> FullOptResult:
    """FullOpt1I (src/workproc.f90:5962-6422; FullOpt1G, :5428-5958, with
    sigma NaN): simultaneous optimisation of functions init_func..final_func
    (positions in `order`), E + P and its gradient from the device.

    Steps as the reference: the overlap penalty is on when
    overlap_threshold < 1 (t2 = threshold^2, b = max_overlap_penalty,
    :6026-6032); functions init_func..final_func move to the end of the
    basis (PermuteFunctions, :6113-6127); the minimiser (DRMNG, dense, with
    the drivers' settings and summary level -1, every scale max(1 / (K^2
    sqrt K), 1e4 eps)) runs until it converges or max_energy_eval energy
    evaluations are spent, a failed energy flagged to it through IV(2) = 1
    (:6236-6262); the parameters it holds at the end (its last point, not
    the best seen) are kept (:6303-6306); the energy without the penalty
    becomes the current energy (:6328-6332) and history['energy'],
    history['num_energy_evals_full_opt'] grow as Glob_History(K) does
    (:6356-6360); the functions that followed final_func move back behind
    them (:6366-6378), restoring the input order.  The penalty is switched
    off on return.

    Not here: the Hessian file (SaveHessian) and the periodic saves of the
    data file during the loop."""
    K = len(order.func_num)
    if not (1 <= init_func <= final_func <= K):
        raise ValueError("FullOpt1: need 1 <= InitFunc <= FinalFunc <= K")
    npt = order.params.shape[1]
    nfo = final_func - init_func + 1
    nfru = K - nfo
    allowed = overlap_threshold < 1.0
    ctx.set_overlap_penalty(allowed, max_overlap_penalty, overlap_threshold ** 2 if allowed else 1.0, nfru)
    try:
        if final_func != K:
            order.permute_functions(init_func, final_func)
        ev = EnergyCalls(ctx, sigma, which, tol, maxiter)
        rc, E0, _, _ = ev.energy(order.params)
        if rc != 0:
            raise RuntimeError("FullOpt1: initial energy failed with status %d" % rc)
        x0 = order.params[nfru:].reshape(-1).copy()

        def set_x(x):
            order.params[nfru:] = x.reshape(nfo, npt)

        def f_only(x):
            set_x(x)
            rc_, e_, _, _ = ev.energy(order.params)
            return rc_ == 0, e_

        def fg(x):
            set_x(x)
            rc_, e_, g_, _ = ev.energy_grad(order.params, nfru)
            return rc_ == 0, e_, g_

        d = np.full(x0.size, drmng.scale_for_basis(K))
        res = drmng.minimise(f_only, fg, x0, d, max_energy_eval, E0, lmax0=max_step, outlev=-1,
                             energy_failure_flag=1, grad_failure_flag=0)
        set_x(res.x_final)
        rc, E, _, info = ev.energy(order.params)
        if rc != 0:
            raise RuntimeError("FullOpt1: energy after the optimisation failed with status %d" % rc)
        P = float(info[6]) if allowed else 0.0
        nev = res.energy_evals
        if history is not None:
            history["energy"] = E - P
            history["num_energy_evals_full_opt"] = history.get("num_energy_evals_full_opt", 0) + nev
        if final_func != K:
            order.permute_functions(init_func, init_func + (K - final_func) - 1)
        return FullOptResult(energy_before=float(E0), energy_after=float(E - P), penalty=P, evaluations=nev,
                             grad_evaluations=res.grad_evals, code=res.code)
    finally:
        ctx.set_overlap_penalty(False, 1.0, 1.0, 0)
