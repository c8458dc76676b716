"""The basis building and optimisation program (BBOP) of an inout.txt file,
run step by step on a device context: the loop of the reference's main
program (src/main.f90:50-221).

Each step sets the shift Glob_ApproxEnergy = current energy x
INVIT_PARAMETER (main.f90:52) and dispatches on its action, with the same
guards as the reference:

* BASIS_ENL m Kstart Kstop Kstep NTrials MaxEnergyEval Q R: when Kstop >
  K and Kstart <= K + 1, grow the basis from K + 1 (BasisEnlI / BasisEnlG,
  `enlarge.basis_enl`);
* OPT_CYCLE m K FuncBegin FuncEnd NumOfFuncToOpt NumOfFuncToShift
  NumCycles MaxEnergyEval Q R: when the first number is the basis size,
  FuncBegin > 0, FuncEnd <= K and fewer than NumCycles cycles are recorded
  for K (OptCycleI / OptCycleG, `optimise.opt_cycle`);
* FULL_OPT1 m K InitFunc FinalFunc MaxEnergyEval Q R ... HessFile: when the
  first number is the basis size (FullOpt1I / FullOpt1G,
  `optimise.full_opt1`; no Hessian file is read or written);
* EXPC_VALS m K: expectation values of the current eigenvector (the
  device part of ExpectationValuesG; run for either method, where the
  reference runs it for 'G' only);
* SAVE_FILE K file: SaveResults into that file;
* the elimination and separation steps are not part of this library, and
  DENSITIES (grids and output files with the equivalent-pair averaging of
  the reference's writer) is left to the caller, who has
  EcgContext.correlation_functions / particle_densities: they are
  reported as skipped.

Method 'I' uses the fixed shift, method 'G' isolates root
WHICH_EIGENVALUE on every energy call.  After every step that changes the
basis the file's fields (parameters, history, current energy) are brought
up to date and, if `path` is given, written there, as SaveResults does.
"""
from __future__ import annotations

import dataclasses
from typing import Callable, List, Optional

import numpy as np

from .enlarge import basis_enl
from .inout import HistoryEntry, InoutFile, write_inout
from .drmng import MAX_SCALED_FIRST_STEP
from .optimise import BasisOrder, full_opt1, opt_cycle

__all__ = ["run_bbop", "BbopReport"]


@dataclasses.dataclass
class BbopReport:
    step: int                 # 1-based BBOP step number
    action: str
    status: str               # "done", "skipped" or "not run" (guard failed)
    energy: float             # current energy after the step
    basis_size: int
    detail: object = None     # the driver's records / result / expectation values


def _hist_dicts(f: InoutFile) -> List[dict]:
    return [dataclasses.asdict(h) for h in f.history] if f.history else \
        [{"energy": float(e), "cycles_done": 0, "init_func_at_last_step": 0, "num_energy_evals_full_opt": 0}
         for e in f.history_energy]


def _store(f: InoutFile, order: BasisOrder, hist: List[dict], energy: float) -> None:
    f.params = np.array(order.params, copy=True)
    f.basis_size = len(order.func_num)
    f.history = [HistoryEntry(float(h["energy"]), int(h["cycles_done"]), int(h["init_func_at_last_step"]),
                              int(h["num_energy_evals_full_opt"])) for h in hist]
    f.history_energy = np.array([h.energy for h in f.history])
    f.current_energy = float(energy)


def run_bbop(ctx, f: InoutFile, rng: np.random.Generator, path: Optional[str] = None, trials: str = "exact",
             steps: Optional[List[int]] = None, maxiter: int = 30, max_step: float = MAX_SCALED_FIRST_STEP,
             on_report: Optional[Callable[[BbopReport], None]] = None) -> List[BbopReport]:
    """Run the BBOP steps of `f` (all, or the 1-based numbers in `steps`)
    on `ctx` (a device context of f.system), updating `f` in place.
    Returns one report per step."""
    reports: List[BbopReport] = []
    order = BasisOrder.identity(f.params.reshape(-1, f.params.shape[-1]) if f.basis_size else
                                np.zeros((0, f.system.npt)))
    hist = _hist_dicts(f)
    tol = f.eigval_tolerance
    which = f.which_eigenvalue
    for si, st in enumerate(f.bbop, start=1):
        if steps is not None and si not in steps:
            continue
        K = len(order.func_num)
        sigma = f.current_energy * f.invit_parameter if st.method == "I" else float("nan")
        a = st.action
        status, detail = "not run", None

        def save():
            _store(f, order, hist, f.current_energy)
            if path:
                write_inout(path, f)

        if a == "BASIS_ENL":
            k_start, k_stop, k_step, n_trials, max_eval = st.ints[:5]
            q, r = st.reals[:2]
            if k_stop > K and k_start <= K + 1:
                def on_step(rec, o):
                    f.current_energy = rec.energy_after
                    save()
                detail = basis_enl(ctx, order, k_stop, k_step, n_trials, max_eval, sigma, rng,
                                   generator_param=f.generator_param, overlap_threshold=q, lin_coeff_threshold=r,
                                   history=hist, which=which, tol=tol, maxiter=maxiter, max_step=max_step,
                                   trials=trials, on_step=on_step)
                status = "done"
        elif a == "OPT_CYCLE":
            kk, fb, fe, nopt, nshift, ncyc, max_eval = st.ints[:7]
            q, r = st.reals[:2]
            if kk == K and fb > 0 and fe <= K and hist[K - 1]["cycles_done"] < ncyc:
                h = hist[K - 1]

                def on_step(rec, o):
                    f.current_energy = rec.energy_after
                detail = opt_cycle(ctx, order, fb, fe, ncyc, max_eval, sigma, q, r, nopt, nshift, history=h,
                                   tol=tol, max_step=max_step, which=which, maxiter=maxiter, on_step=on_step)
                f.current_energy = h.get("energy", f.current_energy)
                save()
                status = "done"
        elif a == "FULL_OPT1":
            kk, fi, ff, max_eval = st.ints[:4]
            q, r = st.reals[:2]
            if kk == K and fi > 0 and ff <= K:
                detail = full_opt1(ctx, order, fi, ff, max_eval, sigma, q, r, history=hist[K - 1], which=which,
                                   tol=tol, maxiter=maxiter, max_step=max_step)
                f.current_energy = detail.energy_after
                save()
                status = "done"
        elif a == "EXPC_VALS":
            if st.ints and st.ints[0] == K and K > 0:
                ctx.set_params(order.params)
                rc, E, c, _ = ctx.energy(which, float("nan"), tol, maxiter, want_vector=True)
                detail = ctx.expectation_values(c)
                status = "done"
        elif a == "SAVE_FILE":
            if st.ints and st.ints[0] == K:
                _store(f, order, hist, f.current_energy)
                write_inout(st.files[0], f)
                status = "done"
        else:
            status = "skipped"
        rep = BbopReport(step=si, action=a, status=status, energy=float(f.current_energy),
                         basis_size=len(order.func_num), detail=detail)
        reports.append(rep)
        if on_report:
            on_report(rep)
    _store(f, order, hist, f.current_energy)
    return reports
