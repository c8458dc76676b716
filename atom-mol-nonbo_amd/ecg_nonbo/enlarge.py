"""Basis enlargement: random trials for new functions, the best kept and
optimised, the loop of the reference's BasisEnlI (src/workproc.f90:3734-4379)
and BasisEnlG (:3079-3730), over the device energy path.

What follows the reference: the candidate generator GenerateTrialParam
(:1375-1487, its two methods and its probabilities), the structure of one
enlargement step (NTrials candidates for functions nfru+1..K, each scored
by EnergyIA(nfru+1, K) — on the device an assembly of the new rows and an
LDL^H update from row nfru+1; the best one kept if it lowers the energy;
its optimisation with at most MaxEnergyEval energy evaluations; the
overlap and linear-coefficient checks with up to
Glob_BadOverlapOrLinCoeffLim attempts), the failure limits
(Glob_MaxFracOfTrialFailsAllowed of the trials, Glob_MaxEnergyFailsAllowed
in the optimisation), the history entries of the new functions and their
numbers.

The optimisation of the kept candidate drives the restated DRMNG
(`drmng.minimise`) with the reference's settings and scale.  What does
not follow the reference: random streams (numpy's PCG64 in place of
random_number and DRNOR, so the candidates are drawn from the same
distributions, not the same numbers).
"""
from __future__ import annotations

import dataclasses
from typing import Callable, List, Optional, Tuple

import numpy as np

from . import drmng
from .optimise import BasisOrder, EnergyCalls

__all__ = ["generate_trial_param", "basis_enl", "EnlStepRecord", "BasisEnlError",
           "BAD_OVERLAP_OR_LIN_COEFF_LIM", "MAX_FRAC_OF_TRIAL_FAILS", "MAX_ENERGY_FAILS"]

# src/globvars.f90:287, 291, 302
MAX_FRAC_OF_TRIAL_FAILS = 0.15
MAX_ENERGY_FAILS = 5
BAD_OVERLAP_OR_LIN_COEFF_LIM = 10


class BasisEnlError(RuntimeError):
    """A condition on which the reference's BasisEnl stops the program."""


def _scale_factor(rng: np.random.Generator, s2: float, max_draws: int = 1_000_000) -> float:
    # method 2's common factor 1 + s2 N(0,1), redrawn while 0.8 < |factor| < 1.2
    # (src/workproc.f90:1446-1449, 1473-1476): near-copies of a function are avoided
    for _ in range(max_draws):
        r = s2 * rng.standard_normal() + 1.0
        if not (0.8 < abs(r) < 1.2):
            return r
    raise BasisEnlError("GenerateTrialParam: no scale factor outside (0.8, 1.2) with s2=%g" % s2)


def generate_trial_param(rng: np.random.Generator, params: np.ndarray, nfun: int, p1: float, s1: float,
                         s2: float, npt: Optional[int] = None) -> Tuple[np.ndarray, int, int]:
    """GenerateTrialParam (src/workproc.f90:1375-1487): parameters of nfun
    trial functions from the current basis `params` (rows = functions).
    Returns (x (nfun x npt), k, method_used): k is the 1-based number of the
    (first) prototype function, 0 with an empty basis; method_used 1 or 2
    (0 with an empty basis).

    * Empty basis: vech L and vech B entries uniform in [-0.5, 0.5)
      (:1420-1430).
    * Otherwise, with probability p1 method 1: every parameter of the
      prototype times 1 + s1 N(0,1), independently; else method 2: all
      parameters times one factor 1 + s2 N(0,1) kept outside (0.8, 1.2).
    * With at least nfun functions, the prototypes are nfun consecutive
      functions from a uniformly drawn k (:1459-1483); with fewer, each
      trial function draws its own prototype (:1432-1458).  In that second
      case the reference's method 2 overwrites the factor with the uniform
      number that picks the prototype (:1450-1455), so the factor is
      uniform in [0, 1) there; this is kept."""
    params = np.asarray(params, dtype=np.float64)
    K = params.shape[0]
    if npt is None:
        npt = params.shape[1]
    x = np.empty((nfun, npt))
    if K == 0:
        x[:] = rng.random((nfun, npt)) - 0.5
        return x, 0, 0
    if K < nfun:
        if rng.random() < p1:
            k = 0
            for i in range(nfun):
                k = int(rng.random() * K) + 1
                x[i] = (s1 * rng.standard_normal(npt) + 1.0) * params[k - 1]
            return x, k, 1
        _scale_factor(rng, s2)
        k = 0
        for i in range(nfun):
            r = rng.random()
            k = int(r * K) + 1
            x[i] = r * params[k - 1]
        return x, k, 2
    k = int(rng.random() * (K - nfun + 1)) + 1
    if rng.random() < p1:
        for i in range(nfun):
            x[i] = (s1 * rng.standard_normal(npt) + 1.0) * params[k - 1 + i]
        return x, k, 1
    r = _scale_factor(rng, s2)
    x[:] = r * params[k - 1:k - 1 + nfun]
    return x, k, 2


@dataclasses.dataclass
class EnlStepRecord:
    first_func: int           # nfru + 1
    nfo: int                  # functions added in the step
    energy_before: float      # energy of the nfru-function basis (inf when empty)
    energy_trials: float      # best trial energy of the last attempt
    energy_after: float       # energy after the optimisation (Glob_CurrEnergy)
    attempts: int             # AttemptToGetGoodFunc - 1
    trial_failures: int       # failed trial energies in the last attempt
    evaluations: int          # energy evaluations of the last optimisation
    prototype: int            # wbfu: prototype of the kept candidate
    method_used: int          # wmu: its generation method
    checks_passed: bool       # overlap and linear-coefficient checks of the last attempt


def _screen_trials(ctx, params, nfru, trials, sigma, tol, maxiter, iters):
    """Energies of nfo = 1 trial functions in one batched call: functions
    1..nfru fixed and factored once (ecg_screen_prepare), each candidate
    scored by `iters` inverse iterations from the eigenvector of the basis
    with the first candidate (ecg_screen).  None if the fixed part is
    singular at sigma."""
    p = params.copy()
    p[nfru] = trials[0][0]
    ctx.set_params(p)
    try:
        ctx.screen_prepare(sigma, tol, maxiter)
    except RuntimeError as e:
        if getattr(e, "code", None) == 1:
            return None
        raise
    E, _ = ctx.screen(np.array([t[0] for t in trials]), iters)
    return E


def basis_enl(ctx, order: BasisOrder, k_stop: int, k_step: int, n_trials: int, max_energy_eval: int,
              sigma: float, rng: np.random.Generator, generator_param=(0.5, 0.1, 0.3),
              overlap_threshold: float = -1.0, lin_coeff_threshold: float = -1.0, optimise: bool = True,
              history: Optional[list] = None, which: int = 1, tol: float = 1e-12, maxiter: int = 30,
              max_step: float = drmng.MAX_SCALED_FIRST_STEP, trials: str = "exact", screen_iters: int = 4,
              bad_lim: int = BAD_OVERLAP_OR_LIN_COEFF_LIM,
              on_step: Optional[Callable[[EnlStepRecord, BasisOrder], None]] = None) -> List[EnlStepRecord]:
    """BasisEnlI (sigma finite: the fixed shift Glob_ApproxEnergy) or
    BasisEnlG (sigma NaN: root `which` isolated on every call): grow the
    basis in `order` (Kstart - 1 functions; Kstart = current size + 1, as
    main.f90:58-61 sets it) to k_stop functions, k_step at a time.

    generator_param: (Glob_RG_p1, Glob_RG_s1, Glob_RG_s2).  optimise: the
    reference's OptimizationType 1 (True) or 0.  history: a list of
    dict-like entries (energy, cycles_done, init_func_at_last_step,
    num_energy_evals_full_opt) to which each new function's entry is
    appended (:4318-4326).  on_step(record, order) runs after each step
    (where the reference calls SaveResults).

    trials: "exact" scores every trial with a full energy call (EnergyIA's
    semantics); "screen" (k_step 1, method 'I') scores them all in one
    batched device call with `screen_iters` inverse iterations each and
    then evaluates the kept one exactly — faster, with the kept candidate
    possibly differing from the exact loop's when two trial energies agree
    to about the screening accuracy.

    As in the reference, a repeated attempt (after a failed overlap or
    coefficient check) must beat the energy of the rejected attempt's
    optimised functions, and after bad_lim attempts the last one is kept
    whatever the checks say.  Raises BasisEnlError where the reference
    stops."""
    p1, s1, s2 = (float(v) for v in generator_param)
    if trials not in ("exact", "screen"):
        raise ValueError("trials must be 'exact' or 'screen'")
    if n_trials < 1 or k_step < 1:
        raise ValueError("BasisEnl needs NTrials >= 1 and Kstep >= 1")
    use_screen = trials == "screen" and k_step == 1 and np.isfinite(sigma)
    npt = order.params.shape[1]
    K = len(order.func_num)
    records: List[EnlStepRecord] = []
    if k_stop <= K:
        return records
    ev = EnergyCalls(ctx, sigma, which, tol, maxiter)
    params = np.zeros((k_stop, npt))
    params[:K] = order.params
    func_num = np.zeros(k_stop, dtype=np.int64)
    func_num[:K] = order.func_num
    if K > 0:
        rc, E_cur, _, _ = ev.energy(params[:K])
        if rc != 0:
            raise BasisEnlError("BasisEnl: initial energy cannot be computed (status %d)" % rc)
    else:
        E_cur = np.inf
    while K < k_stop:
        nfo = min(k_step, k_stop - K)
        nfru = K
        K += nfo
        E_init = E_cur
        bad = True
        attempt = 1
        while bad and attempt <= bad_lim:
            fails = 0
            improved = False
            best = None
            wbfu = wmu = 0
            E_trial_best = np.inf
            cand = [generate_trial_param(rng, params[:nfru], nfo, p1, s1, s2, npt) for _ in range(n_trials)]
            E_scr = _screen_trials(ctx, params[:K], nfru, cand, sigma, tol, maxiter, screen_iters) \
                if use_screen and nfru > 0 else None
            for t, (x, k_t, m_t) in enumerate(cand):
                if E_scr is not None:
                    e = E_scr[t]
                    ok = bool(np.isfinite(e))
                else:
                    params[nfru:K] = x
                    rc, e, _, _ = ev.energy(params[:K])
                    ok = rc == 0
                if not ok:
                    fails += 1
                elif e < E_cur:
                    E_cur = e
                    E_trial_best = e
                    best = x.copy()
                    improved = True
                    wbfu, wmu = k_t, m_t
            if fails / n_trials > MAX_FRAC_OF_TRIAL_FAILS:
                raise BasisEnlError("BasisEnl: the limit of eigenvalue problem failures in the random selection "
                                    "is exceeded (%d of %d trials)" % (fails, n_trials))
            if not improved:
                raise BasisEnlError("BasisEnl: random selection did not result in any energy improvement")
            params[nfru:K] = best
            rc, E_cur, _, _ = ev.energy(params[:K])
            if rc != 0:
                raise BasisEnlError("BasisEnl: energy of the selected trial failed (status %d)" % rc)
            nev = 0
            x_best = best.reshape(-1).copy()
            if optimise and max_energy_eval > 0:
                def set_x(v):
                    params[nfru:K] = v.reshape(nfo, npt)

                def f_only(v):
                    set_x(v)
                    rc_, e_, _, _ = ev.energy(params[:K])
                    return rc_ == 0, e_

                def fg(v):
                    set_x(v)
                    rc_, e_, g_, _ = ev.energy_grad(params[:K], nfru)
                    return rc_ == 0, e_, g_

                # src/workproc.f90:3447-3562: the workspace from its initial
                # copy, every scale the relative energy gain of the
                # selection (1 while the basis is smaller than the group)
                t = drmng.scale_for_enlargement(E_init, E_cur, nfru, nfo)
                res = drmng.minimise(f_only, fg, x_best, np.full(x_best.size, t), max_energy_eval, E_cur,
                                     lmax0=max_step, error=BasisEnlError)
                x_best, nev = res.x_best, res.energy_evals
            params[nfru:K] = x_best.reshape(nfo, npt)
            rc, E_cur, c, _ = ev.energy(params[:K], want_vector=True)
            if rc != 0:
                raise BasisEnlError("BasisEnl: failed to evaluate the energy after the optimisation (status %d)"
                                    % rc)
            bad_overlap = False
            if overlap_threshold > 0.0:
                pen, _ = ctx.overlap_penalty(nfru, 1.0, overlap_threshold ** 2)
                bad_overlap = pen > 0.0
            bad_coeff = False
            if lin_coeff_threshold > 0.0:
                bad_coeff = bool(np.any(np.abs(c) > lin_coeff_threshold))
            bad = bad_overlap or bad_coeff
            attempt += 1
        func_num[nfru:K] = np.arange(nfru + 1, K + 1)
        order.params = params[:K].copy()
        order.func_num = func_num[:K].copy()
        if history is not None:
            for _ in range(nfo):
                history.append({"energy": float(E_cur), "cycles_done": 0, "init_func_at_last_step": 0,
                                "num_energy_evals_full_opt": 0})
        rec = EnlStepRecord(first_func=nfru + 1, nfo=nfo, energy_before=float(E_init),
                            energy_trials=float(E_trial_best), energy_after=float(E_cur), attempts=attempt - 1,
                            trial_failures=fails, evaluations=nev, prototype=wbfu, method_used=wmu,
                            checks_passed=not bad)
        records.append(rec)
        if on_step:
            on_step(rec, order)
    return records
