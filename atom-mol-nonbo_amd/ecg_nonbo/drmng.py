"""ctypes binding of libecg_drmng.so (include/ecg_drmng.h) and the loop
with which the reference's optimisers drive it.

The minimiser is the reverse-communication secant method with a
double-dogleg trust region of PORT (DIVSET / DRMNG, src/dmng.f:1218,
:2510), restated in C++ (csrc/drmng.cpp) with the reference's workspace
layout and arithmetic order.  `reference_settings` loads the defaults and
the changes every driver of the reference makes (src/workproc.f90:4520-4534
for OptCycleG; BasisEnlG/I, OptCycleI and FullOpt1G/I make the same);
`minimise` is the driver loop (src/workproc.f90:4664-4731): the initial
point is evaluated again by the minimiser's first request, the energy
count (not the minimiser's own) ends the loop at MaxEnergyEval, the best
point seen by either kind of call is kept, and a failed call leaves the
previous energy / gradient in place.

There is no fallback: without the built library the calls raise.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import os
from typing import Callable, Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_HERE, "libecg_drmng.so")

LIV = 60   # the reference's drivers: integer, parameter :: LIV=60
EXPORTS = ["ecg_divset", "ecg_drmng"]

# Glob_MaxScStepAllowedInOpt (src/globvars.f90:307) is declared with a
# default-real literal, 0.001, so the value the drivers store is the single
# precision number widened to double.
MAX_SCALED_FIRST_STEP = float(np.float32(0.001))
# Glob_MaxEnergyFailsAllowed (src/globvars.f90:291)
MAX_FAILURES = 5
EPS = float(np.finfo(np.float64).eps)

_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("%s is missing: build it with atom-mol-nonbo_amd/build.py" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    d = ctypes.POINTER(ctypes.c_double)
    i = ctypes.POINTER(ctypes.c_int)
    lib.ecg_divset.argtypes = [ctypes.c_int, i, ctypes.c_int, ctypes.c_int, d]
    lib.ecg_divset.restype = None
    lib.ecg_drmng.argtypes = [d, ctypes.c_double, d, i, ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d]
    lib.ecg_drmng.restype = None
    _lib = lib
    return lib


def lv_for(nvmax: int) -> int:
    """Double workspace length the drivers allocate (LV = 71 + nvmax
    (nvmax + 13) / 2 + 1)."""
    return 71 + nvmax * (nvmax + 13) // 2 + 1


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def divset(alg: int, iv: np.ndarray, v: np.ndarray) -> None:
    load().ecg_divset(int(alg), _ip(iv), iv.size, v.size, _dp(v))


def drmng(d: np.ndarray, fx: float, g: np.ndarray, iv: np.ndarray, v: np.ndarray, x: np.ndarray) -> int:
    """One call; returns IV(1).  d, g, x: float64 arrays of length n
    (x and d are updated in place), iv: int32[LIV], v: float64[lv]."""
    n = x.size
    if d.size != n or g.size != n:
        raise ValueError("d, g and x must have the same length")
    if v.size < 71 + n * (n + 13) // 2:
        raise ValueError("v too short for n = %d" % n)
    for a, t in ((d, np.float64), (g, np.float64), (x, np.float64), (v, np.float64), (iv, np.int32)):
        if a.dtype != t or not a.flags.c_contiguous:
            raise TypeError("workspace arrays must be contiguous %s" % np.dtype(t).name)
    load().ecg_drmng(_dp(d), float(fx), _dp(g), _ip(iv), iv.size, v.size, n, _dp(v), _dp(x))
    return int(iv[0])


def reference_settings(nvmax: int, lmax0: float = MAX_SCALED_FIRST_STEP, outlev: int = 0
                       ) -> Tuple[np.ndarray, np.ndarray]:
    """IV_init / V_init of the reference's drivers: DIVSET(2) defaults,
    then no evaluation or iteration limit of the minimiser's own (10^6),
    printing off (IV(19) = outlev: 0 in the cyclic and enlargement drivers,
    -1 in FullOpt1), no absolute-function-convergence test (V(31) = 0),
    first scaled step at most lmax0 (V(35)), and IV(1) = 12 ("defaults
    loaded")."""
    iv = np.zeros(LIV, np.int32)
    v = np.zeros(lv_for(nvmax), np.float64)
    divset(2, iv, v)
    iv[16] = 1000000
    iv[17] = 1000000
    iv[18] = outlev
    iv[19] = 0
    iv[21] = 0
    iv[22] = -1
    iv[23] = 0
    v[30] = 0.0
    v[34] = lmax0
    iv[0] = 12
    return iv, v


def scale_for_basis(K: int) -> float:
    """The scale d(i) the cyclic and full optimisers give every parameter:
    max(1 / (K^2 sqrt K), 10^4 eps) (src/workproc.f90:4655, :5671)."""
    return max(1.0 / (float(K * K) * math.sqrt(float(K))), 10000 * EPS)


def scale_for_enlargement(e_init: float, e_cur: float, nfru: int, nfo: int) -> float:
    """The scale of a new function's parameters in BasisEnlG/I: the
    relative energy change of the last enlargement, at least 10^4 eps, or
    1 while the basis is smaller than the group (src/workproc.f90:3456-3460)."""
    if nfru >= nfo:
        return max(abs(e_init - e_cur) / (abs(e_init) + abs(e_cur)), 10000 * EPS)
    return 1.0


@dataclasses.dataclass
class MinimiseResult:
    x_best: np.ndarray     # best point seen by an energy or gradient call (the cyclic / enlargement drivers keep it)
    e_best: float
    x_final: np.ndarray    # the minimiser's x when the loop ended (FullOpt1 keeps it)
    energy_evals: int      # NumOfEnergyEval
    grad_evals: int        # NumOfGradEval
    failures: int
    code: int              # IV(1) when the loop ended


def minimise(energy: Callable[[np.ndarray], Tuple[bool, float]],
             energy_grad: Callable[[np.ndarray], Tuple[bool, float, np.ndarray]],
             x0: np.ndarray, d: np.ndarray, max_energy_eval: int, e_start: float,
             lmax0: float = MAX_SCALED_FIRST_STEP, outlev: int = 0, energy_failure_flag: int = 0,
             grad_failure_flag: int = 0, max_failures: int = MAX_FAILURES, nvmax: Optional[int] = None,
             error: type = RuntimeError) -> MinimiseResult:
    """The reference's driver loop around DRMNG (src/workproc.f90:4664-4731
    for OptCycleG; :3478-3562 BasisEnlG; :5690-5825 FullOpt1G).

    energy(x) -> (ok, E); energy_grad(x) -> (ok, E, g).  e_start is the
    energy at x0 (Glob_CurrEnergy), handed to the minimiser until the first
    successful evaluation and the starting value of the best energy.  A
    failed call counts toward max_failures (exceeding it raises `error`,
    where the reference stops), leaves the previous energy and gradient in
    place, and writes energy_failure_flag / grad_failure_flag into IV(2):
    0 in the cyclic and enlargement drivers (which the minimiser does not
    read as a failure), 1 after a failed energy in FullOpt1 (the minimiser
    then shortens its step).  The minimiser's error codes (IV(1) > 10), on
    which the reference's loop would call it again forever, raise `error`."""
    x = np.array(x0, dtype=np.float64, copy=True)
    n = x.size
    d = np.array(d, dtype=np.float64, copy=True)
    iv, v = reference_settings(nvmax or n, lmax0, outlev)
    grad = np.zeros(n, np.float64)
    fx = float(e_start)
    n_e = n_g = fails = 0
    e_best = float(e_start)
    x_best = x.copy()
    code = int(iv[0])
    done = n_e >= max_energy_eval
    while not done:
        code = drmng(d, fx, grad, iv, v, x)
        if code == 1:
            ok, e = energy(x)
            n_e += 1
            if not ok:
                fails += 1
                iv[1] = energy_failure_flag
            else:
                fx = float(e)
                if e < e_best:
                    e_best = float(e)
                    x_best = x.copy()
        elif code == 2:
            ok, e, g = energy_grad(x)
            n_g += 1
            if not ok:
                fails += 1
                iv[1] = grad_failure_flag
            else:
                grad[:] = g
                if e < e_best:
                    e_best = float(e)
                    x_best = x.copy()
        elif 3 <= code <= 10:
            done = True
        else:
            raise error("minimiser returned IV(1) = %d" % code)
        if fails > max_failures:
            raise error("number of failures in energy or gradient calculations during the optimisation of "
                        "nonlinear parameters exceeded %d" % max_failures)
        if n_e >= max_energy_eval:
            done = True
    return MinimiseResult(x_best=x_best, e_best=e_best, x_final=x.copy(), energy_evals=n_e, grad_evals=n_g,
                          failures=fails, code=code)
