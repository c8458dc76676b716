"""Shared test helpers: FP64 error bars for pair elements.

The error bar of a pair element is measured on the oracle itself:
  * the element is recomputed with both parameter vectors perturbed by a
    few units in the last place; the spread is the rounding sensitivity of
    the reference algorithm for that pair (large when C_kl is nearly
    singular, e.g. strongly correlated L);
  * H = T + V may cancel, so rounding of H is relative to |T| + |V|, not
    |H|; T alone is the oracle with all charges set to zero.
Two correct FP64 implementations that round differently can differ by
about that much, so parity is asserted against
    max(1e-12 x magnitude, 16 x sensitivity)
(calibrated on the CPU: the largest observed ratio of deviation to
8 x sensitivity over ~3700 seeded pairs of the three test systems was 1.0).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from oracle import oracle

EPS = np.finfo(float).eps


def _kinetic_only(system):
    return dataclasses.replace(system, q0=0.0, q=np.zeros_like(system.q))


def pair_sensitivity(system, pk, pl, gradflag=3, term=None, trials=6, seed=0):
    """Returns ((H, S, Dk, Dl), sens[4], mag[4]) for one pair (summed over
    the Y^dagger Y terms, or one term)."""
    rng = np.random.default_rng(seed)
    H0, S0, Dk0, Dl0 = oracle.pair(system, pk, pl, gradflag, term=term)
    T0, _, DkT, DlT = oracle.pair(_kinetic_only(system), pk, pl, gradflag, term=term)
    V0 = H0 - T0
    np_ = system.np_
    mag = np.array([abs(T0) + abs(V0), abs(S0),
                    np.max(np.abs(DkT) + np.abs(Dk0 - DkT)), np.max(np.abs(DlT) + np.abs(Dl0 - DlT))])
    s = np.zeros(4)
    for _ in range(trials):
        a = pk * (1 + 4 * EPS * rng.standard_normal(pk.shape))
        b = pl * (1 + 4 * EPS * rng.standard_normal(pl.shape))
        H, S, Dk, Dl = oracle.pair(system, a, b, gradflag, term=term)
        s = np.maximum(s, [abs(H - H0), abs(S - S0), np.max(np.abs(Dk - Dk0)), np.max(np.abs(Dl - Dl0))])
    return (H0, S0, Dk0, Dl0), s, mag


def close(a, b, mag, sens, rel=1e-12, factor=16.0):
    """max |a - b| <= max(rel * mag, factor * sens)."""
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    tol = max(rel * mag, factor * sens)
    return d <= tol, d, tol
