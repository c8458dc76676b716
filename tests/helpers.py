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


def penalty_basis(seeded, n, K, seed, ndup, eps=0.03):
    """A seeded basis whose last ndup functions are perturbed copies of
    functions spread over the rest, so that the overlap penalty has hits
    between fixed and optimised and between two optimised functions."""
    p = seeded(n, K, seed)
    rng = np.random.default_rng(seed + 7)
    src = np.linspace(0, K - ndup - 1, ndup).astype(int)
    for t, k in enumerate(src):
        p[K - ndup + t] = p[k] * (1.0 + eps * rng.standard_normal(p.shape[1]))
    return p


def penalty_threshold(S, lo=0.3, hi=0.97):
    """t2 in (lo, hi) in the middle of the widest gap between consecutive
    |s_ji|^2 of the stored lower triangle, so that no pair sits near the
    threshold; returns (t2, gap, number of pairs above)."""
    K = S.shape[0]
    v = np.sort(np.abs(S[np.tril_indices(K, -1)]) ** 2)
    best = None
    for a, b in zip(v[:-1], v[1:]):
        mid = 0.5 * (a + b)
        if lo < mid < hi and (best is None or b - a > best[1]):
            best = (mid, b - a)
    assert best is not None
    return best[0], best[1], int(np.sum(v > best[0]))
