"""Seeded synthetic ECG bases and the named test systems.

Parameter vector of one basis function: (vech L, vech B) with vech running
down the columns of the lower triangle, L(1,1), L(2,1), ..., L(n,1),
L(2,2), ... (src/matelem.f90:109-122).

Distribution of a seeded function (numpy PCG64, seed recorded by caller):
  L_ii = 10**U(-2, 2);  L_ij (i > j) ~ N(0, (0.1 * L_jj)^2);
  B symmetric, B_ij ~ N(0, 0.05^2).
"""
from __future__ import annotations

import numpy as np

from .system import make_system

__all__ = ["seeded_basis", "vech_index", "SYSTEMS", "named_system", "SEED_BASE"]

SEED_BASE = 20261203

# Masses in electron masses, charges in e.  Particle 1 is the reference.
SYSTEMS = {
    # 3-particle H^- like: nucleus + two electrons, singlet in the electrons
    "hminus": dict(masses=[1836.15267, 1.0, 1.0], charges=[1.0, -1.0, -1.0], young="(1+P23)"),
    # 4-particle H2 like: two protons (1,2), two electrons (3,4)
    "h2": dict(masses=[1836.15267, 1836.15267, 1.0, 1.0], charges=[1.0, 1.0, -1.0, -1.0],
               young="(1+P12)(1+P34)"),
    # 4-particle Li like: nucleus 12786.39 m_e, charge +3, three electrons
    # (doublet: symmetric in 2,3, antisymmetrised against 4)
    "li": dict(masses=[12786.39, 1.0, 1.0, 1.0], charges=[3.0, -1.0, -1.0, -1.0],
               young="(1-P24)(1+P23)"),
    # HD+ as in the reference's example inputs
    "hdplus": dict(masses=[3670.4829652, 1836.15267261, 1.0], charges=[1.0, 1.0, -1.0],
                   young="(P11)"),
}


def named_system(name: str):
    return make_system(**SYSTEMS[name])


def vech_index(n: int, i: int, j: int) -> int:
    """0-based position of L(i, j) (i >= j, 0-based) in vech L."""
    if i < j:
        i, j = j, i
    return j * n - j * (j - 1) // 2 + (i - j)


def seeded_basis(n: int, K: int, seed: int) -> np.ndarray:
    """(K, n(n+1)) float64 parameters drawn from the distribution above."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = n * (n + 1) // 2
    out = np.empty((K, 2 * m))
    for k in range(K):
        diag = 10.0 ** rng.uniform(-2.0, 2.0, size=n)
        L = np.zeros((n, n))
        for j in range(n):
            L[j, j] = diag[j]
            for i in range(j + 1, n):
                L[i, j] = rng.normal(0.0, 0.1 * diag[j])
        B = rng.normal(0.0, 0.05, size=(n, n))
        B = np.tril(B) + np.tril(B, -1).T
        idx = 0
        for j in range(n):
            for i in range(j, n):
                out[k, idx] = L[i, j]
                out[k, m + idx] = B[i, j]
                idx += 1
    return out
