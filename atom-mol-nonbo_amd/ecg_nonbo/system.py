"""Problem set-up for the ECG pair path: mass matrix, pseudo-charges,
transposition matrices and the simplified Y^dagger Y symmetry terms.

Host-side only.  Everything here is integer or O(n^2) real arithmetic that
feeds the pair kernels their constant inputs.

Reference interfaces mirrored (all in src/workproc.f90 of ATOM-MOL-nonBO):
  * mass matrix                     ProgramDataInit :606-610
  * transposition matrices P_ij     ProgramDataInit :673-698
  * Young operator string -> Y, Y^dagger Y terms, identical terms merged
                                    ProgramDataInit :700-1202
  * 2^(3n/2), np, npt               ReadIOFile      :74-76
"""
from __future__ import annotations

import dataclasses
import re

import numpy as np

__all__ = [
    "EcgSystem",
    "transposition_matrices",
    "parse_young_operator",
    "expand_yhy",
    "make_system",
]


def transposition_matrices(n: int) -> dict:
    """Integer n x n matrices of the pair transpositions P_pq acting on the
    ket's exponent (x -> P^T x P), indexed by particle numbers (1-based).

    Follows src/workproc.f90:673-698.  Particle 1 is the reference particle
    (the origin of the internal coordinates).  P_1i (i >= 2) is the unit
    matrix with column i-1 set to -1; P_ij and P_ji (2 <= i < j) swap rows and
    columns i-1 and j-1.  As in the reference, P_i1 (i >= 2) and P_11 are
    left as unit matrices.
    """
    npart = n + 1
    eye = np.eye(n, dtype=np.int64)
    tr = {}
    for i in range(1, npart + 1):
        for j in range(1, npart + 1):
            tr[(i, j)] = eye.copy()
    for i in range(2, npart + 1):
        tr[(1, i)][:, i - 2] = -1
    for i in range(2, npart + 1):
        for j in range(i + 1, npart + 1):
            for key in ((i, j), (j, i)):
                m = tr[key]
                m[i - 2, i - 2] = 0
                m[j - 2, j - 2] = 0
                m[j - 2, i - 2] = 1
                m[i - 2, j - 2] = 1
    return tr


_TERM_RE = re.compile(r"([+-])(\d*)((?:P\d\d)*)")


def parse_young_operator(text: str) -> list:
    """Split a Young operator string such as ``(1+P23)``, ``(P11)``,
    ``(1+P12)(1+P34)`` or ``2(1-P13)*(1+P23)`` into factors.

    Returns a list of factors; each factor is a list of terms
    ``(coefficient, [(p, q), ...])`` where the transpositions are listed in
    the order written.  Spaces and ``*`` are dropped and a leading ``+`` is
    implied, as in src/workproc.f90:700-878.  Raises ``ValueError`` on the
    inputs the reference rejects (foreign symbols, unbalanced brackets).
    """
    s = text.replace(" ", "").replace("*", "")
    if not s:
        raise ValueError("empty Young operator string")
    if re.search(r"[^0-9P+\-()]", s):
        raise ValueError("Young operator contains wrong symbols: %r" % text)
    if s.count("(") != s.count(")"):
        raise ValueError("unbalanced brackets in Young operator: %r" % text)
    # Factors: bracketed groups, plus any bare text between/around them.
    chunks = []
    pos = 0
    for m in re.finditer(r"\(([^()]*)\)", s):
        if m.start() > pos:
            chunks.append(s[pos:m.start()])
        chunks.append(m.group(1))
        pos = m.end()
    if pos < len(s):
        chunks.append(s[pos:])
    factors = []
    for ch in chunks:
        if ch[0] not in "+-":
            ch = "+" + ch
        terms = []
        consumed = 0
        for m in _TERM_RE.finditer(ch):
            if m.start() != consumed:
                raise ValueError("cannot parse factor %r" % ch)
            consumed = m.end()
            if m.group(0) in ("+", "-"):
                raise ValueError("empty term in factor %r" % ch)
            sign = 1 if m.group(1) == "+" else -1
            coeff = sign * (int(m.group(2)) if m.group(2) else 1)
            perms = [(int(t[1]), int(t[2])) for t in re.findall(r"P\d\d", m.group(3))]
            terms.append((coeff, perms))
        if consumed != len(ch):
            raise ValueError("cannot parse factor %r" % ch)
        factors.append(terms)
    return factors


def _merge(terms, n):
    """Add coefficients of identical matrices, drop zero terms; order of
    first appearance is kept (src/workproc.f90:1000-1013)."""
    out = []
    for c, m in terms:
        for k, (c2, m2) in enumerate(out):
            if np.array_equal(m, m2):
                out[k] = (c2 + c, m2)
                break
        else:
            out.append((c, m.copy()))
    return [(c, m) for c, m in out if c != 0]


def _expand(factors, tr, n):
    """Expand a product of factors into merged (coeff, matrix) terms.
    A word P_a P_b ... maps to the matrix product in the written order."""
    cur = [(1, np.eye(n, dtype=np.int64))]
    for fac in factors:
        nxt = []
        for c0, m0 in cur:
            for c1, perms in fac:
                m1 = np.eye(n, dtype=np.int64)
                for pq in perms:
                    m1 = m1 @ tr[pq]
                nxt.append((c0 * c1, m0 @ m1))
        cur = _merge(nxt, n)
    return cur


def expand_yhy(young: str, n: int):
    """Simplified Y and Y^dagger Y operators as integer matrices.

    Y^dagger is Y with the order of factors and of the transpositions inside
    every term reversed (src/workproc.f90:880-908).  Y^dagger Y is expanded
    and identical matrices are merged (src/workproc.f90:1053-1202).

    The reference multiplies the matrices of a word in reversed order; for
    Y^dagger Y this gives the same set of (matrix, coefficient) pairs,
    because Y^dagger Y is self-adjoint: the coefficient of a group element
    equals that of its inverse.  Only the order in which terms are listed
    may differ, which changes sums by rounding only.

    Returns ``(ymats, ycoef, yhymats, yhycoef)`` with matrices stacked as
    (T, n, n) int arrays.
    """
    tr = transposition_matrices(n)
    for fac in parse_young_operator(young):
        for _, perms in fac:
            for p, q in perms:
                if not (1 <= p <= n + 1 and 1 <= q <= n + 1):
                    raise ValueError("transposition P%d%d outside 1..%d" % (p, q, n + 1))
    factors = parse_young_operator(young)
    y_terms = _expand(factors, tr, n)
    dag = [[(c, list(reversed(perms))) for c, perms in fac] for fac in reversed(factors)]
    yhy_terms = _expand(dag + factors, tr, n)
    ym = np.stack([m for _, m in y_terms]).astype(np.int32)
    yc = np.array([c for c, _ in y_terms], dtype=np.int32)
    hm = np.stack([m for _, m in yhy_terms]).astype(np.int32)
    hc = np.array([c for c, _ in yhy_terms], dtype=np.int32)
    return ym, yc, hm, hc


@dataclasses.dataclass
class EcgSystem:
    """Constant inputs of the pair kernels for one physical system."""

    masses: np.ndarray          # (n+1,) particle masses, particle 1 first
    charges: np.ndarray         # (n+1,) particle charges
    young: str                  # Young operator string as in inout.txt
    n: int = 0                  # pseudo-particles
    mass_matrix: np.ndarray = None   # (n, n)
    q0: float = 0.0             # charge of the reference particle
    q: np.ndarray = None        # (n,) pseudo-charges
    yhy_mats: np.ndarray = None  # (T, n, n) int32
    yhy_coef: np.ndarray = None  # (T,) int32

    @property
    def np_(self) -> int:
        return self.n * (self.n + 1) // 2

    @property
    def npt(self) -> int:
        return self.n * (self.n + 1)

    @property
    def T(self) -> int:
        return int(self.yhy_coef.shape[0])

    @property
    def two_3n2(self) -> float:
        # src/workproc.f90:76 -- 2**((3*n)/2.0)
        return 2.0 ** ((3 * self.n) / 2.0)


def make_system(masses, charges, young: str) -> EcgSystem:
    """Build the constant inputs exactly as ProgramDataInit does.

    Mass matrix: M_ij = 1/(2 m_1) for all i, j plus 1/(2 m_{i+1}) on the
    diagonal (src/workproc.f90:606-610).  Pseudo-charges: q0 is the charge
    of particle 1 and q_i that of particle i+1 (src/workproc.f90:91).
    """
    masses = np.asarray(masses, dtype=np.float64)
    charges = np.asarray(charges, dtype=np.float64)
    if masses.shape != charges.shape or masses.ndim != 1 or masses.size < 2:
        raise ValueError("masses and charges must be equal-length vectors of >= 2 particles")
    n = masses.size - 1
    if n > 5:
        raise ValueError("at most 5 pseudo-particles are supported")
    M = np.full((n, n), 0.5 / masses[0])
    for i in range(n):
        M[i, i] += 0.5 / masses[i + 1]
    _, _, hm, hc = expand_yhy(young, n)
    return EcgSystem(masses=masses, charges=charges, young=young, n=n,
                     mass_matrix=M, q0=float(charges[0]), q=charges[1:].copy(),
                     yhy_mats=hm, yhy_coef=hc)
