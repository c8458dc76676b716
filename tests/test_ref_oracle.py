"""The oracle against the reference's own pair routines (CPU, no GPU).

oracle/_ref/libref_matelem.so is MatrixElements and
MatrixElementsForExpcVals (src/matelem.f90:12-750, 949-1983) compiled from
the reference sources by `make -C oracle ref`; oracle/ref.py loads it.
These tests compare, pair by pair on seeded bases of n = 1, 2 and 3
systems with B != 0:
  * H, S and both gradient blocks Dk, Dl of every symmetry term and every
    gradflag (0-3) with the oracle's restatement, to the rounding
    sensitivity of the pair (tests/helpers.py);
  * the term sums ComputeMatElem(AndDeriv) form from them;
  * the expectation-value operators 1/r^2, 1/r, r, r^2, delta(r) of every
    pair (a, b >= a) and H, S, T, V, and the mass-velocity and Darwin pair
    values, with the oracle's expc_pair / relat_pair.
The library exists only where the reference tree was present to build it;
elsewhere the tests skip.
"""
from __future__ import annotations

import numpy as np
import pytest

from ecg_nonbo import named_system, seeded_basis, SEED_BASE
from ecg_nonbo.system import make_system
from oracle import oracle, ref

from helpers import EPS, pair_sensitivity, close

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (reference tree absent)")

SYSTEMS = ("one", "hminus", "hdplus", "h2", "li")
NB = 5


def _system(name):
    if name == "one":
        return make_system([3.0, 1.0], [1.0, -1.0], "(P11)")
    return named_system(name)


def _pairs(K):
    return [(a, b) for a in range(K) for b in range(a + 1)]


@pytest.mark.parametrize("name", SYSTEMS)
def test_pair_terms_against_reference(name):
    sy = _system(name)
    ref.setup(sy)
    p = seeded_basis(sy.n, NB, SEED_BASE + 6100 + len(name))
    worst = 0.0
    for t in range(sy.T):
        for a, b in _pairs(NB):
            (H, S, Dk, Dl), sens, mag = pair_sensitivity(sy, p[a], p[b], 3, term=t, trials=4)
            for g in (0, 1, 2, 3):
                Hr, Sr, Dkr, Dlr = ref.matrix_elements(sy, p[a], p[b], sy.yhy_mats[t], g)
                checks = [(H, Hr, 0), (S, Sr, 1)]
                if g in (1, 3):
                    checks.append((Dk, Dkr, 2))
                else:
                    assert not np.any(Dkr), (name, t, a, b, g)
                if g in (2, 3):
                    checks.append((Dl, Dlr, 3))
                else:
                    assert not np.any(Dlr), (name, t, a, b, g)
                for x, y, i in checks:
                    ok, d, tol = close(x, y, mag[i], sens[i])
                    assert ok, (name, t, a, b, g, i, d, tol)
                    worst = max(worst, d / tol)
    print(name, "worst deviation / tolerance", worst)


@pytest.mark.parametrize("name", SYSTEMS)
def test_term_sums_against_reference(name):
    sy = _system(name)
    ref.setup(sy)
    p = seeded_basis(sy.n, NB, SEED_BASE + 6200 + len(name))
    for a, b in _pairs(NB):
        (H, S, Dk, Dl), sens, mag = pair_sensitivity(sy, p[a], p[b], 3, trials=4)
        Hr, Sr, Dkr, Dlr = ref.pair_sum(sy, p[a], p[b], 3)
        for x, y, i in ((H, Hr, 0), (S, Sr, 1), (Dk, Dkr, 2), (Dl, Dlr, 3)):
            ok, d, tol = close(x, y, mag[i], sens[i])
            assert ok, (name, a, b, i, d, tol)


def _expc_sensitivity(sy, pk, pl, t, trials=4, seed=0):
    """Spread of the oracle's expc_pair and relat_pair values when both
    parameter vectors move by a few units in the last place: the rounding
    sensitivity of each value (large where terms cancel, e.g. a
    mass-velocity value that is small against its parts)."""
    rng = np.random.default_rng(seed)
    o0 = oracle.expc_pair(sy, pk, pl, t)
    r0 = np.array(oracle.relat_pair(sy, pk, pl, t, sy.masses))
    so, sr = np.zeros(o0.size), np.zeros(2)
    for _ in range(trials):
        a = pk * (1 + 4 * EPS * rng.standard_normal(pk.shape))
        b = pl * (1 + 4 * EPS * rng.standard_normal(pl.shape))
        so = np.maximum(so, np.abs(oracle.expc_pair(sy, a, b, t) - o0))
        sr = np.maximum(sr, np.abs(np.array(oracle.relat_pair(sy, a, b, t, sy.masses)) - r0))
    return o0, r0, so, sr


@pytest.mark.parametrize("name", SYSTEMS)
def test_expectation_operators_against_reference(name):
    sy = _system(name)
    n = sy.n
    ref.setup(sy)
    p = seeded_basis(n, NB, SEED_BASE + 6300 + len(name))
    ident = np.eye(n, dtype=np.int32)
    worst = 0.0
    for t in range(sy.T):
        for a, b in _pairs(NB):
            o, rl, so, sr = _expc_sensitivity(sy, p[a], p[b], t)
            r = ref.expc_pair(sy, p[a], p[b], ident, sy.yhy_mats[t])
            # the operators as the driver gathers them, lower triangle (b, a)
            want = []
            for key in ("rm2", "rm", "r", "r2", "delta"):
                want += [r[key][j, i] for i in range(n) for j in range(i, n)]
            want += [r["H"], r["S"], r["T"], r["V"], r["MV"], r["Darwin"]]
            want = np.array(want)
            got = np.concatenate([o, rl])
            sens = np.concatenate([so, sr])
            tol = np.maximum(1e-12 * np.abs(want), 16 * sens)
            dev = np.abs(got - want)
            assert np.all(dev <= tol), (name, t, a, b, int(np.argmax(dev / tol)), np.max(dev / tol))
            worst = max(worst, np.max(dev / tol))
    print(name, "worst deviation / tolerance", worst)


@pytest.mark.parametrize("name,K", [("hminus", 6), ("h2", 5), ("li", 5)])
def test_expectation_contraction_against_reference(name, K):
    """The oracle's expectation_values and relativistic drivers against the
    reference's per-pair values contracted as the driver does
    (src/workproc.f90:8158-8235): over pairs i >= j, factor 1/diagS_i or
    2/sqrt(diagS_i diagS_j), per term with its coefficient, the real part
    of conj(c_i) c_j O_ij.  Tolerance: the sum of each pair's error bar
    (16 x its sensitivity, at least 1e-12 |O|) with the same weights."""
    sy = _system(name)
    n = sy.n
    ref.setup(sy)
    p = seeded_basis(n, K, SEED_BASE + 6400 + K)
    rng = np.random.Generator(np.random.PCG64(6401))
    c = rng.normal(size=K) + 1j * rng.normal(size=K)
    _, _, _, dS = oracle.assemble(sy, p, "G")
    ident = np.eye(n, dtype=np.int32)
    nv = oracle.n_expectation_values(n) + 2
    acc = np.zeros(nv)
    err = np.zeros(nv)
    for i in range(K):
        for j in range(i + 1):
            f = (1.0 if i == j else 2.0) / np.sqrt(dS[i] * dS[j])
            w = np.conj(c[i]) * c[j]
            for t in range(sy.T):
                r = ref.expc_pair(sy, p[i], p[j], ident, sy.yhy_mats[t])
                O = []
                for key in ("rm2", "rm", "r", "r2", "delta"):
                    O += [r[key][b, a] for a in range(n) for b in range(a, n)]
                O = np.array(O + [r["H"], r["S"], r["T"], r["V"], r["MV"], r["Darwin"]])
                _, _, so, sr = _expc_sensitivity(sy, p[i], p[j], t, trials=2)
                wt = f * sy.yhy_coef[t]
                acc += wt * (w * O).real
                err += abs(wt) * abs(w) * np.maximum(16 * np.concatenate([so, sr]), 1e-12 * np.abs(O))
    vals, _ = oracle.expectation_values(sy, p, c, dS)
    rel, _ = oracle.relativistic(sy, p, c, dS, sy.masses)
    got = np.concatenate([vals, rel])
    dev = np.abs(got - acc)
    assert np.all(dev <= err + 1e-300), (name, int(np.argmax(dev / err)), np.max(dev / err))
    print(name, "worst deviation / tolerance", np.max(dev / err))


# ----------------------------------------------------- Drachmanized forms
E_CURR = -1.1   # the current energy the Drachmanized forms take (Glob_CurrEnergy)


def _drach_sensitivity(sy, pk, pl, t, trials=4, seed=0):
    rng = np.random.default_rng(seed)
    o0 = oracle.drach_pair(sy, pk, pl, t, sy.masses, E_CURR)
    s = np.zeros(o0.size)
    for _ in range(trials):
        a = pk * (1 + 4 * EPS * rng.standard_normal(pk.shape))
        b = pl * (1 + 4 * EPS * rng.standard_normal(pl.shape))
        s = np.maximum(s, np.abs(oracle.drach_pair(sy, a, b, t, sy.masses, E_CURR) - o0))
    return o0, s


def _ref_drach(r, n):
    d = [r["drach_delta"][j, i] for i in range(n) for j in range(i, n)]
    return np.array(d + [r["OO"], r["drach_MV"], r["drach_Darwin"]])


@pytest.mark.parametrize("name", SYSTEMS)
def test_drachmanized_pairs_against_reference(name):
    """The oracle's Drachmanized delta, orbit-orbit, Drachmanized
    mass-velocity and Darwin pair values (src/matelem.f90:1276-1478,
    1669-1933) against the compiled reference, every term, to each value's
    rounding bound.  At n = 3 this includes the entries of the 1/(r_ij r_pq)
    table that read the reference's never-written TrCJCJ entries (zero in
    its static storage)."""
    sy = _system(name)
    n = sy.n
    ref.setup(sy, E_CURR)
    p = seeded_basis(n, NB, SEED_BASE + 6700 + len(name))
    ident = np.eye(n, dtype=np.int32)
    worst = 0.0
    for t in range(sy.T):
        for a, b in _pairs(NB):
            o, sens = _drach_sensitivity(sy, p[a], p[b], t)
            want = _ref_drach(ref.expc_pair(sy, p[a], p[b], ident, sy.yhy_mats[t]), n)
            # these values are differences of larger terms (E <1/r>, the Coulomb sums, <r'Xr/r>);
            # the rounding of their parts reaches a few 1e-12 of the value (measured worst 5e-12)
            tol = np.maximum(2e-11 * np.abs(want), 16 * sens)
            dev = np.abs(o - want)
            assert np.all(dev <= tol), (name, t, a, b, int(np.argmax(dev / tol)), np.max(dev / tol))
            worst = max(worst, np.max(dev / tol))
    print(name, "worst deviation / tolerance", worst)


@pytest.mark.parametrize("name", SYSTEMS)
def test_drachmanized_mass_velocity_equals_darwin(name):
    """The reference computes the Drachmanized mass-velocity and Darwin pair
    values with two different sets of formulas (:1669-1865 through
    ME_rXr_over_rij and traces, :1890-1933 through ME_dXd, ME_dXd_dYd and
    ME_dXd_over_rij) that give the same number: an invariant that pins the
    two restatements against each other."""
    sy = _system(name)
    p = seeded_basis(sy.n, NB, SEED_BASE + 6800 + len(name))
    m = sy.n * (sy.n + 1) // 2
    for t in range(sy.T):
        for a, b in _pairs(NB):
            o, sens = _drach_sensitivity(sy, p[a], p[b], t)
            tol = max(2e-11 * abs(o[m + 1]), 16 * (sens[m + 1] + sens[m + 2]))
            assert abs(o[m + 1] - o[m + 2]) <= tol, (name, t, a, b, o[m + 1], o[m + 2], tol)


@pytest.mark.parametrize("name,K", [("hminus", 5), ("h2", 4), ("li", 4)])
def test_drachmanized_contraction_against_reference(name, K):
    """oracle.drachmanized against the reference's per-pair values
    contracted as the driver does (src/workproc.f90:8158-8235)."""
    sy = _system(name)
    n = sy.n
    ref.setup(sy, E_CURR)
    p = seeded_basis(n, K, SEED_BASE + 6900 + K)
    rng = np.random.Generator(np.random.PCG64(6901))
    c = rng.normal(size=K) + 1j * rng.normal(size=K)
    _, _, _, dS = oracle.assemble(sy, p, "G")
    ident = np.eye(n, dtype=np.int32)
    nv = oracle.n_drach_values(n)
    acc = np.zeros(nv)
    err = np.zeros(nv)
    for i in range(K):
        for j in range(i + 1):
            f = (1.0 if i == j else 2.0) / np.sqrt(dS[i] * dS[j])
            w = np.conj(c[i]) * c[j]
            for t in range(sy.T):
                O = _ref_drach(ref.expc_pair(sy, p[i], p[j], ident, sy.yhy_mats[t]), n)
                _, sens = _drach_sensitivity(sy, p[i], p[j], t, trials=2)
                wt = f * sy.yhy_coef[t]
                acc += wt * (w * O).real
                err += abs(wt) * abs(w) * np.maximum(16 * sens, 2e-11 * np.abs(O))
    vals, _ = oracle.drachmanized(sy, p, c, dS, sy.masses, E_CURR)
    dev = np.abs(vals - acc)
    assert np.all(dev <= err + 1e-300), (name, int(np.argmax(dev / err)), np.max(dev / err))
