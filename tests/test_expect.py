"""CPU checks that pin the oracle's expectation values (no GPU).

MatrixElementsForExpcVals (src/matelem.f90:949-1983) gives, per pair and
symmetry term, closed forms of <1/r^2>, <1/r>, <r>, <r^2> and <delta(r)> in
S and tr(C^-1 J); the driver (src/workproc.f90:8158-8300) contracts them
with the eigenvector.  Pinned here without retyping those forms:
  * each operator as an integral, derivative or limit of the overlap with
    C -> C + s J (exp(-s r^2) inserted between the functions), evaluated
    at 30-60 digits with mpmath's own determinant (tests/mp_ecg.py), at
    n = 2 and 3 with B != 0 and non-identity symmetry matrices;
  * n = 1 against radial quadrature of the actual integrals;
  * the contraction: <S> = 1 and <H> = E for the eigenvector of the HD+
    fixture, <H> = <T> + <V>, and the virial ratio -<V>/<T> = 2 of that
    optimised basis.
  * the pair correlation functions CF_q(r) (src/matelem.f90:1934-1946)
    through what fixes them independently of their formula: CF_q(0) is
    <delta(r_q)>, 4 pi int r^2 CF_q dr is <S> and 4 pi int r^4 CF_q dr is
    <r_q^2> (adaptive quadrature of the oracle's CF over r), with random
    complex vectors.
  * the particle densities (src/matelem.f90:1948-1963, vectors Glob_bvc
    from the masses): at n = 1 they are the pair function at a scaled
    radius, Dens_i(r) = CF(r / |b_i|) / |b_i|^3; at n = 2, 3 their
    moments 4 pi int r^2 Dens_i dr = <S> and 4 pi int r^4 Dens_i dr =
    <|b_i . r|^2>, the latter built from the pinned <r_p^2> and
    <|r_p - r_q|^2>.
  * the mass-velocity pair value (src/matelem.f90:1479-1667) against the
    Laplacians of the two Gaussians: at n = 1 the closed form
    <D phi_k|D phi_l> / <phi_k|phi_l> = 60 a^2 b^2 / (a + b)^2, at n = 2, 3
    Gaussian-moment derivatives of det(C + s P + t Q)^(-3/2) at 30 digits
    (mpmath) for every particle; the Darwin pair value against its
    definition through the pinned delta functions.
Not built (and not pinned): the Drachmanized delta functions, the
Drachmanized mass-velocity and Darwin terms and the orbit-orbit term.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from ecg_nonbo import make_system, named_system, read_inout, seeded_basis, SEED_BASE
from oracle import oracle
import mp_ecg


def _pairs_of(n):
    return [(a, b) for a in range(n) for b in range(a, n)]


@pytest.mark.parametrize("name,term", [("hminus", 1), ("h2", 3)])
def test_expc_pair_against_shifted_overlap_integrals(name, term):
    sy = named_system(name)
    p = seeded_basis(sy.n, 6, SEED_BASE + 51)
    pk, pl = p[2], p[4]
    o = oracle.expc_pair(sy, pk, pl, term)
    m = sy.n * (sy.n + 1) // 2
    P = np.asarray(sy.yhy_mats[term])
    for c, (a, b) in enumerate(_pairs_of(sy.n)):
        ref = mp_ecg.expc_by_integrals(sy, pk, pl, P, a, b, dps=24)
        got = [o[q * m + c] for q in range(5)]
        for q in range(5):
            assert abs(got[q] - ref[q]) <= 1e-11 * abs(ref[q]), (name, term, a, b, q, got[q], ref[q])
    H, S = mp_ecg.pair_hs(sy, pk, pl, P)
    assert abs(o[5 * m + 1] - complex(S)) <= 1e-12 * abs(complex(S))
    assert abs(o[5 * m] - complex(H)) <= 1e-11 * abs(complex(H))
    assert abs(o[5 * m] - o[5 * m + 2] - o[5 * m + 3]) <= 1e-14 * abs(o[5 * m])


@pytest.mark.parametrize("pk,pl", [([0.8, 0.3], [1.7, -0.2]), ([0.05, 0.01], [3.0, 1.5]), ([1.2, 0.0], [1.2, 0.0])])
def test_expc_pair_n1_radial_quadrature(pk, pl):
    sy = make_system([1.0, 1836.15267261], [1.0, -1.0], "(P11)")
    o = oracle.expc_pair(sy, np.array(pk), np.array(pl), 0)
    Skk = oracle.expc_pair(sy, np.array(pk), np.array(pk), 0)[6]
    Sll = oracle.expc_pair(sy, np.array(pl), np.array(pl), 0)[6]
    norm = np.sqrt(Skk.real * Sll.real)
    ref = mp_ecg.n1_moments(pk, pl)
    for q in range(5):
        assert abs(o[q] / norm - ref[q]) <= 1e-12 * abs(ref[q]), (q, o[q] / norm, ref[q])


def test_expectation_values_hdplus_fixture(golden_dir):
    """Contraction with the fixture's eigenvector: <S> = 1, <H> = E,
    <H> = <T> + <V>, virial ratio 2, and the proton-deuteron distance."""
    f = read_inout(os.path.join(golden_dir, "hdplus", "inout_1_500.txt"))
    K = f.basis_size
    sigma = f.current_energy * f.invit_parameter
    H, S, dH, dS = oracle.assemble(f.system, f.params, "I", sigma)
    invD = np.zeros(K, np.complex128)
    err, lam, x, rel, nit = oracle.ghepiis(1, H.copy(order="F"), invD, S.copy(order="F"), sigma,
                                           np.ones(K, np.complex128), f.eigval_tolerance)
    Hg, Sg, dHg, dSg = oracle.assemble(f.system, f.params, "G")
    Sf = oracle.full_hermitian(Hg, Sg, dHg)[1]
    x = x / np.sqrt((x.conj() @ Sf @ x).real)
    vals, errs = oracle.expectation_values(f.system, f.params, x, dSg)
    n = f.system.n
    m = n * (n + 1) // 2
    Hx, Sx, Tx, Vx = vals[5 * m:5 * m + 4]
    assert abs(Sx - 1.0) <= 1e-10
    assert abs(Hx - lam) <= 1e-10 * abs(lam)
    assert abs(Hx - Tx - Vx) <= 1e-12
    assert abs(-Vx / Tx - 2.0) <= 1e-6, -Vx / Tx
    # r_1 (particle 2 from particle 1 = the deuteron-proton distance of HD+)
    r12 = vals[2 * m + 0]
    assert 2.0 < r12 < 2.1, r12
    assert np.all(errs >= 0.0) and np.all(errs <= 1e-6 * np.maximum(1.0, np.abs(vals)))


@pytest.mark.parametrize("name,K", [("hminus", 12), ("h2", 10)])
def test_correlation_functions_moments(name, K):
    from scipy.integrate import quad_vec
    sy = named_system(name)
    p = seeded_basis(sy.n, K, SEED_BASE + 71)
    rng = np.random.Generator(np.random.PCG64(72))
    c = rng.normal(size=K) + 1j * rng.normal(size=K)
    _, _, _, dS = oracle.assemble(sy, p, "G")
    vals, _ = oracle.expectation_values(sy, p, c, dS)
    m = sy.n * (sy.n + 1) // 2
    delta, r2, S = vals[4 * m:5 * m], vals[3 * m:4 * m], vals[5 * m + 1]
    cf0 = oracle.correlation_functions(sy, p, c, dS, [0.0])[0]
    np.testing.assert_allclose(cf0, delta, rtol=1e-13, atol=0)
    # moments over r (the map r = x / (1 - x) takes [0, 1) to [0, inf))
    def integrand(x):
        r = x / (1.0 - x)
        jac = 1.0 / (1.0 - x) ** 2
        cf = oracle.correlation_functions(sy, p, c, dS, [r])[0]
        return np.concatenate([4 * np.pi * r ** 2 * cf, 4 * np.pi * r ** 4 * cf]) * jac
    mom, err = quad_vec(integrand, 0.0, 1.0 - 1e-12, epsabs=1e-13, epsrel=1e-11, limit=2000)
    np.testing.assert_allclose(mom[:m], S, rtol=2e-9)
    np.testing.assert_allclose(mom[m:], r2, rtol=2e-9)


def _bvc(masses):
    """Glob_bvc (src/workproc.f90:612-621), columns = particles."""
    ms = np.asarray(masses, dtype=float)
    n = ms.size - 1
    b = np.tile(-ms[1:, None] / ms.sum(), (1, n + 1))
    for i in range(1, n + 1):
        b[i - 1, i] += 1.0
    return b


def test_particle_densities_n1_scaled_pair_function():
    sy = make_system([2.0, 1.0], [1.0, -1.0], "(P11)")   # b = (-1/3, 2/3): both densities resolved
    K = 10
    p = seeded_basis(sy.n, K, SEED_BASE + 73)
    rng = np.random.Generator(np.random.PCG64(74))
    c = rng.normal(size=K) + 1j * rng.normal(size=K)
    _, _, _, dS = oracle.assemble(sy, p, "G")
    grid = np.linspace(0.0, 2.5, 11)
    dens = oracle.particle_densities(sy, p, c, dS, sy.masses, grid)
    b = _bvc(sy.masses)[0]
    for i in range(2):
        cf = oracle.correlation_functions(sy, p, c, dS, grid / abs(b[i]))[:, 0] / abs(b[i]) ** 3
        np.testing.assert_allclose(dens[:, i], cf, rtol=1e-12, atol=1e-14 * np.max(np.abs(cf)))


@pytest.mark.parametrize("name,K", [("hminus", 10), ("h2", 8)])
def test_particle_densities_moments(name, K):
    from scipy.integrate import quad_vec
    sy = named_system(name)
    n = sy.n
    p = seeded_basis(n, K, SEED_BASE + 75)
    rng = np.random.Generator(np.random.PCG64(76))
    c = rng.normal(size=K) + 1j * rng.normal(size=K)
    _, _, _, dS = oracle.assemble(sy, p, "G")
    vals, _ = oracle.expectation_values(sy, p, c, dS)
    m = n * (n + 1) // 2
    r2v, S = vals[3 * m:4 * m], vals[5 * m + 1]
    pairs = [(a, b) for a in range(n) for b in range(a, n)]
    R = np.zeros((n, n))                        # <r_p . r_q>
    for q, (a, b) in enumerate(pairs):
        if a == b:
            R[a, a] = r2v[q]
    for q, (a, b) in enumerate(pairs):
        if a != b:
            R[a, b] = R[b, a] = 0.5 * (R[a, a] + R[b, b] - r2v[q])
    B = _bvc(sy.masses)
    want2 = np.array([B[:, i] @ R @ B[:, i] for i in range(n + 1)])

    def integrand(x):
        r = x / (1.0 - x)
        jac = 1.0 / (1.0 - x) ** 2
        d = oracle.particle_densities(sy, p, c, dS, sy.masses, [r])[0]
        return np.concatenate([4 * np.pi * r ** 2 * d, 4 * np.pi * r ** 4 * d]) * jac
    mom, _ = quad_vec(integrand, 0.0, 1.0 - 1e-12, epsabs=1e-13, epsrel=1e-11, limit=2000)
    np.testing.assert_allclose(mom[:n + 1], S, rtol=2e-9)
    np.testing.assert_allclose(mom[n + 1:], want2, rtol=2e-9)


def test_mass_velocity_n1_closed_form():
    sy = make_system([3.0, 1.0], [1.0, -1.0], "(P11)")
    p = seeded_basis(1, 6, SEED_BASE + 77)
    for k, l in [(0, 0), (1, 3), (4, 2)]:
        mv, _ = oracle.relat_pair(sy, p[k], p[l], 0, sy.masses)
        _, S = oracle.relat_pair(sy, p[k], p[l], 0, sy.masses)[0], oracle.expc_pair(sy, p[k], p[l], 0)[6]
        a = p[k][0] ** 2 - 1j * p[k][1]
        b = p[l][0] ** 2 + 1j * p[l][1]
        want = -(1 / 8) * (1 / 27.0 + 1.0) * 60 * a * a * b * b / (a + b) ** 2
        assert abs(mv / S - want) <= 1e-12 * abs(want), (k, l, mv / S, want)


@pytest.mark.parametrize("name,term", [("hminus", 1), ("h2", 3), ("h2", 0)])
def test_mass_velocity_against_moment_derivatives(name, term):
    sy = named_system(name)
    p = seeded_basis(sy.n, 6, SEED_BASE + 78)
    m = sy.n * (sy.n + 1) // 2
    for k, l in [(2, 4), (5, 1)]:
        mv, dw = oracle.relat_pair(sy, p[k], p[l], term, sy.masses)
        o = oracle.expc_pair(sy, p[k], p[l], term)
        S = o[5 * m + 1]
        want = mp_ecg.mv_ratio_by_derivatives(sy, p[k], p[l], sy.yhy_mats[term], sy.masses)
        assert abs(mv / S - want) <= 1e-10 * abs(want), (name, term, k, l, mv / S, want)
        # Darwin from its definition through the pair delta functions
        delta = o[4 * m:5 * m]
        ms, q0, q = sy.masses, sy.q0, sy.q
        pairs = [(a, b) for a in range(sy.n) for b in range(a, sy.n)]
        d = 0
        for c, (a, b) in enumerate(pairs):
            if a == b:
                d += (1 / ms[0] ** 2 + 1 / ms[a + 1] ** 2) * q0 * q[a] * delta[c]
            else:
                d += (1 / ms[a + 1] ** 2 + 1 / ms[b + 1] ** 2) * q[a] * q[b] * delta[c]
        assert abs(dw - (-np.pi / 2) * d) <= 1e-13 * abs(dw)
