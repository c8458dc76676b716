"""CPU checks that pin the oracle (no GPU).

  * the reference's known answers: CURRENT_ENERGY of the six HD+ example
    inputs (n = 2, K = 500/1000/1500, roots 1 and 2), reproduced by the
    oracle's restatement of ComputeMatElem + GHEPIIS at the reference's own
    shift CURRENT_ENERGY * INVITPARAMETER;
  * two-body (n = 1) S, T, V against radial quadrature of the integrals;
  * n = 2, 3 pairs with B != 0 and non-identity P against an independent
    40-digit mpmath evaluation of the closed forms;
  * the analytic per-pair gradients (Dk, Dl) against 45-digit central
    differences of that evaluation;
  * symmetries that a wrong sign, index or transposition would break;
  * the EnergyGB gradient contraction against finite differences of E.
"""
from __future__ import annotations

import os

import numpy as np
import pytest
import scipy.linalg as sl

from ecg_nonbo import (expand_yhy, make_system, named_system, read_inout, seeded_basis, SEED_BASE,
                       transposition_matrices, parse_young_operator)
from helpers import pair_sensitivity
from oracle import oracle
import mp_ecg

FIXTURES = ["inout_1_500", "inout_2_500", "inout_1_1000", "inout_2_1000", "inout_1_1500", "inout_2_1500"]


# ----------------------------------------------------------- symmetry set-up
def test_transposition_matrices():
    tr = transposition_matrices(3)
    np.testing.assert_array_equal(tr[(1, 2)], [[-1, 0, 0], [-1, 1, 0], [-1, 0, 1]])
    np.testing.assert_array_equal(tr[(3, 4)], [[1, 0, 0], [0, 0, 1], [0, 1, 0]])
    np.testing.assert_array_equal(tr[(2, 4)], [[0, 0, 1], [0, 1, 0], [1, 0, 0]])
    # as in the reference, P_i1 and P_11 stay unit matrices
    np.testing.assert_array_equal(tr[(2, 1)], np.eye(3))
    for key in [(1, 2), (1, 3), (2, 3), (3, 4), (1, 4)]:
        m = tr[key]
        np.testing.assert_array_equal(m @ m, np.eye(3))      # transpositions are involutions


@pytest.mark.parametrize("young,n,expect", [
    ("(P11)", 2, {1: [np.eye(2)]}),
    # (1+P23)^dagger (1+P23) = 2 (1 + P23)
    ("(1+P23)", 2, {2: [np.eye(2), [[0, 1], [1, 0]]]}),
    # (1+P12)(1+P34) for H2: 4 (1 + P12 + P34 + P12 P34)
    ("(1+P12)(1+P34)", 3, {4: [np.eye(3), [[1, 0, 0], [0, 0, 1], [0, 1, 0]], [[-1, 0, 0], [-1, 1, 0], [-1, 0, 1]],
                               [[-1, 0, 0], [-1, 0, 1], [-1, 1, 0]]]}),
])
def test_yhy_hand_derived(young, n, expect):
    _, _, mats, coef = expand_yhy(young, n)
    got = sorted((int(c), tuple(np.asarray(m).ravel())) for c, m in zip(coef, mats))
    want = sorted((c, tuple(np.asarray(m, dtype=int).ravel())) for c, ms in expect.items() for m in ms)
    assert got == want


def test_yhy_doublet_lithium():
    # (1+P23)(1-P24)(1-P24)(1+P23) = 2 [2 + 2 P23 - P24 - P23 P24 - P24 P23 - P34]
    _, _, mats, coef = expand_yhy("(1-P24)(1+P23)", 3)
    assert sorted(coef.tolist()) == [-2, -2, -2, -2, 4, 4]
    assert sum(coef) == 0          # antisymmetriser in P24 annihilates the trace


def test_young_parser_errors():
    with pytest.raises(ValueError):
        parse_young_operator("(1+P23")
    with pytest.raises(ValueError):
        parse_young_operator("(1+Q23)")


def test_mass_matrix_and_charges():
    sy = make_system([3670.4829652, 1836.15267261, 1.0], [1.0, 1.0, -1.0], "(P11)")
    M = sy.mass_matrix
    assert M[0, 1] == M[1, 0] == 0.5 / 3670.4829652
    assert M[1, 1] == 0.5 / 3670.4829652 + 0.5 / 1.0
    assert sy.q0 == 1.0 and list(sy.q) == [1.0, -1.0]
    assert sy.two_3n2 == 8.0


def test_inout_reader(golden_dir):
    f = read_inout(os.path.join(golden_dir, "hdplus", "inout_1_500.txt"))
    assert f.basis_size == 500 and f.params.shape == (500, 6)
    assert f.current_energy == -0.5978979662615874
    assert f.which_eigenvalue == 1 and f.eigval_tolerance == 1e-12
    np.testing.assert_allclose(f.params[0, :2], [-0.6687331086620485, 0.6208759055559284e-01])
    np.testing.assert_allclose(f.params[-1, -1], -0.4444638634946612)


# --------------------------------------------------------- known answers
@pytest.mark.parametrize("fname", FIXTURES)
def test_fixture_energies(golden_dir, fname):
    f = read_inout(os.path.join(golden_dir, "hdplus", fname + ".txt"))
    K = f.basis_size
    sigma = f.current_energy * f.invit_parameter
    H, S, dH, dS = oracle.assemble(f.system, f.params, "I", sigma)
    M = H.copy(order="F")
    invD = np.zeros(K, np.complex128)
    err, lam, x, rel, nit = oracle.ghepiis(1, M, invD, S.copy(order="F"), sigma, np.ones(K, np.complex128),
                                           f.eigval_tolerance)
    assert err in (0, 2)
    assert abs(lam - f.current_energy) <= 1e-10, (lam, f.current_energy)
    # Sylvester inertia: exactly which-1 eigenvalues lie below the shift
    assert int(np.sum((1.0 / invD).real < 0)) == f.which_eigenvalue - 1


def test_fixture_energy_dense_check(golden_dir):
    """The same root from LAPACK on the oracle's full matrices (K = 500,
    where S is conditioned well enough, kappa ~ 3e9, for the dense solver)."""
    f = read_inout(os.path.join(golden_dir, "hdplus", "inout_1_500.txt"))
    H, S, dH, dS = oracle.assemble(f.system, f.params, "G")
    Hf, Sf = oracle.full_hermitian(H, S, dH)
    w = sl.eigh(Hf, Sf, eigvals_only=True, subset_by_index=[0, 0])
    assert abs(w[0] - f.current_energy) <= 1e-9


# ---------------------------------------------------- closed forms, n = 1
@pytest.mark.parametrize("pk,pl", [([0.8, 0.3], [1.7, -0.2]), ([0.05, 0.01], [3.0, 1.5]), ([1.2, 0.0], [1.2, 0.0])])
def test_two_body_quadrature(pk, pl):
    m0, m1, q0, q1 = 1836.15267, 1.0, 1.0, -1.0
    sy = make_system([m0, m1], [q0, q1], "(P11)")
    H, S, _, _ = oracle.pair(sy, np.array(pk), np.array(pl), 0)
    So, To, Vo = mp_ecg.n1_quadrature(m0, m1, q0, q1, pk, pl)
    assert abs(S - So) <= 1e-13 * abs(So)
    assert abs(H - (To + Vo)) <= 1e-12 * (abs(To) + abs(Vo))


# ------------------------------------------------ closed forms, n = 2, 3
@pytest.mark.parametrize("name", ["hminus", "h2", "li"])
def test_pairs_vs_mpmath(name):
    sy = named_system(name)
    B = seeded_basis(sy.n, 8, SEED_BASE + 11)
    for t in range(sy.T):
        for i, j in [(0, 1), (2, 5), (7, 3), (4, 4)]:
            (H, S, _, _), sens, mag = pair_sensitivity(sy, B[i], B[j], 0, term=t)
            Hm, Sm = mp_ecg.pair_hs(sy, B[i], B[j], sy.yhy_mats[t])
            assert abs(S - complex(Sm)) <= max(1e-12 * mag[1], 16 * sens[1]), (t, i, j)
            assert abs(H - complex(Hm)) <= max(1e-12 * mag[0], 16 * sens[0]), (t, i, j)


@pytest.mark.parametrize("name,t", [("hminus", 1), ("h2", 2), ("h2", 3), ("li", 4)])
def test_pair_gradients_vs_mpmath_fd(name, t):
    sy = named_system(name)
    B = seeded_basis(sy.n, 4, SEED_BASE + 13)
    P = sy.yhy_mats[t]
    (H, S, Dk, Dl), sens, mag = pair_sensitivity(sy, B[0], B[1], 3, term=t)
    fk = np.array(mp_ecg.pair_grad_fd(sy, B[0], B[1], P, "k"))
    fl = np.array(mp_ecg.pair_grad_fd(sy, B[0], B[1], P, "l"))
    np_ = sy.np_
    # FD order: (dH/dvechL, dH/dvechB, dS/dvechL, dS/dvechB), as the packing
    for got, want, q in ((Dk, fk, 2), (Dl, fl, 3)):
        err = np.max(np.abs(got - want))
        assert err <= max(1e-11 * mag[q], 16 * sens[q]), (err, mag[q], sens[q])


def test_gradflag_subsets_agree():
    sy = named_system("h2")
    B = seeded_basis(3, 2, SEED_BASE + 17)
    _, _, Dk3, Dl3 = oracle.pair(sy, B[0], B[1], 3, term=2)
    _, _, Dk1, _ = oracle.pair(sy, B[0], B[1], 1, term=2)
    _, _, _, Dl2 = oracle.pair(sy, B[0], B[1], 2, term=2)
    np.testing.assert_array_equal(Dk1, Dk3)
    np.testing.assert_array_equal(Dl2, Dl3)


# ------------------------------------------------------------- symmetries
@pytest.mark.parametrize("name", ["hminus", "h2", "li"])
def test_hermiticity_with_permutations(name):
    """For a transposition P of identical particles (P^2 = 1, [H, P] = 0):
    <k|H P|l> = conj(<l|H P|k>) and the same for S."""
    sy = named_system(name)
    B = seeded_basis(sy.n, 6, SEED_BASE + 19)
    tr = transposition_matrices(sy.n)
    pairs = {"hminus": [(2, 3)], "h2": [(1, 2), (3, 4)], "li": [(2, 3), (2, 4), (3, 4)]}[name]
    for pq in pairs:
        sys1 = make_system(sy.masses, sy.charges, "(P%d%d)" % pq)
        for i, j in [(0, 1), (2, 4), (5, 3)]:
            (H1, S1, _, _), s1, m1 = pair_sensitivity(sys1, B[i], B[j], 0)
            (H2, S2, _, _), s2, m2 = pair_sensitivity(sys1, B[j], B[i], 0)
            assert abs(S1 - np.conj(S2)) <= max(1e-12 * m1[1], 16 * (s1[1] + s2[1]))
            assert abs(H1 - np.conj(H2)) <= max(1e-12 * m1[0], 16 * (s1[0] + s2[0]))


def test_scaling_relations():
    """L -> a L, B -> a^2 B in both functions: S unchanged, T x a^2, V x a
    (to the rounding sensitivity of the pair)."""
    sy = named_system("h2")
    B = seeded_basis(3, 2, SEED_BASE + 23)
    m = sy.np_
    a = 1.7
    sc = np.concatenate([np.full(m, a), np.full(m, a * a)])
    kin = make_system(sy.masses, sy.charges * 0.0, sy.young)
    for t in range(sy.T):
        (H1, S1, _, _), s1, m1 = pair_sensitivity(sy, B[0], B[1], 0, term=t)
        (H2, S2, _, _), s2, m2 = pair_sensitivity(sy, B[0] * sc, B[1] * sc, 0, term=t)
        T1 = oracle.pair(kin, B[0], B[1], 0, term=t)[0]
        T2 = oracle.pair(kin, B[0] * sc, B[1] * sc, 0, term=t)[0]
        V1, V2 = H1 - T1, H2 - T2
        assert abs(S2 - S1) <= max(1e-12 * m1[1], 16 * (s1[1] + s2[1]))
        assert abs(T2 - a * a * T1) <= max(1e-12 * m2[0], 16 * (s2[0] + a * a * s1[0]))
        assert abs(V2 - a * V1) <= max(1e-12 * m2[0], 16 * (s2[0] + a * s1[0]))


def test_overlap_of_function_with_itself():
    sy = named_system("li")
    B = seeded_basis(3, 5, SEED_BASE + 29)
    for k in range(5):
        H, S, _, _ = oracle.pair(sy, B[k], B[k], 0, term=0)   # term 0 is the identity
        assert abs(S - 1.0) <= 1e-9 and abs(H.imag) <= 1e-9 * abs(H)


# ---------------------------------------------------------- energy gradient
def test_energy_gradient_contraction_fd():
    """EnergyGB contraction (stored D) against central differences of the
    energy, K = 16.  The step is relative (1e-4 |p|, floor 1e-6): a plain
    relative 1e-6 step, as in the reference's commented-out check, drowns
    in the eigen solver's rounding (~1e-12 Eh here, ||H|| ~ 1e4) for
    parameters near zero."""
    sy = named_system("hminus")
    K = 16
    p = seeded_basis(sy.n, K, SEED_BASE + 31)

    def energy(params, with_grad=False):
        H, S, dH, dS, D = oracle.assemble_deriv(sy, params, 0, "G")
        Hf, Sf = oracle.full_hermitian(H, S, dH)
        w, v = sl.eigh(Hf, Sf, subset_by_index=[0, 0])
        if not with_grad:
            return w[0]
        return w[0], oracle.gradient(D, v[:, 0], w[0])

    E, g = energy(p, True)
    rng = np.random.default_rng(3)
    for q in rng.choice(p.size, 10, replace=False):
        k, m = divmod(int(q), sy.npt)
        h = 1e-4 * max(abs(p[k, m]), 1e-2)
        pp = p.copy()
        pp[k, m] += h
        pm = p.copy()
        pm[k, m] -= h
        fd = (energy(pp) - energy(pm)) / (2 * h)
        assert abs(fd - g[k * sy.npt + m]) <= 1e-6 * max(1.0, abs(fd)) + 1e-12 / h, (k, m, fd, g[k * sy.npt + m])


def test_ghepiis_matches_dense_solver():
    sy = named_system("hminus")
    K = 30
    p = seeded_basis(sy.n, K, SEED_BASE + 37)
    H, S, dH, dS = oracle.assemble(sy, p, "G")
    Hf, Sf = oracle.full_hermitian(H, S, dH)
    w = sl.eigh(Hf, Sf, eigvals_only=True)
    sigma = w[0] - 1e-3
    Hi, Si, _, _ = oracle.assemble(sy, p, "I", sigma)
    err, lam, x, _, _ = oracle.ghepiis(1, Hi.copy(order="F"), np.zeros(K, np.complex128), Si.copy(order="F"),
                                       sigma, np.ones(K, np.complex128), 1e-12, 50)
    assert abs(lam - w[0]) <= 1e-11 * max(1.0, abs(w[0]))
    assert abs(np.vdot(x, Sf @ x).real - 1.0) <= 1e-10


def test_flop_counts_are_structural():
    """The operation count of the reference algorithm does not depend on the
    parameter values, and grows energy < bra gradient < both gradients."""
    sy = named_system("h2")
    B = seeded_basis(3, 4, SEED_BASE + 41)
    for t in range(sy.T):
        f0 = {oracle.flops_per_pair(sy, B[i], B[j], 0, term=t) for i, j in [(0, 1), (2, 3)]}
        f1 = {oracle.flops_per_pair(sy, B[i], B[j], 1, term=t) for i, j in [(0, 1), (2, 3)]}
        f3 = {oracle.flops_per_pair(sy, B[i], B[j], 3, term=t) for i, j in [(0, 1), (2, 3)]}
        assert len(f0) == len(f1) == len(f3) == 1
        assert min(f0) < min(f1) < min(f3)


# ----------------------------------------------------------- overlap penalty
def _duplicated_basis(sy, K, seed, eps, real):
    """Functions K/2.. are perturbed copies of functions 0.., so that the
    pairs (K/2 + i, i) have large normalised overlaps."""
    p = seeded_basis(sy.n, K, seed)
    rng = np.random.default_rng(seed + 1)
    h = K // 2
    p[h:] = p[:K - h] * (1.0 + eps * rng.standard_normal((K - h, p.shape[1])))
    if real:
        p[:, sy.np_:] = 0.0
    return p


def _threshold_between(S, hits):
    """t2 halfway between the largest |s|^2 of the other pairs and the
    smallest of the wanted ones (asserting that there is such a gap)."""
    K = S.shape[0]
    a2 = np.abs(S) ** 2
    want = [a2[j, i] for j, i in hits]
    other = [a2[j, i] for j in range(K) for i in range(j) if (j, i) not in hits]
    assert max(other) < min(want), (max(other), min(want))
    return 0.5 * (max(other) + min(want))


def test_overlap_penalty_closed_form():
    """At t2 = 0 every pair counts, so P = b * (sum of |s_ji|^2 over the
    strict lower triangle rows j > nfru): a Frobenius norm of a block of
    the stored S, which pins the pair ranges of both loops."""
    sy = named_system("hminus")
    K = 12
    p = seeded_basis(sy.n, K, SEED_BASE + 61)
    H, S, dH, dS = oracle.assemble(sy, p, "G")
    low = np.tril(np.abs(S) ** 2, -1)
    for nfru in (0, 1, 5, K - 1):
        P = oracle.overlap_penalty(S, nfru, 1.3, 0.0)
        ref = 1.3 * low[nfru:, :].sum()
        assert abs(P - ref) <= 1e-13 * max(1.0, ref), (nfru, P, ref)
    # nothing above a threshold of one
    assert oracle.overlap_penalty(S, 0, 1.3, 0.999999) == 0.0


def _penalty_fd(sy, p, nfru, b, t2, params_to_check, tol):
    def pen(q):
        H, S, dH, dS = oracle.assemble(sy, q, "G")
        return oracle.overlap_penalty(S, nfru, b, t2)

    H, S, dH, dS, D = oracle.assemble_deriv(sy, p, nfru, "G")
    P, g = oracle.overlap_penalty_grad(S, dS, D, nfru, b, t2)
    assert abs(P - pen(p)) <= 1e-14 * max(1.0, P)
    assert P > 0.0
    def cd(k, m, h):
        pp = p.copy()
        pp[k, m] += h
        pm = p.copy()
        pm[k, m] -= h
        return (pen(pp) - pen(pm)) / (2 * h)

    for k, m in params_to_check:
        # central differences at h and h/2, Richardson-extrapolated (O(h^4))
        h = 1e-4 * max(abs(p[k, m]), 1e-2)
        fd = (4.0 * cd(k, m, 0.5 * h) - cd(k, m, h)) / 3.0
        an = g[(k - nfru) * sy.npt + m]
        assert abs(fd - an) <= tol * max(1.0, abs(fd)), (k, m, fd, an)


def test_overlap_penalty_gradient_fd_real():
    """Unit diagonal (one identity term) and real overlaps (B = 0): there
    the reference's gradient terms are the exact derivative of P.  Hits in
    both loops (fixed-optimised and optimised-optimised pairs), all vech L
    parameters of the optimised functions checked."""
    sy = named_system("hdplus")
    K, nfru = 10, 2
    p = _duplicated_basis(sy, K, SEED_BASE + 63, 0.02, real=True)
    H, S, _, _ = oracle.assemble(sy, p, "G")
    hits = [(5 + i, i) for i in range(5)]
    t2 = _threshold_between(S, hits)
    checks = [(k, m) for k in range(nfru, K) for m in range(sy.np_)]
    _penalty_fd(sy, p, nfru, 0.7, t2, checks, 1e-7)


def test_overlap_penalty_gradient_fd_complex_fixed_partner():
    """Complex overlaps (B != 0), unit diagonal, hits only between fixed
    and optimised functions (the reference's first loop): again the exact
    derivative; every L and B parameter of the optimised functions."""
    sy = named_system("hdplus")
    K, nfru = 10, 5
    p = _duplicated_basis(sy, K, SEED_BASE + 65, 0.02, real=False)
    H, S, _, _ = oracle.assemble(sy, p, "G")
    hits = [(5 + i, i) for i in range(5)]
    t2 = _threshold_between(S, hits)
    assert np.max(np.abs(np.tril(S, -1).imag)) > 1e-3
    checks = [(k, m) for k in range(nfru, K) for m in range(sy.npt)]
    _penalty_fd(sy, p, nfru, 0.7, t2, checks, 1e-7)


def test_overlap_penalty_stationary_at_duplicates():
    """Exact copies (|s| = 1, the maximum of |s|^2) make P stationary, with
    two terms (T = 2, so S_kk varies with the parameters) and B != 0: the
    overlap-derivative term and the diagonal normalisation term must
    cancel, which a dropped or mis-weighted normalisation term breaks."""
    sy = named_system("hminus")
    K, nfru = 10, 2
    p = _duplicated_basis(sy, K, SEED_BASE + 67, 0.0, real=False)
    H, S, dH, dS, D = oracle.assemble_deriv(sy, p, nfru, "G")
    assert np.ptp(dS) > 1e-3                      # S_kk is not constant
    hits = [(5 + i, i) for i in range(5)]
    t2 = _threshold_between(S, hits)
    P, g = oracle.overlap_penalty_grad(S, dS, D, nfru, 0.7, t2)
    assert P > 0.0
    # size of the individual terms: 2 pc |D_S| summed
    scale = 2 * 0.7 / (1 - t2) * np.abs(D[sy.npt:]).max()
    assert np.max(np.abs(g)) <= 1e-12 * scale, (np.max(np.abs(g)), scale)
