"""GPU parity of the incremental refactorisation (tail update).

Reference behaviour: EnergyIA(Nmin, ...) / EnergyIB (src/workproc.f90:
1954-2154) recompute matrix elements of functions Nmin..K only and call
GHEPIIS with k = Nmin, whose LDLHF completes the factor from row Nmin
(src/linalg.f90:316-455); OptCycleI (src/workproc.f90:5202-5240) calls them
with Nmin = K for every trial point of the function being optimised.  Each
test compares the tail path (info[7] = rows refactored) with a fresh
context's full factorisation of the same parameters, whose own parity with
the oracle is covered elsewhere: energies to 1e-11 relative (a dropped or
mis-signed term of the update moves E by ~1e-3), the inertia exactly, and
gradients to 1e-8 of their largest entry (at least 1e-15).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from ecg_nonbo import named_system, read_inout, seeded_basis, SEED_BASE
from ecg_nonbo.hip import EcgContext

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print("[case]", *a, flush=True)


def _fixture(golden_dir, name):
    return read_inout(os.path.join(golden_dir, "hdplus", name + ".txt"))


def _perturb(p, rows, seed, scale=1e-3):
    q = p.copy()
    rng = np.random.Generator(np.random.PCG64(seed))
    q[rows] *= 1.0 + scale * rng.normal(size=q[rows].shape)
    return q


def _fresh_energy(sy, p, which, sigma):
    ctx = EcgContext(sy)
    ctx.set_params(p)
    rc, E, _, info = ctx.energy(which, sigma, TOL, 20)
    ctx.close()
    assert info[7] == -1
    return E, info


def _fresh_grad(sy, p, nfru, which, sigma, stored=False):
    ctx = EcgContext(sy)
    ctx.set_params(p)
    fn = ctx.energy_grad_stored if stored else ctx.energy_grad
    rc, E, g, info = fn(nfru, which, sigma, TOL, 20)
    ctx.close()
    assert info[7] == -1
    return E, g


def _close(a, b, rel=1e-11):
    assert abs(a - b) <= rel * abs(b), (a, b, a - b)


def _grad_close(g1, g2):
    """1e-8 of the largest entry, floored at 1e-15: at an optimised basis the
    gradient is ~1e-11 while its terms are O(1), so two correct evaluations
    differ by their rounding (~1e-16 of the terms), not by 1e-8 of g."""
    np.testing.assert_allclose(g1, g2, rtol=0, atol=max(1e-8 * np.max(np.abs(g2)), 1e-15))


@pytest.mark.parametrize("r", [1, 3, 17, 64, 65])
def test_tail_energy_matches_full(golden_dir, r):
    say("ecg_energy tail", "inout_1_500", r)
    f = _fixture(golden_dir, "inout_1_500")
    sigma = f.current_energy * f.invit_parameter
    K = f.basis_size
    ctx = EcgContext(f.system)
    ctx.set_params(f.params)
    _, E0, _, info0 = ctx.energy(1, sigma, TOL, 20)
    assert info0[7] == -1
    _close(E0, f.current_energy, 1e-12)
    p = _perturb(f.params, slice(K - r, K), 100 + r)
    ctx.set_params(p)
    _, E1, _, info1 = ctx.energy(1, sigma, TOL, 20)
    assert info1[7] == (r if r <= 64 else -1)
    E2, info2 = _fresh_energy(f.system, p, 1, sigma)
    print("E0", E0, "tail", E1, "full", E2, "diff", E1 - E2)
    _close(E1, E2)
    assert info1[4] == info2[4] == 0
    ctx.close()


def test_tail_unchanged_parameters(golden_dir):
    say("ecg_energy tail", "inout_1_500", "unchanged")
    f = _fixture(golden_dir, "inout_1_500")
    sigma = f.current_energy * f.invit_parameter
    ctx = EcgContext(f.system)
    ctx.set_params(f.params)
    _, E0, _, _ = ctx.energy(1, sigma, TOL, 20)
    ctx.set_params(f.params)
    _, E1, _, info = ctx.energy(1, sigma, TOL, 20)
    assert info[7] == 0
    _close(E1, E0, 1e-13)
    _, E2, _, info = ctx.energy(1, sigma * (1 + 1e-9), TOL, 20)   # another shift: full factorisation
    assert info[7] == -1
    _close(E2, E0, 1e-12)
    ctx.close()


def test_tail_optimisation_sequence_no_drift(golden_dir):
    """An OptCycle-like run: the last function changes 12 times; every step
    is a one-row tail update on the same leading factor."""
    say("ecg_energy_grad tail sequence", "inout_1_500")
    f = _fixture(golden_dir, "inout_1_500")
    sigma = f.current_energy * f.invit_parameter
    K = f.basis_size
    ctx = EcgContext(f.system)
    p = f.params.copy()
    ctx.set_params(p)
    ctx.energy_grad(K - 1, 1, sigma, TOL, 20)
    for step in range(12):
        p = _perturb(p, slice(K - 1, K), 300 + step, 3e-3)
        ctx.set_params(p)
        _, E1, g1, info = ctx.energy_grad(K - 1, 1, sigma, TOL, 20)
        assert info[7] == 1, step
        if step % 4 == 3:
            E2, g2 = _fresh_grad(f.system, p, K - 1, 1, sigma)
            print("step", step, "E", E1, E2, "dg", np.max(np.abs(g1 - g2)), np.max(np.abs(g2)))
            _close(E1, E2)
            _grad_close(g1, g2)
    ctx.close()


def test_tail_update_params_entry(golden_dir):
    say("ecg_update_params tail", "inout_1_500")
    f = _fixture(golden_dir, "inout_1_500")
    sigma = f.current_energy * f.invit_parameter
    K = f.basis_size
    ctx = EcgContext(f.system)
    ctx.set_params(f.params)
    ctx.energy(1, sigma, TOL, 20)
    p = _perturb(f.params, slice(K - 5, K - 2), 401)
    ctx.update_params(K - 4, p[K - 5:K - 2])
    _, E1, _, info = ctx.energy(1, sigma, TOL, 20)
    assert info[7] == 5
    E2, _ = _fresh_energy(f.system, p, 1, sigma)
    _close(E1, E2)
    ctx.close()


@pytest.mark.parametrize("count", [10, 50])
def test_update_params_few_and_many(golden_dir, count):
    """ecg_update_params with few functions (their values travel as a kernel
    argument: 10 x 6 doubles at n = 2) and with more than one kernel
    argument holds (50 x 6 doubles: a copy), then ecg_set_params with an
    interior range changed: each energy equals a fresh context's."""
    say("ecg_update_params", "inout_1_500", count)
    f = _fixture(golden_dir, "inout_1_500")
    sigma = f.current_energy * f.invit_parameter
    K = f.basis_size
    ctx = EcgContext(f.system)
    ctx.set_params(f.params)
    ctx.energy(1, sigma, TOL, 20)
    first = K - count - 2   # 1-based
    p = _perturb(f.params, slice(first - 1, first - 1 + count), 500 + count)
    ctx.update_params(first, p[first - 1:first - 1 + count])
    _, E1, _, info = ctx.energy(1, sigma, TOL, 20)
    assert info[7] == K - first + 1   # rows refactored from the first changed one
    E2, _ = _fresh_energy(f.system, p, 1, sigma)
    _close(E1, E2)
    q = _perturb(p, slice(100, 100 + count), 600 + count)
    ctx.set_params(q)
    _, E3, _, _ = ctx.energy(1, sigma, TOL, 20)
    E4, _ = _fresh_energy(f.system, q, 1, sigma)
    _close(E3, E4)
    ctx.close()


def test_tail_basis_grows_and_shrinks(golden_dir):
    say("ecg_energy tail", "inout_1_1000", "grow/shrink")
    f = _fixture(golden_dir, "inout_1_1000")
    sigma = f.current_energy * f.invit_parameter
    ctx = EcgContext(f.system)
    ctx.set_params(f.params[:600])
    ctx.energy(1, sigma, TOL, 20)
    # capacity 640 for K = 600: 641 moves the matrices and factor to a
    # larger layout, which keeps their leading rows
    for K, rows in [(601, 1), (641, 40), (637, 0), (600, 0), (700, 100)]:
        p = f.params[:K]
        ctx.set_params(p)
        _, E1, _, info = ctx.energy(1, sigma, TOL, 20)
        assert info[7] == (rows if rows <= 64 else -1), (K, info[7])
        E2, _ = _fresh_energy(f.system, p, 1, sigma)
        print("K", K, "tail rows", rows, "E", E1, E2)
        _close(E1, E2)
    ctx.close()


def test_tail_second_root_inertia(golden_dir):
    say("ecg_energy tail", "inout_2_500", "root 2")
    f = _fixture(golden_dir, "inout_2_500")
    sigma = f.current_energy * f.invit_parameter
    K = f.basis_size
    ctx = EcgContext(f.system)
    ctx.set_params(f.params)
    _, E0, _, info0 = ctx.energy(2, sigma, TOL, 20)
    assert info0[4] == 1
    p = _perturb(f.params, slice(K - 9, K), 502)
    ctx.set_params(p)
    _, E1, _, info1 = ctx.energy(2, sigma, TOL, 20)
    assert info1[7] == 9 and info1[4] == 1
    E2, info2 = _fresh_energy(f.system, p, 2, sigma)
    _close(E1, E2)
    ctx.close()


def test_tail_stored_path_h2(golden_dir):
    """n = 3 (four symmetry terms), the stored-derivative path."""
    say("ecg_energy_grad_stored tail", "h2", 300)
    sy = named_system("h2")
    K = 300
    p = seeded_basis(sy.n, K, SEED_BASE + 3)
    ctx = EcgContext(sy)
    ctx.set_params(p)
    _, Er, _, _ = ctx.energy(1, float("nan"), TOL, 40)
    sigma = Er - 1e-3 * abs(Er)
    ctx.energy_grad_stored(K - 2, 1, sigma, TOL, 20)
    q = _perturb(p, slice(K - 2, K), 601)
    ctx.set_params(q)
    _, E1, g1, info = ctx.energy_grad_stored(K - 2, 1, sigma, TOL, 20)
    assert info[7] == 2
    E2, g2 = _fresh_grad(sy, q, K - 2, 1, sigma, stored=True)
    _close(E1, E2)
    _grad_close(g1, g2)
    ctx.close()


def test_tail_state_reset_by_other_paths(golden_dir):
    """Paths that overwrite H, S or the factor leave the next call a full
    factorisation."""
    say("tail state reset", "inout_1_500")
    f = _fixture(golden_dir, "inout_1_500")
    sigma = f.current_energy * f.invit_parameter
    K = f.basis_size
    ctx = EcgContext(f.system)
    ctx.set_params(f.params)
    ctx.energy(1, sigma, TOL, 20)
    H, S, dH, dS = ctx.ComputeMatElem(1, K, "G")   # same values as the energy path's: factor kept
    _, E0, _, info = ctx.energy(1, sigma, TOL, 20)
    assert info[7] == 0
    ctx.ComputeMatElem(2, K, "G", H=H, S=S, diagH=dH, diagS=dS)   # the caller's diagS: nothing kept
    _, E1, _, info = ctx.energy(1, sigma, TOL, 20)
    assert info[7] == -1
    _close(E1, f.current_energy, 1e-12)
    H, S, dH, dS = ctx.ComputeMatElem(1, K, "I", sigma)
    invD = np.zeros(K, np.complex128)
    ctx.GHEPIIS(1, H, invD, S, sigma, np.ones(K, np.complex128), TOL)
    _, E2, _, info = ctx.energy(1, sigma, TOL, 20)
    assert info[7] == -1
    _close(E2, f.current_energy, 1e-12)
    ctx.close()


def test_tail_no_reassembly_then_update(golden_dir):
    """A call that reassembles nothing records no assembly events; the
    calls after it still succeed."""
    say("no reassembly then ecg_update_params", "inout_1_500")
    f = _fixture(golden_dir, "inout_1_500")
    sigma = f.current_energy * f.invit_parameter
    K = f.basis_size
    ctx = EcgContext(f.system)
    ctx.set_params(f.params)
    ctx.energy(1, sigma, TOL, 20)
    _, E0, g0, info = ctx.energy_grad(K - 1, 1, sigma, TOL, 20)
    assert info[7] == 0
    p = _perturb(f.params, slice(K - 1, K), 701)
    ctx.update_params(K, p[K - 1])
    _, E1, g1, info = ctx.energy_grad(K - 1, 1, sigma, TOL, 20)
    assert info[7] == 1
    E2, g2 = _fresh_grad(f.system, p, K - 1, 1, sigma)
    _close(E1, E2)
    _grad_close(g1, g2)
    ctx.close()
