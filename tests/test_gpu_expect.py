"""GPU parity of the expectation values (ecg_expectation_values).

Reference behaviour: MatrixElementsForExpcVals (src/matelem.f90:949-1983)
per pair and symmetry term, summed by the driver over the lower triangle
with the eigenvector (src/workproc.f90:8158-8300).  The device values are
compared with the oracle's restatement of both (itself pinned on the CPU
by tests/test_expect.py): each value to 1e-11 relative plus ten times the
driver's own rounding bound, the bounds themselves to 1e-6 relative.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from ecg_nonbo import make_system, named_system, read_inout, seeded_basis, SEED_BASE
from ecg_nonbo.hip import EcgContext
from oracle import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _compare(vals, errs, ov, oe):
    tol = 1e-11 * np.abs(ov) + 10.0 * oe + 1e-300
    bad = np.abs(vals - ov) > tol
    assert not np.any(bad), (np.nonzero(bad), vals[bad], ov[bad])
    np.testing.assert_allclose(errs, oe, rtol=1e-6, atol=1e-300)


@pytest.mark.parametrize("name,K", [("n1", 24), ("hminus", 40), ("h2", 33), ("li", 30)])
def test_expectation_values_vs_oracle(name, K):
    say("ecg_expectation_values", name, K)
    sy = make_system([1.0, 1836.15267261], [1.0, -1.0], "(P11)") if name == "n1" else named_system(name)
    p = seeded_basis(sy.n, K, SEED_BASE + 61)
    rng = np.random.Generator(np.random.PCG64(62))
    c = rng.normal(size=K) + 1j * rng.normal(size=K)
    _, _, _, dS = oracle.assemble(sy, p, "G")
    ov, oe = oracle.expectation_values(sy, p, c, dS)
    ctx = EcgContext(sy)
    ctx.set_params(p)
    vals, errs = ctx.expectation_values(c)
    ctx.close()
    print("max rel dev", np.max(np.abs(vals - ov) / (np.abs(ov) + 1e-300)))
    _compare(vals, errs, ov, oe)


def test_expectation_values_hdplus_last_eigenvector(golden_dir):
    say("ecg_expectation_values", "inout_1_500", "last eigenvector")
    f = read_inout(os.path.join(golden_dir, "hdplus", "inout_1_500.txt"))
    sigma = f.current_energy * f.invit_parameter
    ctx = EcgContext(f.system)
    ctx.set_params(f.params)
    _, E, c, _ = ctx.energy(1, sigma, f.eigval_tolerance, 20, want_vector=True)
    vals, errs = ctx.expectation_values()
    vals2, _ = ctx.expectation_values(c)
    ctx.close()
    np.testing.assert_array_equal(vals, vals2)
    _, _, _, dS = oracle.assemble(f.system, f.params, "G")
    ov, oe = oracle.expectation_values(f.system, f.params, c, dS)
    _compare(vals, errs, ov, oe)
    m = 3
    H, S, T, V = vals[5 * m:5 * m + 4]
    print("HD+ <R>", vals[2 * m], "<H>", H, "E", E, "virial", -V / T)
    assert abs(S - 1.0) <= 1e-10
    assert abs(H - E) <= 1e-10 * abs(E)
    assert abs(-V / T - 2.0) <= 1e-6


def test_expectation_values_needs_a_vector():
    sy = named_system("hminus")
    ctx = EcgContext(sy)
    ctx.set_params(seeded_basis(sy.n, 8, SEED_BASE + 63))
    with pytest.raises(Exception):
        ctx.expectation_values()
    ctx.close()


@pytest.mark.parametrize("name,K,ngrid", [("n1", 24, 7), ("hminus", 40, 9), ("h2", 33, 4), ("li", 30, 5)])
def test_correlation_functions_vs_oracle(name, K, ngrid):
    """ecg_correlation_functions against the oracle (itself pinned on the
    CPU by its r = 0 value and radial moments): 1e-12 relative to the
    largest value of each pair's function, over grids whose length is and
    is not a multiple of the device's four-point chunks."""
    say("ecg_correlation_functions", name, K, ngrid)
    sy = make_system([1.0, 1836.15267261], [1.0, -1.0], "(P11)") if name == "n1" else named_system(name)
    p = seeded_basis(sy.n, K, SEED_BASE + 61)
    rng = np.random.Generator(np.random.PCG64(63))
    c = rng.normal(size=K) + 1j * rng.normal(size=K)
    grid = np.linspace(0.0, 3.0, ngrid)
    _, _, _, dS = oracle.assemble(sy, p, "G")
    ref = oracle.correlation_functions(sy, p, c, dS, grid)
    ctx = EcgContext(sy)
    ctx.set_params(p)
    got = ctx.correlation_functions(grid, c)
    vals, _ = ctx.expectation_values(c)
    ctx.close()
    m = sy.n * (sy.n + 1) // 2
    scale = np.max(np.abs(ref), axis=0)
    print("max rel dev", np.max(np.abs(got - ref) / scale))
    assert np.all(np.abs(got - ref) <= 1e-12 * scale + 1e-300)
    np.testing.assert_allclose(got[0], vals[4 * m:5 * m], rtol=1e-12)   # CF(0) = <delta>


def test_correlation_functions_hdplus_last_eigenvector(golden_dir):
    say("ecg_correlation_functions", "inout_1_500", "last eigenvector")
    f = read_inout(os.path.join(golden_dir, "hdplus", "inout_1_500.txt"))
    sigma = f.current_energy * f.invit_parameter
    ctx = EcgContext(f.system)
    ctx.set_params(f.params)
    _, E, c, _ = ctx.energy(1, sigma, -1e-9, 20, want_vector=True)
    grid = np.linspace(0.0, 4.0, 6)
    got = ctx.correlation_functions(grid)
    got2 = ctx.correlation_functions(grid, c)
    ctx.close()
    np.testing.assert_array_equal(got, got2)
    _, _, _, dS = oracle.assemble(f.system, f.params, "G")
    ref = oracle.correlation_functions(f.system, f.params, c, dS, grid)
    scale = np.max(np.abs(ref), axis=0)
    assert np.all(np.abs(got - ref) <= 1e-11 * scale)


@pytest.mark.parametrize("name,K,ngrid", [("n1", 24, 5), ("hminus", 40, 6), ("h2", 33, 4), ("li", 30, 3)])
def test_particle_densities_vs_oracle(name, K, ngrid):
    """ecg_particle_densities against the oracle (pinned on the CPU by the
    n = 1 scaling identity and the radial moments)."""
    say("ecg_particle_densities", name, K, ngrid)
    sy = make_system([2.0, 1.0], [1.0, -1.0], "(P11)") if name == "n1" else named_system(name)
    p = seeded_basis(sy.n, K, SEED_BASE + 64)
    rng = np.random.Generator(np.random.PCG64(65))
    c = rng.normal(size=K) + 1j * rng.normal(size=K)
    grid = np.linspace(0.0, 2.0, ngrid)
    _, _, _, dS = oracle.assemble(sy, p, "G")
    ref = oracle.particle_densities(sy, p, c, dS, sy.masses, grid)
    ctx = EcgContext(sy)
    ctx.set_params(p)
    got = ctx.particle_densities(grid, c)
    with pytest.raises(Exception):
        ctx.particle_densities(grid, c, masses=np.r_[sy.masses[:-1], -1.0])
    ctx.close()
    scale = np.max(np.abs(ref), axis=0)
    print("max rel dev", np.max(np.abs(got - ref) / scale))
    assert np.all(np.abs(got - ref) <= 1e-12 * scale + 1e-300)


@pytest.mark.parametrize("name,K", [("n1", 24), ("hminus", 40), ("h2", 33), ("li", 30)])
def test_relativistic_corrections_vs_oracle(name, K):
    """ecg_relativistic_corrections (MV, Darwin) against the oracle, which
    restates the reference's loops (pinned on the CPU by the Laplacians of
    the Gaussians and the delta functions); the device uses the same
    traces in rank-one form, so values agree to rounding and the bounds to
    1e-6 relative."""
    say("ecg_relativistic_corrections", name, K)
    sy = make_system([2.0, 1.0], [1.0, -1.0], "(P11)") if name == "n1" else named_system(name)
    p = seeded_basis(sy.n, K, SEED_BASE + 66)
    rng = np.random.Generator(np.random.PCG64(67))
    c = rng.normal(size=K) + 1j * rng.normal(size=K)
    _, _, _, dS = oracle.assemble(sy, p, "G")
    ov, oe = oracle.relativistic(sy, p, c, dS, sy.masses)
    ctx = EcgContext(sy)
    ctx.set_params(p)
    vals, errs = ctx.relativistic_corrections(c)
    ctx.close()
    print("MV", vals[0], ov[0], "Darwin", vals[1], ov[1])
    _compare(vals, errs, ov, oe)


@pytest.mark.parametrize("name,K", [("n1", 12), ("hminus", 16), ("h2", 12), ("li", 10)])
def test_drachmanized_corrections_vs_oracle(name, K):
    """ecg_drachmanized_corrections (Drachmanized delta of every pair,
    orbit-orbit, Drachmanized mass-velocity and Darwin) against the oracle's
    restatement (pinned pair by pair on the CPU against the compiled
    reference, tests/test_ref_oracle.py).  These values are differences of
    larger terms, so the device's rounding is compared with the size of
    those sums: each value to 1e-10 of the driver's absolute-sum scale
    (its rounding bound divided by epsilon) plus 10 bounds; the bounds to
    1e-6 relative.  The Drachmanized mass-velocity and Darwin values, two
    formulas for one quantity, agree with each other to the same bar."""
    say("ecg_drachmanized_corrections", name, K)
    sy = make_system([2.0, 1.0], [1.0, -1.0], "(P11)") if name == "n1" else named_system(name)
    p = seeded_basis(sy.n, K, SEED_BASE + 68)
    rng = np.random.Generator(np.random.PCG64(69))
    c = rng.normal(size=K) + 1j * rng.normal(size=K)
    E = -1.1
    _, _, _, dS = oracle.assemble(sy, p, "G")
    ov, oe = oracle.drachmanized(sy, p, c, dS, sy.masses, E)
    ctx = EcgContext(sy)
    ctx.set_params(p)
    vals, errs = ctx.drachmanized_corrections(E, c)
    ctx.close()
    scale = oe / np.finfo(float).eps
    tol = 1e-10 * scale + 10.0 * oe + 1e-300
    dev = np.abs(vals - ov)
    print("max dev / scale", np.max(dev / scale), "values", vals)
    assert np.all(dev <= tol), (np.nonzero(dev > tol), vals, ov)
    np.testing.assert_allclose(errs, oe, rtol=1e-6, atol=1e-300)
    m = sy.n * (sy.n + 1) // 2
    assert abs(vals[m + 1] - vals[m + 2]) <= 1e-10 * (scale[m + 1] + scale[m + 2])


def test_drachmanized_corrections_hdplus_last_eigenvector(golden_dir):
    """On the HD+ fixture's eigenvector at its own energy: the last energy
    call's vector (c = NULL) gives what the explicit vector gives, and the
    oracle agrees."""
    say("ecg_drachmanized_corrections", "inout_1_500", "last eigenvector")
    f = read_inout(os.path.join(golden_dir, "hdplus", "inout_1_500.txt"))
    sigma = f.current_energy * f.invit_parameter
    ctx = EcgContext(f.system)
    ctx.set_params(f.params)
    _, E, c, _ = ctx.energy(1, sigma, f.eigval_tolerance, 20, want_vector=True)
    vals, errs = ctx.drachmanized_corrections(E)
    vals2, _ = ctx.drachmanized_corrections(E, c)
    ctx.close()
    np.testing.assert_array_equal(vals, vals2)
    _, _, _, dS = oracle.assemble(f.system, f.params, "G")
    ov, oe = oracle.drachmanized(f.system, f.params, c, dS, f.system.masses, E)
    scale = oe / np.finfo(float).eps
    print("drach delta, OO, drach MV, drach Darwin", vals, "max dev / scale", np.max(np.abs(vals - ov) / scale))
    assert np.all(np.abs(vals - ov) <= 1e-10 * scale + 10.0 * oe)
