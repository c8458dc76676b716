"""GPU checks of the context's call-order state through the C-ABI.

The post-processing entries (ecg_expectation_values and the grid / MV
values with c = NULL, ecg_eigenvector) use the eigenvector of the last
energy call, and the screening entries use the session of the last
ecg_screen_prepare.  Both must refuse (ECG_ERR_STATE) once the parameters
they were built from have changed, instead of silently mixing the old
vector or factor with the new basis.  GHEPIIS on host arrays of another
size leaves the context without usable parameters until they are set
again.
"""
from __future__ import annotations

import numpy as np
import pytest

from ecg_nonbo import named_system, seeded_basis, SEED_BASE
from ecg_nonbo.hip import EcgContext, EcgError, ECG_ERR_STATE, ECG_OK

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _state_error(fn, *a, **kw):
    with pytest.raises(EcgError) as ei:
        fn(*a, **kw)
    assert ei.value.code == ECG_ERR_STATE


def test_vector_follows_parameters():
    sy = named_system("h2")
    K = 40
    p = seeded_basis(sy.n, K, SEED_BASE + 1700)
    say("vector state", "h2", K)
    ctx = EcgContext(sy)
    ctx.set_params(p)
    rc, E, c, _ = ctx.energy(1, float("nan"), 1e-12, 60, want_vector=True)
    assert rc == ECG_OK
    v0, _ = ctx.expectation_values()
    np.testing.assert_allclose(ctx.eigenvector(), c, rtol=0, atol=1e-15)
    # the same parameters again: nothing changed, the vector stays valid
    ctx.set_params(p.copy())
    v1, _ = ctx.expectation_values()
    np.testing.assert_array_equal(v0, v1)
    # changed parameters at the same K: the old vector must not be used
    q = p.copy()
    q[7] *= 1.01
    ctx.set_params(q)
    _state_error(ctx.expectation_values)
    _state_error(ctx.eigenvector)
    _state_error(ctx.correlation_functions, np.array([0.5, 1.0]))
    # an explicit vector is still accepted
    ctx.expectation_values(c)
    # one function replaced
    rc, E2, c2, _ = ctx.energy(1, float("nan"), 1e-12, 60, want_vector=True)
    assert rc == ECG_OK and abs(E2 - E) > 0
    ctx.expectation_values()
    ctx.update_params(K, q[K - 1:K] * 1.001)
    _state_error(ctx.expectation_values)
    # the energy call restores it, and <H> of the new vector is its energy
    rc, E3, _, _ = ctx.energy(1, float("nan"), 1e-12, 60)
    vals, _ = ctx.expectation_values()
    m = sy.n * (sy.n + 1) // 2
    assert abs(vals[5 * m] - E3) <= 1e-12 * abs(E3)
    ctx.close()


def test_screening_session_closes_on_parameter_change():
    sy = named_system("h2")
    K = 48
    p = seeded_basis(sy.n, K, SEED_BASE + 1701)
    cand = seeded_basis(sy.n, 4, SEED_BASE + 1702)
    say("screening state", "h2", K)
    ctx = EcgContext(sy)
    ctx.set_params(p)
    rc, E, _, _ = ctx.energy(1, float("nan"), 1e-12, 60)
    sigma = E - 1e-4 * abs(E)
    ctx.screen_prepare(sigma)
    Ea, _ = ctx.screen(cand)
    q = p.copy()
    q[3] *= 0.99
    ctx.set_params(q)
    _state_error(ctx.screen, cand)
    _state_error(ctx.screen_loop, cand[None, :, :])
    ctx.screen_prepare(sigma)
    Eb, _ = ctx.screen(cand)
    assert np.all(np.isfinite(Eb)) and np.max(np.abs(Eb - Ea)) > 0
    ctx.update_params(2, q[1:2] * 1.01)
    _state_error(ctx.screen, cand)
    ctx.close()


def test_ghepiis_of_another_size_needs_new_parameters():
    sy = named_system("hminus")
    K = 30
    p = seeded_basis(sy.n, K, SEED_BASE + 1703)
    say("ghepiis state", "hminus", K)
    ctx = EcgContext(sy)
    ctx.set_params(p)
    rc, E, _, _ = ctx.energy(1, float("nan"), 1e-12, 60)
    assert rc == ECG_OK
    # a GHEPIIS call on a 20 x 20 problem of the caller's
    from oracle import oracle
    H, S, dH, dS = oracle.assemble(sy, p[:20], "I", E)
    rc2, lam, x, _, _ = ctx.GHEPIIS(1, H.copy(order="F"), np.zeros(20, np.complex128), S.copy(order="F"),
                                    E, np.ones(20, np.complex128), 1e-12, 40)
    _state_error(ctx.energy, 1, float("nan"), 1e-12, 60)
    _state_error(ctx.expectation_values)
    ctx.set_params(p)
    rc, E2, _, _ = ctx.energy(1, float("nan"), 1e-12, 60)
    assert rc == ECG_OK and abs(E2 - E) <= 1e-12 * abs(E)
    ctx.close()
