"""GPU parity at the full sizes of two benchmark configurations.

C3: H2-like (n = 3, T = 4) energy + gradient at K = 4000 in both memory
modes (the fused recompute pass and the reference's stored-derivative
scheme, 6.1 GB of Glob_D on the device).
C5: the 32-problem batch (H-, H2, Li round robin, K ~ U[200, 2000]).

The oracle at these sizes: its assembly of H and S (the restated
ComputeMatElem / StoreHS, src/matform.f90:11-230) on the host cores, then
LAPACK on the oracle's own matrices for the eigen step (a Cholesky
factorisation of H - sigma S, which also proves that sigma lies below the
lowest root, and inverse iteration with it), and the restated
ComputeMatElemAndDeriv + EnergyGB contraction (src/workproc.f90:1925-1940)
for the gradient of the last functions (nfru = K - 8), whose rows are the
only ones the derivative assembly needs beyond H and S.  The device's
whole gradient is compared between its two memory modes.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg

from ecg_nonbo import named_system, seeded_basis, SEED_BASE
from ecg_nonbo.hip import EcgBatch, EcgContext, ECG_OK, ECG_NOT_CONVERGED
from oracle import oracle

pytestmark = pytest.mark.gpu

NTAIL = 8   # functions whose gradient the oracle recomputes at full size


def say(*a):
    print("[case]", *a, flush=True)


def oracle_energy_tail_grad(sy, params, sigma, ntail=NTAIL, iters=4):
    """(E, x, gradient of the last ntail functions) from the oracle's H, S.

    H - sigma S must be positive definite (Cholesky succeeds): then every
    root lies above sigma, and inverse iteration at sigma converges to the
    lowest one; E is the Rayleigh quotient of the converged vector."""
    K = params.shape[0]
    H, S, dH, dS = oracle.assemble(sy, params, "I", sigma)
    # method 'I' storage: lower H holds (H - sigma S) f, its diagonal
    # H_kk / S_kk - sigma; S is stored in both triangles with a unit diagonal
    M = np.tril(H, -1)
    M = M + M.conj().T + np.diag(np.diag(H).real.astype(np.complex128))
    Sf = np.array(S, copy=True)
    np.fill_diagonal(Sf, 1.0)
    cf = scipy.linalg.cho_factor(M, lower=True, check_finite=False)
    x = np.ones(K, np.complex128)
    for _ in range(iters):
        x = scipy.linalg.cho_solve(cf, Sf @ x, check_finite=False)
        x /= np.max(np.abs(x))
    xs = Sf @ x
    t = np.vdot(x, xs).real
    E = np.vdot(x, M @ x).real / t + sigma
    x = x / np.sqrt(t)
    nfru = K - ntail
    _, _, _, _, D = oracle.assemble_deriv(sy, params, nfru, "I", sigma, 0, H.copy(order="F"),
                                          S.copy(order="F"), dH.copy(), dS.copy())
    g = oracle.gradient(D, x, E, nfru)
    return E, x, g


def test_c3_k4000_both_memory_modes():
    sy = named_system("h2")
    K = 4000
    params = seeded_basis(sy.n, K, SEED_BASE + 3 + K)
    say("C3", "h2", K, "seed", SEED_BASE + 3 + K)
    ctx = EcgContext(sy)
    ctx.set_params(params)
    rc, E0, _, info0 = ctx.energy(1, float("nan"), 1e-12, 60)
    assert rc == ECG_OK and int(info0[4]) == 0
    sigma = E0 - 1e-6 * abs(E0)
    rc, E1, g1, _ = ctx.energy_grad(0, 1, sigma, 1e-12, 30)
    assert rc == ECG_OK
    say("C3 stored mode", K)
    rc, E2, g2, _ = ctx.energy_grad_stored(0, 1, sigma, 1e-12, 30)
    assert rc == ECG_OK
    ctx.close()
    gmax = np.max(np.abs(g1))
    print("recompute vs stored: dE", E1 - E2, "max dg", np.max(np.abs(g1 - g2)), "max|g|", gmax)
    assert abs(E1 - E2) <= 1e-12 * abs(E1)
    assert np.max(np.abs(g1 - g2)) <= 1e-10 * gmax
    say("C3 oracle", K)
    Eo, _, go = oracle_energy_tail_grad(sy, params, sigma)
    print("E", E1, "oracle", Eo, "dE", E1 - Eo, "tail grad max dev", np.max(np.abs(g1[-go.size:] - go)),
          "max|g tail|", np.max(np.abs(go)))
    assert abs(E1 - Eo) <= 1e-10 * abs(Eo)
    np.testing.assert_allclose(g1[-go.size:], go, rtol=1e-7, atol=1e-8 * np.max(np.abs(go)))
    np.testing.assert_allclose(g2[-go.size:], go, rtol=1e-7, atol=1e-8 * np.max(np.abs(go)))


BATCH_SYSTEMS = ("hminus", "h2", "li")


def test_c5_batch_32_problems():
    nprob = 32
    rng = np.random.default_rng(SEED_BASE + 5)
    Ks = rng.integers(200, 2001, nprob)
    say("C5", nprob, "problems, K", int(Ks.min()), "..", int(Ks.max()))
    ctxs, sig, ref, probs = [], [], [], []
    for i in range(nprob):
        sy = named_system(BATCH_SYSTEMS[i % 3])
        K = int(Ks[i])
        p = seeded_basis(sy.n, K, SEED_BASE + 5 + 100 * (i + 1))
        c = EcgContext(sy)
        c.set_params(p)
        rc, E0, _, info0 = c.energy(1, float("nan"), 1e-12, 60)
        # H- bases of this generator reach kappa(S) ~ 1e12, where GHEPIIS's
        # measure levels off above 1e-12: code 2, as the reference returns
        assert (rc == ECG_OK or (rc == ECG_NOT_CONVERGED and info0[1] < 1e-8)) and int(info0[4]) == 0, \
            (i, rc, info0[1])
        s = E0 - 1e-6 * abs(E0)
        rc, E, g, info = c.energy_grad(0, 1, s, 1e-12, 30)
        assert rc == ECG_OK or (rc == ECG_NOT_CONVERGED and info[1] < 1e-8), (i, rc, info[1])
        ctxs.append(c)
        sig.append(s)
        ref.append((E, g, info[1]))
        probs.append((sy, p))
    b = EcgBatch(ctxs)
    for packed, ns in ((2, 4), (1, 4), (0, 4)):
        say("C5 batch", "packed", packed, "streams", ns)
        b.set_config(packed, ns)
        rc, E, grads, lat = b.energy_grad(sig, 1e-12, 30)
        for i, (Er, gr, nd) in enumerate(ref):
            # the batch starts inverse iteration from another vector than the
            # single run; where the iteration stops at its rounding floor
            # (measure nd, up to ~4e-10 on the H- bases with kappa(S) ~ 1e12)
            # the vectors agree to about nd, and the gradient, a near-stationary
            # sum with heavy cancellation, to a few hundred times that
            fl = max(1e-9, 1000.0 * nd)
            assert abs(E[i] - Er) <= max(1e-11, fl * fl) * max(1.0, abs(Er)), (i, E[i], Er)
            np.testing.assert_allclose(grads[i], gr, rtol=fl, atol=fl * np.max(np.abs(gr)))
        assert np.all(lat > 0)
    b.close()
    for c in ctxs:
        c.close()
    # the oracle on the largest H2 problem and the largest Li problem
    for name in ("h2", "li"):
        idx = [i for i in range(nprob) if BATCH_SYSTEMS[i % 3] == name]
        i = max(idx, key=lambda q: Ks[q])
        sy, p = probs[i]
        say("C5 oracle", name, "problem", i, "K", int(Ks[i]))
        Eo, _, go = oracle_energy_tail_grad(sy, p, sig[i])
        Er, gr, _ = ref[i]
        print(name, "E", Er, "oracle", Eo, "dE", Er - Eo, "tail grad max dev", np.max(np.abs(gr[-go.size:] - go)))
        assert abs(Er - Eo) <= 1e-10 * abs(Eo)
        np.testing.assert_allclose(gr[-go.size:], go, rtol=1e-7, atol=1e-8 * np.max(np.abs(go)))
