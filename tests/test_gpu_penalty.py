"""GPU parity of the overlap penalty through the C-ABI.

Reference behaviour: ComputeOverlapPenalty and
ComputeOverlapPenaltyAndAddGradient (src/workproc.f90:1491-1693), called
from EnergyGA/EnergyIA (the energy becomes E + P, :1759-1764, 1996-2001)
and EnergyGB/EnergyIB (E + P and the penalty's gradient added,
:1942-1945, 2144-2147).  The oracle restates both routines on its own
assembled S and stored derivatives D.
"""
from __future__ import annotations

import numpy as np
import pytest

from ecg_nonbo import named_system, seeded_basis, SEED_BASE
from ecg_nonbo.hip import EcgBatch, EcgContext
from helpers import penalty_basis, penalty_threshold
from oracle import oracle

pytestmark = pytest.mark.gpu

BMAX = 0.37


def say(*a):
    print("[case]", *a, flush=True)


def _oracle_penalty(sy, p, nfru, t2):
    H, S, dH, dS, D = oracle.assemble_deriv(sy, p, nfru, "G")
    return oracle.overlap_penalty_grad(S, dS, D, nfru, BMAX, t2)


def _threshold(sy, p):
    H, S, _, _ = oracle.assemble(sy, p, "G")
    t2, gap, nhit = penalty_threshold(S)
    assert nhit >= 2 and gap > 1e-4, (nhit, gap)
    return t2


@pytest.mark.parametrize("name,K,nfru", [("hminus", 48, 0), ("hminus", 48, 17), ("h2", 40, 9), ("li", 33, 0)])
def test_overlap_penalty_vs_oracle(name, K, nfru):
    say("ecg_overlap_penalty", name, K, nfru)
    sy = named_system(name)
    p = penalty_basis(seeded_basis, sy.n, K, SEED_BASE + 900, 6)
    t2 = _threshold(sy, p)
    Po, go = _oracle_penalty(sy, p, nfru, t2)
    ctx = EcgContext(sy)
    ctx.set_params(p)
    ctx.ComputeMatElem(1, K, "G")
    P1, _ = ctx.overlap_penalty(nfru, BMAX, t2)
    g0 = np.linspace(-1.0, 1.0, sy.npt * (K - nfru))
    P2, g2 = ctx.overlap_penalty(nfru, BMAX, t2, grad=g0)
    print("t2", t2, "P", P1, P2, "oracle", Po, "max|g|", np.max(np.abs(go)), "dev",
          np.max(np.abs(g2 - g0 - go)))
    assert Po > 0.0
    assert abs(P1 - Po) <= 1e-12 * max(1.0, Po)
    assert P1 == P2
    np.testing.assert_allclose(g2 - g0, go, rtol=0, atol=1e-10 * max(1.0, np.max(np.abs(go))))
    ctx.close()


@pytest.mark.parametrize("mode", ["recompute", "stored"])
def test_energy_grad_adds_penalty(mode):
    """E + P and grad + dP from the energy-and-gradient entry points, and
    E + P from the energy-only one (with its own nfru)."""
    sy = named_system("hminus")
    K, nfru = 64, 21
    say("energy_grad with penalty", mode, K, nfru)
    p = penalty_basis(seeded_basis, sy.n, K, SEED_BASE + 910, 5)
    t2 = _threshold(sy, p)
    ctx = EcgContext(sy)
    ctx.set_params(p)
    rc, E0, _, _ = ctx.energy(1, float("nan"), 1e-12, 60)
    sigma = E0 - 1e-6 * abs(E0)
    f = ctx.energy_grad if mode == "recompute" else ctx.energy_grad_stored
    rc, Eoff, goff, ioff = f(nfru, 1, sigma, 1e-12, 40)
    ctx.set_overlap_penalty(True, BMAX, t2, nfru=3)
    rc, Eon, gon, ion = f(nfru, 1, sigma, 1e-12, 40)
    Po, go = _oracle_penalty(sy, p, nfru, t2)
    print("E off", Eoff, "on", Eon, "diff", Eon - Eoff, "oracle P", Po, "info", ion[6])
    assert ioff[6] == 0.0
    assert abs(ion[6] - Po) <= 1e-12 * max(1.0, Po)
    assert abs((Eon - Eoff) - Po) <= 1e-12 * max(1.0, abs(Eon))
    scale = max(1.0, np.max(np.abs(goff)), np.max(np.abs(go)))
    np.testing.assert_allclose(gon - goff, go, rtol=0, atol=1e-10 * scale)
    # energy only: the penalty over the pairs of the nfru given to the setter
    rc, Ea, _, ia = ctx.energy(1, sigma, 1e-12, 40)
    ctx.set_overlap_penalty(False)
    rc, Eb, _, ib = ctx.energy(1, sigma, 1e-12, 40)
    P3, _ = _oracle_penalty(sy, p, 3, t2)
    print("energy only", Ea, Eb, Ea - Eb, P3)
    assert ib[6] == 0.0
    assert abs((Ea - Eb) - P3) <= 1e-12 * max(1.0, abs(Ea))
    ctx.close()


@pytest.mark.parametrize("packed", [1, 0])
def test_batch_adds_penalty(packed):
    say("batch with penalty", packed)
    probs = [("hminus", 64, True), ("h2", 40, False), ("li", 36, True)]
    ctxs, sig, ref = [], [], []
    for i, (name, K, on) in enumerate(probs):
        sy = named_system(name)
        p = penalty_basis(seeded_basis, sy.n, K, SEED_BASE + 920 + i, 4)
        t2 = _threshold(sy, p)
        c = EcgContext(sy)
        c.set_params(p)
        rc, E0, _, _ = c.energy(1, float("nan"), 1e-12, 60)
        s = E0 - 1e-6 * abs(E0)
        c.set_overlap_penalty(on, BMAX, t2, nfru=5)
        rc, E, g, info = c.energy_grad(0, 1, s, 1e-12, 40)
        assert (info[6] > 0.0) == on
        ctxs.append(c)
        sig.append(s)
        ref.append((E, g))
    b = EcgBatch(ctxs)
    b.set_config(packed, 2)
    rc, E, grads, lat = b.energy_grad(sig, 1e-12, 40)
    for i, (Er, gr) in enumerate(ref):
        print(probs[i], E[i], Er, E[i] - Er, np.max(np.abs(grads[i] - gr)))
        assert abs(E[i] - Er) <= 1e-11 * max(1.0, abs(Er))
        np.testing.assert_allclose(grads[i], gr, rtol=1e-8, atol=1e-9 * np.max(np.abs(gr)))
    b.close()
    for c in ctxs:
        c.close()


def test_penalty_needs_assembled_overlaps():
    sy = named_system("hminus")
    ctx = EcgContext(sy)
    ctx.set_params(seeded_basis(sy.n, 10, SEED_BASE + 930))
    with pytest.raises(Exception) as e:
        ctx.overlap_penalty(0, BMAX, 0.5)
    assert "-5" in str(e.value)
    ctx.close()
