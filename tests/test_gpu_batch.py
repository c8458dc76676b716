"""GPU parity of packed multi-problem batches (ecg_batch_energy_grad): every
problem's energy and gradient against the same problem alone
(ecg_energy_grad), for packed and per-problem pair launches and one or
several concurrent eigen streams."""
from __future__ import annotations

import numpy as np
import pytest

from ecg_nonbo import named_system, seeded_basis, SEED_BASE
from ecg_nonbo.hip import EcgBatch, EcgContext, ECG_OK

pytestmark = pytest.mark.gpu

MIX = [("hminus", 70), ("h2", 130), ("li", 45), ("h2", 64), ("hminus", 200), ("li", 96), ("h2", 17)]


@pytest.mark.parametrize("packed,nstreams", [(1, 1), (1, 3), (0, 2), (2, 1), (2, 3)])
def test_batch_matches_single_problems(packed, nstreams):
    print("[case] batch", packed, nstreams, MIX, flush=True)
    ctxs, sig, ref = [], [], []
    for i, (name, K) in enumerate(MIX):
        sy = named_system(name)
        p = seeded_basis(sy.n, K, SEED_BASE + 800 + i)
        c = EcgContext(sy)
        c.set_params(p)
        rc, E0, _, _ = c.energy(1, float("nan"), 1e-12, 60)
        s = E0 - 1e-6 * abs(E0)
        rc, E, g, _ = c.energy_grad(0, 1, s, 1e-12, 40)
        ctxs.append(c)
        sig.append(s)
        ref.append((E, g))
    b = EcgBatch(ctxs)
    b.set_config(packed, nstreams)
    for rep in range(2):   # a second call reuses the descriptor tables
        rc, E, grads, lat = b.energy_grad(sig, 1e-12, 40)
        for i, (Er, gr) in enumerate(ref):
            print(rep, MIX[i], E[i], Er, E[i] - Er, np.max(np.abs(grads[i] - gr)), lat[i])
            assert abs(E[i] - Er) <= 1e-11 * max(1.0, abs(Er))
            np.testing.assert_allclose(grads[i], gr, rtol=1e-8, atol=1e-9 * np.max(np.abs(gr)))
        assert np.all(lat > 0)
    b.close()
    for c in ctxs:
        c.close()
