"""The cyclic optimiser (OptCycleI's loop, src/workproc.f90:4902-5423) on
the device energy path: every step's energy is at most the one before it,
the steps run on tail updates (a few changed rows, not a full
refactorisation), the basis ends in its input order with only the
optimised functions changed, and the final energy agrees with a fresh
context's on the final parameters.  The minimiser is the restated DRMNG
with the reference's settings (checked call by call against the
reference's own in test_drmng.py)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from ecg_nonbo import read_inout
from ecg_nonbo.hip import EcgContext
from ecg_nonbo.optimise import BasisOrder, opt_cycle

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def test_opt_cycle_hdplus_last_functions(golden_dir):
    say("opt_cycle", "inout_1_500 first 200 functions, functions 185..200")
    f = read_inout(os.path.join(golden_dir, "hdplus", "inout_1_500.txt"))
    K = 200
    p0 = f.params[:K].copy()
    ctx = EcgContext(f.system)
    ctx.set_params(p0)
    _, E0, _, _ = ctx.energy(1, float("nan"), 1e-12, 60)
    sigma = E0 - 1e-4 * abs(E0)
    order = BasisOrder.identity(p0)
    hist = {"cycles_done": 0, "init_func_at_last_step": 0}
    # GHEPIIS's floor mode: on this truncated HD+ basis the reference's own
    # iteration levels off above the file's tolerance of 1e-12
    recs = opt_cycle(ctx, order, 185, 200, 1, max_energy_eval=10, sigma=sigma, overlap_threshold=0.99,
                     lin_coeff_threshold=1e6, history=hist, tol=-1e-9)
    energies = [r.energy_after for r in recs]
    print("E0", E0, "after", energies[-1], "steps", len(recs), "accepted", sum(r.accepted for r in recs))
    assert [r.func for r in recs] == list(range(185, 201))
    assert all(r.energy_after <= r.energy_before + 1e-12 * abs(r.energy_before) for r in recs)
    assert all(b.energy_before <= a.energy_after + 1e-12 for a, b in zip(recs, recs[1:]))
    assert energies[-1] <= E0 + 1e-12
    tails = [t for r in recs for t in r.tail_rows]
    assert max(tails) <= 64 and sum(1 for t in tails if t < 0) == 0, tails
    np.testing.assert_array_equal(order.func_num, np.arange(1, K + 1))
    np.testing.assert_array_equal(order.params[:184], p0[:184])
    assert hist["cycles_done"] == 1 and hist["init_func_at_last_step"] == 0
    fresh = EcgContext(f.system)
    fresh.set_params(order.params)
    _, Ef, _, info = fresh.energy(1, sigma, 1e-12, 30)
    fresh.close()
    ctx.close()
    assert info[7] == -1
    assert abs(Ef - energies[-1]) <= 1e-11 * abs(Ef), (Ef, energies[-1])
