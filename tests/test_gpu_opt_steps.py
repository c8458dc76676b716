"""OptCycleG steps on the device against the reference's own logs.

example_output/out_{1,2}_1500 log method-'G' OPT_CYCLE runs on the HD+
1500-function bases (roots 1 and 2; OPT_CYCLE G 1500 1 1500 1 1 .. 120
0.99 50: one function per step, MaxEnergyEval 120, overlap threshold 0.99,
coefficient threshold 50).  Per step the log holds the function's
parameters before and after, the numbers of energy and gradient
evaluations, and the energy after (tests/golden/replay/*.npz).  Undoing
the records from the end rebuilds the basis before any logged step; from
there one step of `opt_cycle` runs (the restated DRMNG with the drivers'
settings, the device energy and gradient with root isolation), and its
result is compared with the log's.

What agrees: the function optimised, the energy after the step (within
1e-13 Eh of the logged one; the per-step gains are 1e-15..1e-13, the
two codes' energies differ by a few 1e-15), and the evaluation counts in
aggregate (within 10 % for the energies, 35 % for the gradients).  Step
by step the counts agree only sometimes: the minimiser makes its
decisions on energy differences of the size of the two codes' rounding
differences, so its paths part after a few evaluations (the minimiser
itself is checked bit for bit against the reference's in test_drmng.py).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from ecg_nonbo import read_inout
from ecg_nonbo.hip import EcgContext
from ecg_nonbo.optimise import BasisOrder, opt_cycle

pytestmark = pytest.mark.gpu

# log: records taken from its end, eigenvalue tolerance of the device's
# inverse iteration (root 2: GHEPIIS's floor mode, accepted at 1e-10)
CASES = {"1": (24, 1e-12), "2": (16, -1e-10)}


class _Stop(Exception):
    pass


@pytest.mark.parametrize("tag", ["1", "2"])
def test_opt_cycle_steps_against_log(golden_dir, tag):
    nrec_run, tol = CASES[tag]
    f = read_inout(os.path.join(golden_dir, "hdplus", "inout_%s_1500.txt" % tag))
    d = np.load(os.path.join(golden_dir, "replay", "replay_%s_1500.npz" % tag))
    func, before, after, Elog, nev = d["func"], d["before"], d["after"], d["E"], d["nevals"]
    nrec = len(func)
    P = f.params.copy()
    for r in range(nrec):
        P[func[r] - 1] = after[r]
    which = f.which_eigenvalue
    ctx = EcgContext(f.system)
    tot = np.zeros(4, np.int64)
    exact = 0
    try:
        for r in range(nrec - 1, nrec - 1 - nrec_run, -1):
            Pr = P.copy()
            for q in range(nrec - 1, r - 1, -1):
                Pr[func[q] - 1] = before[q]
            fn = int(func[r])
            print("[case] log %s record %d function %d" % (tag, r, fn), flush=True)
            order = BasisOrder.identity(Pr)
            got = {}

            def on_step(rec, o):
                got["rec"] = rec
                raise _Stop()

            with pytest.raises(_Stop):
                opt_cycle(ctx, order, 1, 1500, 58, 120, float("nan"), overlap_threshold=0.99,
                          lin_coeff_threshold=50.0, history={"cycles_done": 57, "init_func_at_last_step": fn - 1},
                          tol=tol, which=which, maxiter=60, on_step=on_step)
            rec = got["rec"]
            assert rec.func == fn
            assert 1 <= rec.evaluations <= 120
            assert rec.energy_after <= rec.energy_before + 1e-14 * abs(rec.energy_before)
            assert abs(rec.energy_after - Elog[r]) <= 1e-13, (r, rec.energy_after, Elog[r])
            tot += [nev[r][0], nev[r][1], rec.evaluations, rec.grad_evaluations]
            exact += (rec.evaluations, rec.grad_evaluations) == tuple(nev[r])
    finally:
        ctx.close()
    print("log %s: %d steps, evaluations log %d / %d, device %d / %d (energy / gradient), %d steps with equal counts"
          % (tag, nrec_run, tot[0], tot[1], tot[2], tot[3], exact))
    assert abs(tot[2] / tot[0] - 1.0) <= 0.10
    assert abs(tot[3] / tot[1] - 1.0) <= 0.35
