"""GPU replay of the reference's cyclic optimisation logs.

example_output/out_{1,2}_1500 log method-'G' OPT_CYCLE runs on the HD+
1500-function bases (roots 1 and 2): per optimised function its parameters
before and after, and the energy E with the "after" parameters (every
earlier record applied).  tests/golden/replay/replay_{1,2}_1500.npz hold
those records (tests/golden/replay/make_replay.py).  The state after the
last record is the reference's inout_1_1500.txt for log 1; for log 2 the
shipped file is from a slightly later point of the same run (53 functions
of 19..110 re-optimised once more), so the state is rebuilt from the log's
own "after" parameters and its energy checked against the log's last E
first.  The replay then undoes the records one by one, the function's
"before" parameters through ecg_update_params, and compares each energy
with the E the log printed for the previous record (and, after the first
record is undone, with the CURRENT_ENERGY the log's header printed): about
2300 reference-held energies.  The device path is the reference's method
'I' at the fixed shift E_min (1 + 1e-12): reassembly of the changed row
and refactorisation, from that row when fewer than 65 rows follow it
(the tail update of LDLHF from Nmin), else in full.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from ecg_nonbo import read_inout
from ecg_nonbo.hip import EcgContext

pytestmark = pytest.mark.gpu

TOL_E = 1e-10   # Eh, the reference's printed energies


@pytest.mark.parametrize("tag", ["1", "2"])
def test_opt_cycle_log_replay(golden_dir, tag):
    f = read_inout(os.path.join(golden_dir, "hdplus", "inout_%s_1500.txt" % tag))
    d = np.load(os.path.join(golden_dir, "replay", "replay_%s_1500.npz" % tag))
    func, before, after, Elog = d["func"], d["before"], d["after"], d["E"]
    E_init = float(d["E_initial"])
    nrec = len(func)
    P = f.params.copy()
    for r in range(nrec):
        P[func[r] - 1] = after[r]
    if tag == "1":
        assert np.array_equal(P, f.params)   # the shipped file is the state after the last record
    which = f.which_eigenvalue
    sigma = min(float(np.min(Elog)), E_init) * (1.0 + 1e-12)
    print("[case] replay log", tag, "records", nrec, "root", which, "shift", sigma, flush=True)
    ctx = EcgContext(f.system)
    ctx.set_params(P)
    rc, E, _, info = ctx.energy(which, sigma, 1e-12, 40)
    assert int(info[4]) == which - 1
    devs = [abs(E - Elog[-1])]
    tails = 0
    for r in range(nrec - 1, -1, -1):
        ctx.update_params(int(func[r]), before[r])
        rc, E, _, info = ctx.energy(which, sigma, 1e-12, 40)
        expect = Elog[r - 1] if r > 0 else E_init
        devs.append(abs(E - expect))
        tails += int(info[7] >= 1)
        assert abs(E - expect) <= TOL_E, (r, int(func[r]), E, expect)
        assert int(info[4]) == which - 1, r
    ctx.close()
    devs = np.array(devs)
    print("replay log %s: %d energies, max |dE| %.2e, mean %.2e, tail updates %d" %
          (tag, devs.size, devs.max(), devs.mean(), tails))
