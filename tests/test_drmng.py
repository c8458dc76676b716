"""The optimisers' minimiser (csrc/drmng.cpp through ecg_nonbo.drmng)
against the reference's own PORT routines (src/dmng.f DIVSET / DRMNG,
compiled into oracle/_ref by `make -C oracle ref`), call by call: after
every reverse-communication return the whole integer and double workspaces,
the point and the scale vector must be identical bit for bit.  The cases
walk the minimiser's paths: the defaults and the settings of the
reference's drivers, the summary-printing settings (whose workspace side
effects are restated), evaluations that fail (with and without the
IV(2) = 1 flag), the evaluation and iteration limits, parameter errors and
n = 1.  Without the reference library those tests skip; the properties at
the end (convergence, the driver loop's accounting) run everywhere."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "atom-mol-nonbo_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from ecg_nonbo import drmng as dm  # noqa: E402
from oracle import ref  # noqa: E402

need_ref = pytest.mark.skipif(not ref.dmng_available(), reason="reference minimiser not built (no reference tree)")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not os.path.exists(dm.LIB_PATH):
        pytest.skip("libecg_drmng.so not built")
    dm.load()


def rosenbrock(x):
    a, b = x[0::2], x[1::2]
    f = float(np.sum(100.0 * (b - a * a) ** 2 + (1.0 - a) ** 2))
    g = np.zeros_like(x)
    g[0::2] = -400.0 * a * (b - a * a) - 2.0 * (1.0 - a)
    g[1::2] = 200.0 * (b - a * a)
    return f, g


def wavy1(x):
    return 3.0 * (x[0] - 2.0) ** 2 + math.sin(x[0]), np.array([6.0 * (x[0] - 2.0) + math.cos(x[0])])


def quadratic(x):
    # ill-conditioned convex quadratic, minimum at c
    s = np.array([1.0, 10.0, 100.0, 1e3, 1e4, 1e5])[: x.size]
    c = np.linspace(-1.0, 1.0, x.size)
    r = x - c
    return float(0.5 * np.sum(s * r * r)), s * r


def _settings(kind, n, ours):
    iv = np.zeros(dm.LIV, np.int32)
    v = np.zeros(dm.lv_for(n), np.float64)
    (dm.divset if ours else ref.ref_divset)(2, iv, v)
    if kind == "defaults":
        return iv, v
    if kind in ("driver", "fullopt", "mxfcal", "mxiter", "badv"):
        iv[16] = iv[17] = 1000000
        iv[18] = -1 if kind == "fullopt" else 0
        iv[19] = 0
        iv[21] = 0
        iv[22] = -1
        iv[23] = 0
        v[30] = 0.0
        v[34] = dm.MAX_SCALED_FIRST_STEP
        iv[0] = 12
    if kind == "mxfcal":
        iv[16] = 7
    if kind == "mxiter":
        iv[17] = 3
    if kind == "badv":
        v[34] = -1.0
    return iv, v


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.int64), np.ascontiguousarray(b).view(np.int64))


CASES = [
    # name, function, x0, d, settings, failure ball radius, flag failures
    ("rosen2-defaults", rosenbrock, [-1.2, 1.0], [1.0, 1.0], "defaults", None, False),
    ("rosen6-driver", rosenbrock, [-1.2, 1.0, -1.2, 1.0, 0.5, 0.3], [1e-2] * 6, "driver", None, False),
    ("rosen4-fullopt", rosenbrock, [-1.2, 1.0, 0.3, 0.2], [0.5] * 4, "fullopt", None, False),
    ("quad6-driver", quadratic, [0.3, -0.2, 0.9, 0.1, -0.4, 2.0], [1.0, 3.0, 10.0, 30.0, 100.0, 300.0], "driver",
     None, False),
    ("quad6-tiny-scale", quadratic, [0.3, -0.2, 0.9, 0.1, -0.4, 2.0], [dm.scale_for_basis(1500)] * 6, "driver",
     None, False),
    ("rosen4-toobig", rosenbrock, [-1.2, 1.0, 0.5, 0.5], [0.1] * 4, "driver", 2.0, True),
    ("rosen4-failure-kept", rosenbrock, [-1.2, 1.0, 0.5, 0.5], [0.1] * 4, "driver", 2.0, False),
    ("rosen2-mxfcal", rosenbrock, [-1.2, 1.0], [1.0, 1.0], "mxfcal", None, False),
    ("rosen2-mxiter", rosenbrock, [-1.2, 1.0], [1.0, 1.0], "mxiter", None, False),
    ("rosen2-zero-scale", rosenbrock, [-1.2, 1.0], [1.0, 0.0], "driver", None, False),
    ("rosen2-bad-lmax0", rosenbrock, [-1.2, 1.0], [1.0, 1.0], "badv", None, False),
    ("wavy1", wavy1, [10.0], [1.0], "driver", None, False),
]


def compare(name):
    """Drive both minimisers through case `name`; returns the IV(1) codes."""
    _, fg, x0, d0, kind, radius, toobig = next(c for c in CASES if c[0] == name)
    n = len(x0)
    ivA, vA = _settings(kind, n, True)
    ivB, vB = _settings(kind, n, False)
    assert np.array_equal(ivA, ivB) and _same(vA, vB), "DIVSET defaults differ"
    xA = np.array(x0, float)
    xB = xA.copy()
    dA = np.array(d0, float)
    dB = dA.copy()
    gA = np.zeros(n)
    gB = np.zeros(n)
    fA = fB = 0.0
    codes = []
    for call in range(700):
        cA = dm.drmng(dA, fA, gA, ivA, vA, xA)
        cB = ref.ref_drmng(dB, fB, gB, ivB, vB, xB)
        assert cA == cB, "call %d: IV(1) %d vs %d" % (call, cA, cB)
        assert np.array_equal(ivA, ivB), "call %d: IV differs at %s" % (call, np.nonzero(ivA != ivB)[0] + 1)
        assert _same(vA, vB), "call %d: V differs at %s" % (call, np.nonzero(vA != vB)[0] + 1)
        assert _same(xA, xB) and _same(dA, dB), "call %d: x or d differs" % call
        codes.append(cA)
        if cA not in (1, 2):
            break
        if radius is not None and np.linalg.norm(xA) > radius:
            ivA[1] = ivB[1] = 1 if toobig else 0
            continue
        f, g = fg(xA)
        if cA == 1:
            fA = fB = f
        else:
            gA[:] = g
            gB[:] = g
    return codes


@need_ref
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_call_by_call_against_reference(case):
    name, radius = case[0], case[5]
    if case[4] in ("defaults", "fullopt") or name.endswith(("zero-scale", "bad-lmax0")):
        # the reference writes its reports to standard output from its own
        # runtime: run the comparison in a child process that keeps them
        code = ("import sys; sys.path.insert(0, %r); import test_drmng as t; t.dm.load(); "
                "print('CODES', t.compare(%r))" % (os.path.dirname(os.path.abspath(__file__)), name))
        out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr[-2000:]
        line = [ln for ln in out.stdout.splitlines() if ln.startswith("CODES")][0]
        codes = [int(c) for c in line[6:].strip("[] ").split(",")]
    else:
        codes = compare(name)
    # each case reaches the ending it was built for
    end = codes[-1]
    if name.endswith("mxfcal"):
        assert end == 9
    elif name.endswith("mxiter"):
        assert end == 10
    elif name.endswith("zero-scale"):
        assert end == 18
    elif name.endswith("bad-lmax0"):
        assert end == 35   # the index of the out-of-range V entry
    elif radius is None:
        assert 3 <= end <= 8, codes[-5:]


@need_ref
def test_driver_loop_against_reference_loop():
    """minimise() against the same loop written around the reference's
    DRMNG: same evaluation counts, best point and energy, with failed
    evaluations in the middle."""
    x0 = np.array([-1.2, 1.0, 0.5, 0.5, 0.2, 0.1])
    d = np.full(6, 0.05)
    fails = []

    def energy(x):
        if np.linalg.norm(x) > 2.2:
            fails.append(1)
            return False, 0.0
        return True, rosenbrock(x)[0]

    def energy_grad(x):
        f, g = rosenbrock(x)
        return True, f, g

    e0 = rosenbrock(x0)[0]
    r = dm.minimise(energy, energy_grad, x0, d, 60, e0)

    iv, v = _settings("driver", 6, False)
    x = x0.copy()
    g = np.zeros(6)
    fx = e0
    n_e = n_g = 0
    e_best, x_best = e0, x.copy()
    while n_e < 60:
        c = ref.ref_drmng(d, fx, g, iv, v, x)
        if c == 1:
            n_e += 1
            ok, e = energy(x)
            if ok:
                fx = e
                if e < e_best:
                    e_best, x_best = e, x.copy()
            else:
                iv[1] = 0
        elif c == 2:
            n_g += 1
            _, e, gg = energy_grad(x)
            g[:] = gg
            if e < e_best:
                e_best, x_best = e, x.copy()
        else:
            break
    assert (r.energy_evals, r.grad_evals) == (n_e, n_g)
    assert r.e_best == e_best and _same(r.x_best, x_best)


def test_exports():
    lib = dm.load()
    for name in dm.EXPORTS:
        assert hasattr(lib, name)


def test_converges_on_rosenbrock():
    x0 = np.array([-1.2, 1.0, -1.2, 1.0])

    def e(x):
        return True, rosenbrock(x)[0]

    def eg(x):
        f, g = rosenbrock(x)
        return True, f, g

    r = dm.minimise(e, eg, x0, np.ones(4), 500, rosenbrock(x0)[0], lmax0=1.0)
    assert 3 <= r.code <= 8
    assert np.max(np.abs(r.x_best - 1.0)) < 1e-6
    assert r.e_best < 1e-12


def test_quadratic_minimum_and_energy_budget():
    x0 = np.zeros(6)

    def e(x):
        return True, quadratic(x)[0]

    def eg(x):
        f, g = quadratic(x)
        return True, f, g

    r = dm.minimise(e, eg, x0, np.ones(6), 400, quadratic(x0)[0], lmax0=1.0)
    assert np.max(np.abs(r.x_best - np.linspace(-1.0, 1.0, 6))) < 1e-6
    # a budget of k energies stops the loop after exactly k energy calls
    # when the minimiser has not converged before
    r3 = dm.minimise(e, eg, x0, np.ones(6), 3, quadratic(x0)[0], lmax0=1.0)
    assert r3.energy_evals == 3
    assert r3.e_best <= quadratic(x0)[0]


def test_failures_beyond_the_limit_raise():
    # every point but the start fails: the minimiser keeps seeing the
    # start's energy, shortens its step, and the sixth failure stops the loop
    x0 = np.ones(3)

    def e(x):
        return (True, 1.0) if np.array_equal(x, x0) else (False, 0.0)

    def eg(x):
        return True, 1.0, np.array([1.0, -2.0, 0.5])

    with pytest.raises(RuntimeError, match="failures"):
        dm.minimise(e, eg, x0, np.ones(3), 100, 1.0)


def test_scales():
    # max(1 / (K^2 sqrt K), 1e4 eps): the floor takes over near K = 40000
    assert dm.scale_for_basis(1500) == 1.0 / (1500.0 * 1500.0 * math.sqrt(1500.0))
    assert dm.scale_for_basis(100000) == 10000 * np.finfo(float).eps
    assert dm.scale_for_enlargement(-1.0, -1.0, 5, 1) == 10000 * np.finfo(float).eps
    assert dm.scale_for_enlargement(-1.0, -1.1, 0, 1) == 1.0
    assert dm.MAX_SCALED_FIRST_STEP == 0.0010000000474974513
