"""CPU checks of the work figures bench.py divides by measured time."""
from __future__ import annotations

import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_factor_flops_match_ldlh_closed_form():
    """Sum over the trailing-update launches = flops of a blocked complex
    LDL^H minus the first diagonal block and panel, which run in their own
    launches: total complex FMAs of an unpivoted LDL^H of order K are
    K(K-1)(K+1)/6 + O(K^2) ~ K^3/6; the look-ahead block inverses add
    (K/64) * 64^3/6.  Compared in 8-flop complex FMAs, to 2 %."""
    b = _bench()
    for K in (640, 2000, 4000):
        f = b.factor_flops(K)
        assert len(f) == (K + 63) // 64 - 1
        total_cfma = sum(f) / 8.0
        nb = 64
        ldl = K ** 3 / 6.0
        inv = ((K + nb - 1) // nb - 1) * nb ** 3 / 6.0
        first = nb ** 3 / 6.0 + (K - nb) * nb * (nb + 1) / 2.0   # first block and panel (own launches)
        # the update counts rank-64 updates of lower triangles incl. diagonal blocks: O(K^2 nb) extra
        expect = ldl + inv - first
        assert abs(total_cfma - expect) / expect < 0.05, (K, total_cfma, expect)


def test_pmc_traffic_file_parses():
    b = _bench()
    v = b.pmc_traffic("k_ldl_update")
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if os.path.exists(path):
        assert v is not None and v > 1e6
    assert b.pmc_traffic("no_such_kernel") is None
