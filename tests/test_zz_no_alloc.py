"""ecg_matrix_elements_noalloc against ecg_matrix_elements through the C-ABI.

The no-allocation variant stages the pair in the library's fixed device
scratch (one array per GPU, reused by every call) instead of hipMalloc'ed
buffers; its results must equal the allocating entry point's bit for bit.
Two calls in a row on different systems, the second smaller (n = 3 then
n = 2), check that nothing of the first call's staging leaks into the
second; then the first system again, to catch the reverse.
"""
from __future__ import annotations

import numpy as np
import pytest

from ecg_nonbo import named_system, seeded_basis, SEED_BASE
from ecg_nonbo.hip import EcgContext

pytestmark = pytest.mark.gpu


def _pairs(name, seed):
    sy = named_system(name)
    p = seeded_basis(sy.n, 2, seed)
    return sy, p


@pytest.mark.parametrize("gradflag", [3, 1, 0])
def test_noalloc_equals_allocating_entry(gradflag):
    cases = [_pairs("h2", SEED_BASE + 1800), _pairs("hminus", SEED_BASE + 1801), _pairs("li", SEED_BASE + 1802),
             _pairs("h2", SEED_BASE + 1803)]
    ctxs = [EcgContext(sy) for sy, _ in cases]
    jobs = [(ctx, p, P) for ctx, (sy, p) in zip(ctxs, cases) for P in sy.yhy_mats]
    print("[case] noalloc gradflag", gradflag, "calls", len(jobs), flush=True)
    ref = [ctx.MatrixElements(p[1], p[0], P, gradflag) for ctx, p, P in jobs]
    # the no-allocation calls back to back: n = 3 (T = 4), then n = 2, n = 3
    # (Li, T = 2) and n = 3 again, each reusing the same scratch
    got = [ctx.MatrixElements(p[1], p[0], P, gradflag, noalloc=True) for ctx, p, P in jobs]
    for a, b in zip(ref, got):
        assert a[0] == b[0] and a[1] == b[1]
        np.testing.assert_array_equal(a[2], b[2])
        np.testing.assert_array_equal(a[3], b[3])
        assert np.isfinite(a[0]) and abs(a[1]) > 0
    for ctx in ctxs:
        ctx.close()
