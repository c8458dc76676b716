"""Benchmark of the ECG variational hot path on one MI355X.

Workload (one step): a full energy plus analytic gradient for the
4-particle H2-like non-BO system (two protons, two electrons; n = 3
pseudo-particles; Y = (1+P12)(1+P34), T = 4 simplified Y^dagger Y terms) with
K = 2000 seeded complex ECGs (numpy PCG64, seed 20261203 + 3; L_ii =
10**U(-2,2), L_ij ~ N(0, (0.1 L_jj)^2), B_ij ~ N(0, 0.05^2)), FP64 complex:
  assembly of all K(K+1)/2 = 2,001,000 pair elements H_kl, S_kl (x T terms)
  -> LDL^H of H - sigma S and inverse iteration for the lowest root
  -> fused adjoint pass: dE/d(vech L_k, vech B_k) for all 12 K parameters.
Each step starts from new parameters (every parameter scaled by 1 + 1e-7 u,
u ~ U(-1,1), as an optimiser step would) and uses the previous step's
energy for the shift (sigma = E_prev - 1e-6 |E_prev|, the reference's
method 'I' with its Glob_ApproxEnergy); nothing of a previous step's H, S,
factor or gradient is reused.

Metric: pair H,S+grad elements per second = K(K+1)/2 / (seconds per E+grad).
Prints one JSON line with a `roofline` object for the dominant kernel (the
one with the largest share of the step, timed with HIP events on the
library's stream) and a `cpu_baseline` object (the oracle's pair algebra on
the host cores over a bounded random sample of the same pairs, plus the
host eigen step at full size).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "atom-mol-nonbo_amd"))
sys.path.insert(0, ROOT)

from ecg_nonbo import named_system, seeded_basis, SEED_BASE  # noqa: E402
from ecg_nonbo.hip import EcgContext  # noqa: E402

FP64_PEAK_TFLOPS = 78.6       # MI355X dense FP64 (vector and matrix), AMD datasheet
HBM_PEAK_GBS = 8000.0


class DeviceBuffer:
    """A copy of a host array in device memory, through the HIP runtime the
    library already loaded (hipMalloc / hipMemcpy / hipDeviceSynchronize)."""

    def __init__(self, arr):
        import ctypes
        self._hip = ctypes.CDLL("libamdhip64.so.7")
        p = ctypes.c_void_p()
        if self._hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(arr.nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        if self._hip.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(arr.nbytes), 1) != 0:
            raise RuntimeError("hipMemcpy failed")   # 1 = hipMemcpyHostToDevice
        self.ptr = p.value

    def sync(self):
        self._hip.hipDeviceSynchronize()

    def free(self):
        import ctypes
        if self.ptr:
            self._hip.hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = 0


def algorithmic_flops(system, params, K):
    """Reference-algorithm flops (oracle operation count, convention of the
    flop table: add/sub/mul/div/sqrt = 1, per (pair, term), per-function set-up
    and integer permutation products excluded) of one full E+grad pair pass:
    T x [K(K-1)/2 x f(gradflag 3) + K x f(gradflag 1)], with f averaged over
    the T symmetry matrices."""
    from oracle import oracle
    f3 = np.mean([oracle.flops_per_pair(system, params[0], params[1], 3, term=t) for t in range(system.T)])
    f1 = np.mean([oracle.flops_per_pair(system, params[0], params[0], 1, term=t) for t in range(system.T)])
    f0 = np.mean([oracle.flops_per_pair(system, params[0], params[1], 0, term=t) for t in range(system.T)])
    return f0, f1, f3, system.T * (K * (K - 1) / 2 * f3 + K * f1)


def factor_flops(K, nb=64):
    """Algorithmic flops (complex FMA = 8) of each trailing-update launch of
    the blocked LDL^H (one per block column j but the last): the rank-64
    Hermitian update of the m x m trailing lower triangle (m = K - 64 (j+1)),
    the look-ahead LDL^H and unit-lower inverse of the next b x b diagonal
    block (b = min(64, m)), and the next panel W = A X^H over the m - b rows
    below it (X lower triangular)."""
    out = []
    for j0 in range(0, K - nb, nb):
        m = K - (j0 + nb)
        b = min(nb, m)
        upd = 8.0 * nb * m * (m + 1) / 2
        diag = 8.0 * b ** 3 / 6 + 8.0 * b ** 3 / 6
        panel = 8.0 * (m - b) * b * (b + 1) / 2
        out.append(upd + diag + panel)
    return out


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (profiles/pmc_traffic.txt, lines 'kernel bytes_per_launch'),
    or None when it has not been measured."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel and not line.startswith("#"):
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def pmc_flops(kernel):
    """Executed FP64 flops per launch of `kernel` from the committed
    rocprofv3 PMC summary (profiles/pmc_flops.txt: 64 x
    SQ_INSTS_VALU_FLOPS_FP64 + 512 x SQ_INSTS_VALU_MFMA_MOPS_F64 per
    dispatch), or None when it has not been measured."""
    path = os.path.join(ROOT, "profiles", "pmc_flops.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel and not line.startswith("#"):
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


ROCNAME = {"ldl_update": "k_ldl_update", "adjoint": "k_adjoint<3>", "trsv_fwd": "k_xsolve_fwd",
           "trsv_bwd": "k_xsolve_bwd", "offdiag": "k_offdiag<3>", "ldl_diag": "k_ldl_diag",
           "ldl_panel": "k_ldl_panel"}


def kernel_table(system, base, K, trace, nsteps, flops_adj):
    """Per-kernel device time per E+grad step (traced launches) and each
    kernel's roofline figures; the dominant kernel is the one with the
    largest share.  Compute-bound kernels carry two rates: `achieved` from
    the algorithmic work (the LDL^H closed form for the factorisation; the
    reference algorithm's operation count for the pair kernels, whose device
    formulation does fewer operations, so that rate is an `equivalent` one),
    and `executed` from the FP64 operations the kernel actually issues
    (rocprofv3 counters, profiles/pmc_flops.txt); `frac` is the executed
    rate over the peak (a measured fraction, at most 1), and for the
    factorisation also the algorithmic one.  The pair kernels run on the
    FP64 vector units (bound "valu"), the factorisation on MFMA."""
    upd = factor_flops(K)
    nf = max(1, trace["ldl_update"][0] // len(upd)) if trace["ldl_update"][0] else 1
    rows = {}
    for name, (cnt, tot) in trace.items():
        if cnt == 0:
            continue
        avg = tot / cnt * 1e-3   # s per launch
        r = {"launches_per_step": cnt / nsteps, "ms_per_step": tot / nsteps, "avg_us": avg * 1e6}
        ex = pmc_flops(ROCNAME.get(name, name))
        if ex is not None:
            r.update(executed_flops_per_launch=ex, executed_TFLOPs=ex / avg / 1e12)
        if name == "adjoint":
            r.update(bound="valu", work=flops_adj, unit="TFLOP/s", equivalent_TFLOPs=flops_adj / avg / 1e12,
                     peak=FP64_PEAK_TFLOPS)
        elif name == "ldl_update":
            w = float(np.mean(upd))
            r.update(bound="mfma", work=w, unit="TFLOP/s", achieved=w / avg / 1e12, peak=FP64_PEAK_TFLOPS)
        elif name in ("trsv_fwd", "trsv_bwd"):
            w = 16.0 * K * (K + 1) / 2     # the strict upper triangle (L^-H) streamed once
            r.update(bound="hbm", work=w, unit="GB/s", achieved=w / avg / 1e9, peak=HBM_PEAK_GBS)
        elif name == "offdiag":
            f0 = np.mean([__import__("oracle").oracle.flops_per_pair(system, base[0], base[1], 0, term=t)
                          for t in range(system.T)])
            w = system.T * f0 * K * (K - 1) / 2
            r.update(bound="valu", work=w, unit="TFLOP/s", equivalent_TFLOPs=w / avg / 1e12, peak=FP64_PEAK_TFLOPS)
        if "executed_TFLOPs" in r and r.get("unit") == "TFLOP/s":
            r["frac"] = r["executed_TFLOPs"] / r["peak"]
        if "achieved" in r:
            r["frac_algorithmic" if "frac" in r else "frac"] = r["achieved"] / r["peak"]
        rows[name] = r
    dominant = max(rows, key=lambda k: rows[k]["ms_per_step"])
    return rows, dominant, nf


def conditioning(ctx, K, E):
    """kappa(S) of the normalised overlap matrix of the current basis
    (eigenvalues of the device's S, copied out) and a first-order FP64
    bound of the energy: with c the eigenvector (c^H S c = 1) and
    element-wise relative rounding errors of at most eta in the assembled
    matrices (so ||dX||_2 <= ||dX||_F <= eta ||X||_F),
    |dE| = |c^H (dH - E dS) c| <= c^H c eta (||H||_F + |E| ||S||_F),
    eta = 64 eps (an allowance for the ~10^3 rounded operations of one pair
    element)."""
    c = ctx.eigenvector()
    H, S, dH, dS = ctx.ComputeMatElem(1, K, "G", 0.0)
    Sf = np.tril(S, -1)
    Sf = Sf + Sf.conj().T + np.eye(K, dtype=np.complex128)
    w = np.linalg.eigvalsh(Sf)
    Hl = np.tril(H, -1)
    hF = float(np.sqrt(2.0 * np.sum(np.abs(Hl) ** 2) + np.sum(dH ** 2)))
    sF = float(np.linalg.norm(Sf))
    cc = float(np.vdot(c, c).real)
    eta = 64.0 * np.finfo(float).eps
    return {"kappa_S": float(w[-1] / w[0]), "lambda_min_S": float(w[0]), "lambda_max_S": float(w[-1]),
            "cHc": cc, "eta": eta, "E_error_bound": cc * eta * (hF + abs(E) * sF),
            "note": "first-order bound |dE| <= c^H c eta (||H||_F + |E| ||S||_F) for element-wise relative "
                    "errors <= eta; the device's E agrees with the oracle's to 9e-16 on this basis "
                    "(tests/test_gpu_parity.py)"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(system, params, K, seconds_target=15.0):
    """Oracle pair algebra on all host threads over a random sample of the
    workload's pairs, then the host eigen step (LAPACK zhegvx via scipy) at
    the full K on the oracle's own H, S."""
    from oracle import oracle
    import scipy.linalg.lapack as lapack
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(7)
    n_probe = 4000
    kk = rng.integers(0, K, n_probe)
    ll = (rng.random(n_probe) * (kk + 1)).astype(np.int64)
    t0 = time.perf_counter()
    oracle.pair_batch(system, params, kk, ll, threads)
    rate = n_probe / (time.perf_counter() - t0)
    n_s = int(min(K * (K + 1) // 2, max(20000, rate * seconds_target * 0.7)))
    kk = rng.integers(0, K, n_s)
    ll = (rng.random(n_s) * (kk + 1)).astype(np.int64)
    t0 = time.perf_counter()
    oracle.pair_batch(system, params, kk, ll, threads)
    t_pairs = time.perf_counter() - t0
    rate = n_s / t_pairs
    H, S, dH, dS = oracle.assemble(system, params, "G", nthreads=threads)
    Hf, Sf = oracle.full_hermitian(H, S, dH)
    t0 = time.perf_counter()
    out = lapack.zhegvx(Hf, Sf, itype=1, jobz="V", uplo="U", il=1, iu=1, abstol=2 * np.finfo(float).tiny,
                        **{"range": "I"})
    t_eig = time.perf_counter() - t0
    npairs = K * (K + 1) // 2
    s_per_eg = npairs / rate + t_eig
    return {
        "value": npairs / s_per_eg,
        "unit": "pair elements/s",
        "cores": threads,
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": "%d random pairs of the K=%d H2 basis with gradients (T=4 terms each) in %.1f s, "
                  "plus one LAPACK zhegvx (lowest root, eigenvector) at K=%d in %.1f s; "
                  "s per E+grad = %.2f" % (n_s, K, t_pairs, K, t_eig, s_per_eg),
        "s_per_energy_grad": s_per_eg,
        "pair_rate": rate,
        "zhegvx_E": float(out[0][0]) if out[-1] == 0 else None,
    }


def screening_leg(system, base, K, ncand=256, nsteps=100, iters=4, plain_sample=32, tail_sample=64):
    """Batched candidate screening for the last function (K functions, the
    first K-1 fixed) against per-candidate scoring on the same GPU.

    Batched: nsteps optimisation steps of ncand seeded candidates each,
    one hipGraph replay per step (rows, Schur complements, `iters`
    inverse iterations, Rayleigh quotients, accept-if-better).
    Per candidate (the baseline of speedup_vs_per_candidate): the
    reference's own scheme on the device, i.e. replace function K
    (ecg_update_params), refactor from row K (the tail update of LDLHF from
    Nmin = K) and iterate at the same shift (ecg_energy), on a sample.
    Also reported: ecg_screen with one candidate per call (the batched
    kernels at C = 1) and a full factorisation per candidate (row K
    reassembled, LDL^H of the whole matrix at a new shift)."""
    ctx = EcgContext(system)
    ctx.set_params(base)
    st, E0, _, _ = ctx.energy(1, float("nan"), 1e-12, 60)
    sigma = E0 - 1e-6 * abs(E0)
    t0 = time.perf_counter()
    ctx.screen_prepare(sigma, 1e-12, 40)
    t_prep = time.perf_counter() - t0
    cand = np.stack([seeded_basis(system.n, ncand, SEED_BASE + 4 + 1000 * (s + 1)) for s in range(nsteps)])
    rng = np.random.default_rng(SEED_BASE + 4)
    for s in range(nsteps):   # a few near-copies of the current function, so that steps can improve
        cand[s, :4] = base[-1] * (1 + 1e-4 * rng.standard_normal((4, system.npt)))
    ctx.screen_loop(cand[:2], iters=iters)        # warm-up (graph build, first launches)
    ctx.set_params(base)
    ctx.screen_prepare(sigma, 1e-12, 40)
    t0 = time.perf_counter()
    Eb, ib = ctx.screen_loop(cand, iters=iters)
    t_loop = time.perf_counter() - t0
    tl = ctx.timings()
    dev_loop = tl[0] * 1e-3
    gemm_tile = int(tl[1])
    cand_per_block = int(tl[2])
    # accuracy of `iters` iterations: one step's candidates at iters and at 4 x iters
    ctx.set_params(base)
    ctx.screen_prepare(sigma, 1e-12, 40)
    Ea, ia = ctx.screen(cand[0], iters=iters)
    Ec, ic = ctx.screen(cand[0], iters=4 * iters)
    err = float(np.max(np.abs(Ea - Ec)))
    # batched kernels at one candidate per call
    t0 = time.perf_counter()
    for j in range(plain_sample):
        ctx.screen(cand[0, j:j + 1], iters=iters)
    t_plain = (time.perf_counter() - t0) / plain_sample
    # the reference's per-candidate scheme on the device: function K
    # replaced, LDL^H completed from row K, inverse iteration at sigma
    ctx.set_params(base)
    ctx.energy(1, sigma, 1e-12, 40)
    Et = np.zeros(tail_sample)
    t0 = time.perf_counter()
    for j in range(tail_sample):
        ctx.update_params(K, cand[0, j:j + 1])
        _, Et[j], _, _ = ctx.energy(1, sigma, 1e-12, 40)
    t_tail = (time.perf_counter() - t0) / tail_sample
    dE_tail = float(np.max(np.abs(Et - Ea[:tail_sample])))
    # a full factorisation per candidate (a new shift each time, so no tail update)
    full_sample = 4
    t0 = time.perf_counter()
    for j in range(full_sample):
        p = base.copy()
        p[-1] = cand[0, j]
        ctx.set_params(p)
        ctx.energy(1, sigma * (1.0 + 1e-15 * (j + 1)), 1e-12, 40)   # a new shift: no tail update
    t_full = (time.perf_counter() - t0) / full_sample
    ctx.close()
    rate = ncand * nsteps / t_loop
    return {
        "workload": "H2-like n=3 (T=4), K=%d, %d seeded candidates for function K per step, %d steps in one "
                    "hipGraph loop, %d inverse iterations per candidate" % (K, ncand, nsteps, iters),
        "candidates_per_s": rate,
        "ms_per_step": 1e3 * t_loop / nsteps,
        "device_ms_per_step": 1e3 * dev_loop / nsteps,
        "gemm_row_tile": gemm_tile,
        "candidates_per_block": cand_per_block,
        "prepare_ms": 1e3 * t_prep,
        "per_candidate_candidates_per_s": 1.0 / t_tail,
        "per_candidate_ms": 1e3 * t_tail,
        "speedup_vs_per_candidate": rate * t_tail,
        "max_dE_vs_per_candidate": dE_tail,
        "one_candidate_per_call_per_s": 1.0 / t_plain,
        "full_factorisation_candidates_per_s": 1.0 / t_full,
        "accepted_steps": int(np.sum(ib >= 0)),
        "E_start": float(E0), "E_end": float(Eb[ib >= 0][-1]) if np.any(ib >= 0) else float(E0),
        "max_dE_vs_%d_iterations" % (4 * iters): err,
    }


def optimise_leg(system, base, K, nevals=60, warm=3):
    """One function optimised at a time, as OptCycleI does with
    NumOfFuncToOpt = 1 (src/workproc.f90:5202-5240): each trial point
    changes only function K (ecg_update_params), and E with dE/d(params of
    function K) is evaluated at the same shift, so rows are reassembled and
    refactored from row K only.  Reported beside the same call after every
    function has changed (full assembly and factorisation)."""
    ctx = EcgContext(system)
    ctx.set_params(base)
    st, E0, _, _ = ctx.energy(1, float("nan"), 1e-12, 60)
    sigma = E0 - 1e-6 * abs(E0)
    ctx.energy_grad(K - 1, 1, sigma, 1e-12, 30)
    rng = np.random.default_rng(SEED_BASE + 77)
    trials = base[-1] * (1.0 + 1e-4 * rng.standard_normal((nevals + warm, system.npt)))
    for j in range(warm):
        ctx.update_params(K, trials[j])
        ctx.energy_grad(K - 1, 1, sigma, 1e-12, 30)
    rows, its = [], []
    t0 = time.perf_counter()
    for j in range(warm, warm + nevals):
        ctx.update_params(K, trials[j])
        st, e, g, inf = ctx.energy_grad(K - 1, 1, sigma, 1e-12, 30)
        rows.append(inf[7])
        its.append(inf[0])
    t_eval = (time.perf_counter() - t0) / nevals
    t0 = time.perf_counter()
    for j in range(warm, warm + nevals):
        ctx.update_params(K, trials[j])
        ctx.energy(1, sigma, 1e-12, 30)
    t_energy = (time.perf_counter() - t0) / nevals
    ctx.trace(True)
    for j in range(warm):
        ctx.update_params(K, trials[j])
        ctx.energy_grad(K - 1, 1, sigma, 1e-12, 30)
    trace = ctx.trace_read()
    ctx.trace(False)
    # the same call after every function changed
    nfull = 5
    t_full = 0.0
    for j in range(nfull):
        p = base * (1.0 + 1e-9 * rng.uniform(-1, 1, base.shape))
        p[-1] = trials[j]
        ctx.set_params(p)
        t0 = time.perf_counter()
        st, e_full, _, inf_full = ctx.energy_grad(K - 1, 1, sigma, 1e-12, 30)
        t_full += time.perf_counter() - t0
        assert inf_full[7] == -1
    t_full /= nfull
    # check: the last trial point's energy from a fresh context
    p = base.copy()
    p[-1] = trials[warm + nevals - 1]
    fresh = EcgContext(system)
    fresh.set_params(p)
    _, e_ref, _, _ = fresh.energy_grad(K - 1, 1, sigma, 1e-12, 30)
    fresh.close()
    ctx.close()
    return {
        "workload": "H2-like n=3 (T=4), K=%d: %d trial points of function K (parameters x (1 + 1e-4 N(0,1))), "
                    "E and the gradient of function K at a fixed shift" % (K, nevals),
        "ms_per_eval": 1e3 * t_eval,
        "evals_per_s": 1.0 / t_eval,
        "ms_per_energy_only": 1e3 * t_energy,
        "ms_full_path_same_call": 1e3 * t_full,
        "speedup_vs_full_path": t_full / t_eval,
        "tail_rows": sorted(set(int(r) for r in rows)),
        "mean_iterations": float(np.mean(its)),
        "dE_vs_fresh_context": float(e - e_ref),
        "trace_ms_per_eval": {k: v[1] / warm for k, v in trace.items() if v[0]},
    }


def opt_cycle_leg(system, base, K, nfunc=16, max_evals=12):
    """One cycle of the cyclic optimiser (OptCycleI's loop) over the last
    nfunc functions, at most max_evals energy evaluations each, at a fixed
    shift; every evaluation runs on the incremental path."""
    from ecg_nonbo.optimise import BasisOrder, opt_cycle
    ctx = EcgContext(system)
    ctx.set_params(base)
    st, E0, _, _ = ctx.energy(1, float("nan"), 1e-12, 60)
    sigma = E0 - 1e-6 * abs(E0)
    order = BasisOrder.identity(base)
    t0 = time.perf_counter()
    # GHEPIIS's floor mode (tol < 0): iterate to the rounding floor, accept a measure <= 1e-9
    recs = opt_cycle(ctx, order, K - nfunc + 1, K, 1, max_evals, sigma, tol=-1e-9)
    t = time.perf_counter() - t0
    ctx.close()
    tails = [x for r in recs for x in r.tail_rows]
    nev = len(tails)                       # device energy calls of the steps (one tail entry each)
    return {
        "workload": "H2-like n=3 (T=4), K=%d: one optimisation cycle over functions %d..%d, <= %d energy "
                    "evaluations each (the restated DRMNG with the reference's settings)" % (K, K - nfunc + 1, K, max_evals),
        "s_per_function": t / len(recs),
        "ms_per_energy_call": 1e3 * t / nev,
        "energy_calls": nev,
        "energy_gradient_evaluations": [int(sum(r.evaluations for r in recs)), int(sum(r.grad_evaluations for r in recs))],
        "E_start": float(E0), "E_end": float(recs[-1].energy_after),
        "accepted": int(sum(r.accepted for r in recs)),
        "max_tail_rows": int(max(tails)), "full_factorisations": int(sum(1 for x in tails if x < 0)),
    }


def enlarge_leg(system, base, K, n_trials=64, max_evals=6):
    """One step of the basis-enlargement loop (BasisEnlI's) for function K:
    n_trials candidates from the reference's generator around functions
    1..K-1, scored one energy call each (assembly of row K, tail update,
    inverse iterations) or all in one batched screening call, the best one
    kept and optimised with at most max_evals energy evaluations."""
    from ecg_nonbo.enlarge import basis_enl
    from ecg_nonbo.optimise import BasisOrder
    out = {"workload": "H2-like n=3 (T=4): function %d added to %d, %d trials (generator p1=0.7, s1=0.3, "
                       "s2=3.0), <= %d energy evaluations of optimisation, fixed shift" % (K, K - 1, n_trials, max_evals)}
    for mode in ("exact", "screen"):
        ctx = EcgContext(system)
        try:
            ctx.set_params(base[:K - 1])
            _, E0, _, _ = ctx.energy(1, float("nan"), 1e-12, 60)
            sigma = E0 - 1e-6 * abs(E0)
            order = BasisOrder.identity(base[:K - 1])
            rng = np.random.Generator(np.random.PCG64(SEED_BASE + 9))
            t0 = time.perf_counter()
            rec = basis_enl(ctx, order, K, 1, n_trials, max_evals, sigma, rng, generator_param=(0.7, 0.3, 3.0),
                            tol=-1e-9, trials=mode)[0]
            t = time.perf_counter() - t0
            out[mode] = {"s_per_step": t, "ms_per_trial_incl_optimisation": 1e3 * t / n_trials,
                         "E_before": rec.energy_before, "E_best_trial": rec.energy_trials,
                         "E_after": rec.energy_after, "prototype": rec.prototype, "trial_failures": rec.trial_failures}
        except RuntimeError as e:   # the loop's own stop conditions (as the reference's)
            out[mode] = {"error": str(e)}
        finally:
            ctx.close()
    return out


def expectation_leg(system, base, K, reps=3):
    """Expectation values (<1/r^2>, <1/r>, <r>, <r^2>, <delta(r)> of every
    pseudo-particle pair, <H>, <S>, <T>, <V>) in the lowest eigenvector:
    one fused pass over the K(K+1)/2 pairs x T terms after an energy call."""
    ctx = EcgContext(system)
    ctx.set_params(base)
    st, E, _, _ = ctx.energy(1, float("nan"), 1e-12, 60)
    vals, errs = ctx.expectation_values()
    t0 = time.perf_counter()
    for _ in range(reps):
        vals, errs = ctx.expectation_values()
    t = (time.perf_counter() - t0) / reps
    ctx.close()
    m = system.n * (system.n + 1) // 2
    npairs = K * (K + 1) // 2
    return {
        "workload": "H2-like n=3 (T=4), K=%d: %d expectation values in the lowest eigenvector" % (K, len(vals)),
        "ms": 1e3 * t,
        "pair_terms_per_s": npairs * system.T / t,
        "E": float(E), "H": float(vals[5 * m]), "S": float(vals[5 * m + 1]),
        "virial": float(-vals[5 * m + 3] / vals[5 * m + 2]),
        "max_rel_error_estimate": float(np.max(errs / np.maximum(np.abs(vals), 1e-300))),
    }


def memory_modes_leg(system, Ks=(1000, 4000), reps=3):
    """E+grad at K=1000 and K=4000 in the two memory modes: recompute (the
    fused adjoint pass; O(K) extra device memory) and stored (per-pair
    derivatives in Glob_D's layout, 2 npt x K x K complex, then the
    contraction pass).  Every repetition starts from new parameters (all
    functions moved by 1e-9 relative, uploaded before its timer starts), so
    nothing of the previous call's matrices or factor is reused."""
    out = []
    rng = np.random.default_rng(SEED_BASE + 44)
    for K in Ks:
        base = seeded_basis(system.n, K, SEED_BASE + 3 + K)
        sets = [base * (1.0 + 1e-9 * rng.uniform(-1, 1, base.shape)) for _ in range(reps + 1)]
        ctx = EcgContext(system)
        ctx.set_params(base)
        st, E0, _, _ = ctx.energy(1, float("nan"), 1e-12, 60)
        sigma = E0 - 1e-6 * abs(E0)
        row = {"K": K, "pairs": K * (K + 1) // 2, "D_bytes": 2 * system.npt * K * K * 16}
        for mode in ("recompute", "stored"):
            fn = ctx.energy_grad if mode == "recompute" else ctx.energy_grad_stored
            ctx.set_params(sets[reps])
            fn(0, 1, sigma, 1e-12, 30)   # warm-up (allocation)
            ts, tm = [], []
            for r in range(reps):
                ctx.set_params(sets[r])
                t0 = time.perf_counter()
                st, E, g, inf = fn(0, 1, sigma, 1e-12, 30)
                ts.append(time.perf_counter() - t0)
                tm.append(ctx.timings())
            tm = np.mean(tm, axis=0)
            row[mode] = {"ms": 1e3 * float(np.mean(ts)), "pair_elements_per_s": row["pairs"] / float(np.mean(ts)),
                         "assembly_ms": float(tm[0]), "eigen_ms": float(tm[1]), "gradient_pass_ms": float(tm[2]),
                         "E": float(E)}
            if mode == "recompute":
                g_ref = g
            else:
                row["max_grad_diff"] = float(np.max(np.abs(g - g_ref)))
        ctx.close()
        out.append(row)
    return out


BATCH_SYSTEMS = ("hminus", "h2", "li")


def batch_leg(nprob=32, reps=3, config=None):
    """32 independent problems (n=2 H^- and n=3 H2 / Li, round robin), K
    drawn uniformly from [200, 2000] (seeded), energy + full gradient each,
    as one packed batch.  A small autotuner times the launch structures
    (mode 2: packed pair launches and batched factorisations; 1: packed
    pair launches; 0: per problem, x 1/2/4 concurrent per-problem streams)
    on this batch and keeps the fastest; the timed runs use it.  config =
    (mode, streams) fixes the structure instead (no autotuning)."""
    from ecg_nonbo.hip import EcgBatch
    rng = np.random.default_rng(SEED_BASE + 5)
    Ks = rng.integers(200, 2001, nprob)
    ctxs, sig, pairs = [], [], 0
    for i in range(nprob):
        sy = named_system(BATCH_SYSTEMS[i % 3])
        K = int(Ks[i])
        c = EcgContext(sy)
        c.set_params(seeded_basis(sy.n, K, SEED_BASE + 5 + 100 * (i + 1)))
        st, E0, _, _ = c.energy(1, float("nan"), 1e-12, 60)
        sig.append(E0 - 1e-6 * abs(E0))
        ctxs.append(c)
        pairs += K * (K + 1) // 2
    b = EcgBatch(ctxs)
    tune = []
    for packed in (() if config else (2, 1, 0)):
        for ns in (1, 2, 4):
            b.set_config(packed, ns)
            ts = []
            for r in range(2):
                t0 = time.perf_counter()
                b.energy_grad(sig, 1e-12, 30)
                ts.append(time.perf_counter() - t0)
            tune.append({"packed": packed, "streams": ns, "ms": 1e3 * min(ts)})
    best = min(tune, key=lambda d: d["ms"]) if tune else {"packed": config[0], "streams": config[1]}
    b.set_config(best["packed"], best["streams"])
    ts, lats = [], []
    for r in range(reps):
        t0 = time.perf_counter()
        st, E, grads, lat = b.energy_grad(sig, 1e-12, 30)
        ts.append(time.perf_counter() - t0)
        lats.append(lat)
    b.close()
    for c in ctxs:
        c.close()
    lat = np.concatenate(lats)
    t = float(np.mean(ts))
    return {
        "workload": "%d problems (H- n=2, H2 and Li n=3, round robin), K ~ U[200, 2000] seed %d, energy + full "
                    "gradient each" % (nprob, SEED_BASE + 5),
        "pairs": int(pairs),
        "batch_ms": 1e3 * t,
        "pair_elements_per_s": pairs / t,
        "latency_ms_p50": float(np.percentile(lat, 50)),
        "latency_ms_p99": float(np.percentile(lat, 99)),
        "config": best,
        "autotune": tune,
    }


def copy_calibration(nbytes=256 << 20, reps=3):
    """Device-to-device copies of a known size through the HIP runtime (its
    copy kernel reads and writes nbytes each): the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print("# copy calibration: %d x %d bytes" % (reps, nbytes), file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--K", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-screen", action="store_true", help="skip the candidate-screening leg")
    ap.add_argument("--no-modes", action="store_true", help="skip the stored/recompute memory-mode leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the multi-problem batch leg")
    ap.add_argument("--no-opt", action="store_true", help="skip the one-function optimisation leg")
    ap.add_argument("--no-expect", action="store_true", help="skip the expectation-value leg")
    ap.add_argument("--no-cycle", action="store_true", help="skip the optimisation-cycle leg")
    ap.add_argument("--no-enlarge", action="store_true", help="skip the basis-enlargement leg")
    ap.add_argument("--batch-config", default=None, metavar="MODE,STREAMS",
                    help="run the batch leg with this launch structure instead of autotuning it")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first issue device copies of a known size (reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the path on a single MI355X "
              "(multi-GPU sharding is a design note only); --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
    K = args.K
    system = named_system("h2")
    base = seeded_basis(system.n, K, SEED_BASE + 3)
    rng = np.random.default_rng(SEED_BASE + 33)
    ctx = EcgContext(system)
    ctx.set_params(base)
    # cold start (untimed): locate the lowest root without a shift
    st, E, _, info = ctx.energy(1, float("nan"), 1e-12, 60)
    sigma = E - 1e-6 * abs(E)
    print("# cold start E=%.15f factorisations=%d iterations=%d" % (E, info[3], info[0]), file=sys.stderr,
          flush=True)

    # each step's parameters: an optimiser's small move of the basis.  They
    # are staged in HBM before the timed region; a step reads its set from
    # there (ecg_set_params_device).
    ntr = 3
    nsets = args.warmup + args.steps + ntr
    psets = np.ascontiguousarray(np.stack([base * (1.0 + 1e-7 * rng.uniform(-1, 1, base.shape))
                                           for _ in range(nsets)]))
    dsets = DeviceBuffer(psets)
    set_bytes = base.size * 8
    order = {"warmup": 0, "timed": args.warmup, "traced": args.warmup + args.steps}

    def step(i, phase="timed", host=False):
        nonlocal sigma
        j = order[phase] + i
        if host:
            ctx.set_params(psets[j])
        else:
            ctx.set_params_device(K, dsets.ptr + j * set_bytes)
        st, e, g, inf = ctx.energy_grad(0, 1, sigma, 1e-12, 30)
        sigma = e - 1e-6 * abs(e)
        return e, g, inf

    for i in range(args.warmup):
        e, g, inf = step(i, "warmup")
        print("# warmup %d E=%.15f iters=%d" % (i, e, inf[0]), file=sys.stderr, flush=True)
    times = []
    adj_ms = []
    off_ms = []
    eig_ms = []
    iters = []
    lines = []
    dsets.sync()
    t_all0 = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        e, g, inf = step(i)
        times.append(time.perf_counter() - t0)
        tm = ctx.timings()
        adj_ms.append(tm[4])
        off_ms.append(tm[5])
        eig_ms.append(tm[1])
        iters.append(inf[0])
        lines.append("# step %d E=%.15f |g|=%.3e ms=%.3f (assembly %.3f eigen %.3f adjoint %.3f)"
                     % (i, e, np.linalg.norm(g), 1e3 * times[-1], tm[0], tm[1], tm[2]))
    dsets.sync()
    t_all = time.perf_counter() - t_all0
    print("\n".join(lines), file=sys.stderr, flush=True)
    s_per = t_all / args.steps
    npairs = K * (K + 1) // 2
    f0, f1, f3, flops_adj = algorithmic_flops(system, base, K)
    # traced pass (not timed): per-kernel device time of the same steps
    ctx.trace(True)
    for i in range(ntr):
        step(i, "traced")
    trace = ctx.trace_read()
    ctx.trace(False)
    # the same steps with the parameters handed over in host memory (the
    # Fortran host's path: upload over PCIe inside each step)
    dsets.sync()
    t_h0 = time.perf_counter()
    for i in range(args.steps):
        e_last, _, _ = step(i, "timed", host=True)
    s_per_host = (time.perf_counter() - t_h0) / args.steps
    cond = conditioning(ctx, K, e_last)
    rows, dominant, _ = kernel_table(system, base, K, trace, ntr, flops_adj)
    dk = rows[dominant]
    names = {"ldl_update": "k_ldl_update (trailing update + look-ahead diagonal block + next panel + L^-1 tiles)",
             "adjoint": "k_adjoint<3> (fused E+grad pair pass)",
             "trsv_fwd": "k_xsolve_fwd (z = D^-1 L^-1 b)", "trsv_bwd": "k_xsolve_bwd + k_xsolve_reduce (x = L^-H z)",
             "offdiag": "k_offdiag<3>"}
    rocname = {"ldl_update": "k_ldl_update", "adjoint": "k_adjoint<3>", "trsv_fwd": "k_xsolve_fwd",
               "trsv_bwd": "k_xsolve_bwd", "offdiag": "k_offdiag<3>"}
    out = {
        "metric": "FP64 pair H,S+grad elements/s (E+grad, K=2000, n=3)",
        "value": npairs / s_per,
        "unit": "pair elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * s_per,
        "s_per_energy_grad": s_per,
        "host_params": {"value": npairs / s_per_host, "ms_per_step": 1e3 * s_per_host,
                        "note": "same steps with the parameters uploaded from host memory in each step"},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded ECG basis; each step's parameters staged in HBM before timing)",
        "config": {"workload": "H2-like n=3 (T=4) energy+gradient, K=2000 complex ECGs, seed %d" % (SEED_BASE + 3),
                   "K": K, "n": 3, "T": system.T, "pairs": npairs},
        "phases_ms": {"offdiag_kernel": float(np.mean(off_ms)), "eigen_step": float(np.mean(eig_ms)),
                      "adjoint_kernel": float(np.mean(adj_ms)), "inverse_iterations": float(np.mean(iters))},
        "roofline": {
            "bound": dk["bound"],
            "kernel": names.get(dominant, dominant),
            "achieved": dk.get("achieved", dk.get("equivalent_TFLOPs")),
            "peak": dk["peak"],
            "unit": dk["unit"],
            "frac": dk.get("achieved", dk.get("equivalent_TFLOPs")) / dk["peak"],
            "executed_TFLOPs": dk.get("executed_TFLOPs"),
            "executed_frac": dk.get("executed_TFLOPs", float("nan")) / dk["peak"],
            "traffic": pmc_traffic(rocname.get(dominant, dominant)),
            "work_per_launch": dk["work"],
            "avg_launch_us": dk["avg_us"],
            "share_of_step": dk["ms_per_step"] / (1e3 * s_per),
        },
        "kernels": rows,
        "flops_per_pair_term": {"energy": f0, "grad_bra": f1, "grad_both": f3},
        "conditioning": cond,
    }
    if not args.no_opt:
        out["optimise_one_function"] = optimise_leg(system, base, K)
    if not args.no_expect:
        out["expectation_values"] = expectation_leg(system, base, K)
    if not args.no_cycle:
        out["opt_cycle"] = opt_cycle_leg(system, base, K)
    if not args.no_enlarge:
        out["basis_enlargement"] = enlarge_leg(system, base, K)
    if not args.no_screen:
        out["screening"] = screening_leg(system, base, K)
    if not args.no_modes:
        out["memory_modes"] = memory_modes_leg(system)
    if not args.no_batch:
        out["batch"] = batch_leg(config=tuple(int(v) for v in args.batch_config.split(",")) if args.batch_config
                                 else None)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(system, base, K)
    ctx.close()
    dsets.free()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
