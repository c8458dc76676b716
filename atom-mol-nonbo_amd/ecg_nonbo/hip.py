"""ctypes binding of libecg_hip.so (the C-ABI in include/ecg.h).

Mirrors the reference's subroutine interfaces on the Python side:
``MatrixElements`` (src/matelem.f90:12), ``ComputeMatElem`` /
``ComputeMatElemAndDeriv`` (src/matform.f90:119, 234), ``GHEPIIS``
(src/linalg.f90:1781) and the energy evaluators ``EnergyGA``/``EnergyIA``
and ``EnergyGB``/``EnergyIB`` (src/workproc.f90:1697-2154).

There is no fallback: if the shared library is missing or no GPU is
present, the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# ECG_HIP_LIB selects another build of the library (A/B comparisons)
LIB_PATH = os.environ.get("ECG_HIP_LIB", os.path.join(_HERE, "libecg_hip.so"))

ECG_OK = 0
ECG_ERR_ARG = -1
ECG_ERR_HIP = -2
ECG_ERR_NOMEM = -3
ECG_ERR_NONFINITE = -4
ECG_ERR_STATE = -5
ECG_ERR_SYNC = -6
ECG_SINGULAR = 1
ECG_NOT_CONVERGED = 2
ECG_ROOT_NOT_FOUND = 3

_NAMES = {
    ECG_ERR_ARG: "invalid argument", ECG_ERR_HIP: "HIP runtime error", ECG_ERR_NOMEM: "device allocation failed",
    ECG_ERR_NONFINITE: "non-finite pair element", ECG_ERR_STATE: "call order violated",
    ECG_ERR_SYNC: "device-side wait timed out",
    ECG_SINGULAR: "H - sigma S singular", ECG_NOT_CONVERGED: "inverse iteration not converged",
    ECG_ROOT_NOT_FOUND: "requested root not isolated",
}

EXPORTS = [
    "ecg_ctx_create", "ecg_ctx_destroy", "ecg_set_params", "ecg_update_params", "ecg_matrix_elements",
    "ecg_compute_mat_elem", "ecg_compute_mat_elem_and_deriv", "ecg_ghepiis", "ecg_energy", "ecg_energy_grad",
    "ecg_energy_grad_stored", "ecg_batch_create", "ecg_batch_destroy", "ecg_batch_set_config",
    "ecg_batch_energy_grad", "ecg_trace", "ecg_trace_read", "ecg_get_params", "ecg_screen_prepare", "ecg_screen", "ecg_screen_loop", "ecg_last_timings",
    "ecg_set_overlap_penalty", "ecg_overlap_penalty", "ecg_set_params_device", "ecg_expectation_values",
    "ecg_correlation_functions", "ecg_particle_densities", "ecg_relativistic_corrections", "ecg_drachmanized_corrections", "ecg_eigenvector",
    "ecg_matrix_elements_noalloc", "ecg_screen_set_shapes",
]


class EcgError(RuntimeError):
    def __init__(self, func, code):
        super().__init__("%s returned %d (%s)" % (func, code, _NAMES.get(code, "?")))
        self.code = code


_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)
_lib = None


def load(path: str = LIB_PATH):
    """Load the library and declare the prototypes (no GPU is touched)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError("%s is not built; run atom-mol-nonbo_amd/build.py" % path)
    L = ctypes.CDLL(path)
    c_int, c_double, c_char = ctypes.c_int, ctypes.c_double, ctypes.c_char
    vp = ctypes.c_void_p
    L.ecg_ctx_create.argtypes = [c_int, c_int, _i, _i, _d, c_double, _d, ctypes.POINTER(vp)]
    L.ecg_ctx_destroy.argtypes = [vp]
    L.ecg_set_params.argtypes = [vp, c_int, _d]
    L.ecg_update_params.argtypes = [vp, c_int, c_int, _d]
    L.ecg_matrix_elements.argtypes = [vp, _d, _d, _i, c_int, _d, _d, _d, _d]
    L.ecg_matrix_elements_noalloc.argtypes = [vp, _d, _d, _i, c_int, _d, _d, _d, _d]
    L.ecg_compute_mat_elem.argtypes = [vp, c_int, c_int, c_char, c_double, _d, _d, c_int, _d, _d]
    L.ecg_compute_mat_elem_and_deriv.argtypes = [vp, c_int, c_char, c_double, _d, _d, c_int, _d, _d, _d, c_int]
    L.ecg_ghepiis.argtypes = [vp, c_int, c_int, _d, c_int, _d, _d, c_int, c_double, _d, c_double, c_int, c_int,
                              _d, _d, _d, _i]
    L.ecg_energy.argtypes = [vp, c_int, c_double, c_double, c_int, _d, _d, _d]
    L.ecg_energy_grad.argtypes = [vp, c_int, c_int, c_double, c_double, c_int, _d, _d, _d]
    L.ecg_energy_grad_stored.argtypes = [vp, c_int, c_int, c_double, c_double, c_int, _d, _d, _d]
    L.ecg_screen_set_shapes.argtypes = [vp, c_int, c_int]
    L.ecg_batch_create.argtypes = [c_int, ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.ecg_batch_destroy.argtypes = [vp]
    L.ecg_batch_set_config.argtypes = [vp, c_int, c_int]
    L.ecg_batch_energy_grad.argtypes = [vp, _d, c_double, c_int, _d, _d, _d]
    L.ecg_trace.argtypes = [vp, c_int]
    L.ecg_trace_read.argtypes = [vp, _d, _d, c_int]
    L.ecg_get_params.argtypes = [vp, _d]
    L.ecg_screen_prepare.argtypes = [vp, c_double, c_double, c_int, _d]
    L.ecg_screen.argtypes = [vp, c_int, _d, c_int, _d, _d]
    L.ecg_screen_loop.argtypes = [vp, c_int, c_int, _d, c_int, _d, _i]
    L.ecg_last_timings.argtypes = [vp, _d, c_int]
    L.ecg_set_overlap_penalty.argtypes = [vp, c_int, c_double, c_double, c_int]
    L.ecg_set_params_device.argtypes = [vp, c_int, vp]
    L.ecg_overlap_penalty.argtypes = [vp, c_int, c_double, c_double, _d, _d]
    L.ecg_expectation_values.argtypes = [vp, _d, _d, _d]
    L.ecg_correlation_functions.argtypes = [vp, _d, c_int, _d, _d]
    L.ecg_particle_densities.argtypes = [vp, _d, _d, c_int, _d, _d]
    L.ecg_relativistic_corrections.argtypes = [vp, _d, _d, _d, _d]
    L.ecg_drachmanized_corrections.argtypes = [vp, _d, _d, ctypes.c_double, _d, _d]
    L.ecg_eigenvector.argtypes = [vp, _d]
    for name in EXPORTS:
        getattr(L, name).restype = c_int
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(_d)


def _check(func, rc, ok=(ECG_OK,)):
    if rc not in ok:
        raise EcgError(func, rc)
    return rc


class EcgContext:
    """Device state of one system and basis (one ecg_ctx)."""

    def __init__(self, system):
        L = load()
        self.system = system
        n = system.n
        self.n = n
        self.npt = system.npt
        P = np.ascontiguousarray(np.transpose(system.yhy_mats, (0, 2, 1)).reshape(-1), dtype=np.int32)
        coef = np.ascontiguousarray(system.yhy_coef, dtype=np.int32)
        M = np.ascontiguousarray(system.mass_matrix.T.reshape(-1), dtype=np.float64)
        q = np.ascontiguousarray(system.q, dtype=np.float64)
        h = ctypes.c_void_p()
        _check("ecg_ctx_create", L.ecg_ctx_create(n, system.T, P.ctypes.data_as(_i), coef.ctypes.data_as(_i),
                                                   _p(M), float(system.q0), _p(q), ctypes.byref(h)))
        self._h = h
        self.K = 0

    def close(self):
        if self._h is not None and self._h.value:
            load().ecg_ctx_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --------------------------------------------------------- parameters
    def set_params(self, params):
        """params: (K, npt) array, row k = (vech L_k, vech B_k)."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        K = p.shape[0]
        _check("ecg_set_params", load().ecg_set_params(self._h, K, _p(p)))
        self.K = K

    def set_params_device(self, K, dptr):
        """Parameters of K functions from device memory (an integer device
        address, e.g. a torch tensor's data_ptr())."""
        _check("ecg_set_params_device", load().ecg_set_params_device(self._h, int(K), ctypes.c_void_p(int(dptr))))
        self.K = int(K)

    def update_params(self, first, params):
        p = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
        _check("ecg_update_params", load().ecg_update_params(self._h, first, p.shape[0], _p(p)))

    # --------------------------------------------------- drop-in routines
    def MatrixElements(self, paramk, paraml, P, gradflag=0, noalloc=False):
        """(Hkl, Skl, Dk, Dl) of one pair and one symmetry matrix P
        (noalloc: through ecg_matrix_elements_noalloc, which stages the pair
        in the library's fixed device scratch instead of allocating)."""
        n = self.n
        pk = np.ascontiguousarray(paramk, dtype=np.float64)
        pl = np.ascontiguousarray(paraml, dtype=np.float64)
        Pf = np.ascontiguousarray(np.asarray(P).T.reshape(-1), dtype=np.int32)
        H = np.zeros(1, np.complex128)
        S = np.zeros(1, np.complex128)
        Dk = np.zeros(2 * n * (n + 1), np.complex128)
        Dl = np.zeros(2 * n * (n + 1), np.complex128)
        name = "ecg_matrix_elements_noalloc" if noalloc else "ecg_matrix_elements"
        _check(name, getattr(load(), name)(
            self._h, _p(pk), _p(pl), Pf.ctypes.data_as(_i), gradflag, _p(H.view(np.float64)),
            _p(S.view(np.float64)), _p(Dk.view(np.float64)), _p(Dl.view(np.float64))))
        return H[0], S[0], Dk, Dl

    def ComputeMatElem(self, nmin=1, nmax=None, method="G", approx_energy=0.0, H=None, S=None, diagH=None,
                       diagS=None):
        K = self.K
        nmax = K if nmax is None else nmax
        if H is None:
            H = np.zeros((K, K), np.complex128, order="F")
            S = np.zeros((K, K), np.complex128, order="F")
            diagH = np.zeros(K)
            diagS = np.zeros(K)
        assert H.flags.f_contiguous and S.flags.f_contiguous
        _check("ecg_compute_mat_elem", load().ecg_compute_mat_elem(
            self._h, nmin, nmax, method.encode(), float(approx_energy), H.ctypes.data_as(_d),
            S.ctypes.data_as(_d), H.shape[0], _p(diagH), _p(diagS)))
        return H, S, diagH, diagS

    def ComputeMatElemAndDeriv(self, nfru=0, method="G", approx_energy=0.0, H=None, S=None, diagH=None,
                               diagS=None, ldD=None):
        """Rows nfru+1..K with derivatives.  Rows <= nfru are not recomputed:
        pass the arrays of an earlier ComputeMatElem (their diagS normalises
        the new rows), as the reference's global state provides.  ldD: the
        second extent of the returned D (Glob_D's NumOfFuncToOpt, at least
        K - nfru); entries beyond K - nfru stay zero."""
        K = self.K
        if H is None:
            if nfru > 0:
                H, S, diagH, diagS = self.ComputeMatElem(1, K, method, approx_energy)
            else:
                H = np.zeros((K, K), np.complex128, order="F")
                S = np.zeros((K, K), np.complex128, order="F")
                diagH = np.zeros(K)
                diagS = np.zeros(K)
        ldD = K - nfru if ldD is None else int(ldD)
        D = np.zeros((2 * self.npt, ldD, K), np.complex128, order="F")
        _check("ecg_compute_mat_elem_and_deriv", load().ecg_compute_mat_elem_and_deriv(
            self._h, nfru, method.encode(), float(approx_energy), H.ctypes.data_as(_d), S.ctypes.data_as(_d),
            K, _p(diagH), _p(diagS), D.ctypes.data_as(_d), ldD))
        return H, S, diagH, diagS, D

    def GHEPIIS(self, k, M, invD, B, apprlambda, v, tol, maxiter=20, specif_norm=0):
        """Returns (ErrorCode, lambda, x, RelAcc, NumIter); M and invD are
        updated in place like the reference's arguments."""
        n = M.shape[0]
        assert M.flags.f_contiguous and B.flags.f_contiguous
        v = np.ascontiguousarray(v, dtype=np.complex128)
        x = np.zeros(n, np.complex128)
        lam = ctypes.c_double()
        rel = ctypes.c_double()
        nit = ctypes.c_int()
        rc = load().ecg_ghepiis(self._h, k, n, M.ctypes.data_as(_d), M.shape[0], invD.ctypes.data_as(_d),
                                B.ctypes.data_as(_d), B.shape[0], float(apprlambda), v.ctypes.data_as(_d),
                                float(tol), maxiter, specif_norm, ctypes.byref(lam), x.ctypes.data_as(_d),
                                ctypes.byref(rel), ctypes.byref(nit))
        _check("ecg_ghepiis", rc, ok=(ECG_OK, ECG_SINGULAR, ECG_NOT_CONVERGED))
        return rc, lam.value, x, rel.value, nit.value

    # --------------------------------------------------- device-resident
    def energy(self, which=1, sigma=float("nan"), tol=1e-12, maxiter=20, want_vector=False):
        """(status, E, c or None, info) for root `which`."""
        E = ctypes.c_double()
        info = np.zeros(8)
        c = np.zeros(self.K, np.complex128) if want_vector else None
        rc = load().ecg_energy(self._h, which, float(sigma), float(tol), maxiter, ctypes.byref(E),
                               c.ctypes.data_as(_d) if c is not None else None, _p(info))
        _check("ecg_energy", rc, ok=(ECG_OK, ECG_NOT_CONVERGED))
        return rc, E.value, c, info

    def energy_grad(self, nfru=0, which=1, sigma=float("nan"), tol=1e-12, maxiter=20):
        """(status, E, grad, info); grad ordered per function (dE/dvechL,
        dE/dvechB) for functions nfru+1..K."""
        E = ctypes.c_double()
        info = np.zeros(8)
        g = np.zeros(self.npt * (self.K - nfru))
        rc = load().ecg_energy_grad(self._h, nfru, which, float(sigma), float(tol), maxiter, ctypes.byref(E),
                                    _p(g), _p(info))
        _check("ecg_energy_grad", rc, ok=(ECG_OK, ECG_NOT_CONVERGED))
        return rc, E.value, g, info

    def energy_grad_stored(self, nfru=0, which=1, sigma=float("nan"), tol=1e-12, maxiter=20):
        """As energy_grad, with the per-pair derivatives stored on the device
        and contracted afterwards (the reference's scheme)."""
        E = ctypes.c_double()
        info = np.zeros(8)
        g = np.zeros(self.npt * (self.K - nfru))
        rc = load().ecg_energy_grad_stored(self._h, nfru, which, float(sigma), float(tol), maxiter,
                                           ctypes.byref(E), _p(g), _p(info))
        _check("ecg_energy_grad_stored", rc, ok=(ECG_OK, ECG_NOT_CONVERGED))
        return rc, E.value, g, info

    def screen_set_shapes(self, row_tile=0, cand_per_block=0):
        """Fix the screening step's GEMM row tile (32 or 64) and border-row
        candidates per workgroup (1, 2, 4, 8); 0 leaves a shape timed."""
        _check("ecg_screen_set_shapes", load().ecg_screen_set_shapes(self._h, int(row_tile), int(cand_per_block)))

    def set_overlap_penalty(self, allowed, max_pair_penalty=1.0, threshold2=1.0, nfru=0):
        """Glob_OverlapPenaltyAllowed / Glob_MaxOverlapPenalty /
        Glob_OverlapPenaltyThreshold2 (and nfru for energy-only calls)."""
        _check("ecg_set_overlap_penalty",
               load().ecg_set_overlap_penalty(self._h, int(bool(allowed)), float(max_pair_penalty),
                                              float(threshold2), int(nfru)))

    def overlap_penalty(self, nfru, max_pair_penalty, threshold2, grad=None):
        """(P, grad): the penalty of the device S; grad (if given, npt*(K -
        nfru)) is returned with the penalty's gradient added."""
        P = ctypes.c_double()
        g = None if grad is None else np.array(grad, dtype=np.float64)
        rc = load().ecg_overlap_penalty(self._h, int(nfru), float(max_pair_penalty), float(threshold2),
                                        ctypes.byref(P), _p(g) if g is not None else None)
        _check("ecg_overlap_penalty", rc)
        return P.value, g

    TRACE_LABELS = ("offdiag", "adjoint", "ldl_update", "ldl_diag", "ldl_panel", "trsv_fwd", "trsv_bwd", "gemv",
                    "form_shifted", "ldl_tail")

    def expectation_values(self, c=None):
        """(values, errors) of ecg_expectation_values: 5 m + 4 reals each,
        m = n(n+1)/2; c None uses the last energy call's eigenvector."""
        nobs = 5 * self.system.n * (self.system.n + 1) // 2 + 4
        vals = np.zeros(nobs)
        errs = np.zeros(nobs)
        cc = None if c is None else np.ascontiguousarray(c, dtype=np.complex128)
        _check("ecg_expectation_values", load().ecg_expectation_values(
            self._h, cc.ctypes.data_as(_d) if cc is not None else None, _p(vals), _p(errs)))
        return vals, errs

    def correlation_functions(self, grid, c=None):
        """ecg_correlation_functions: (ngrid, m) array, row g = the
        pseudo-particle pair correlation functions at radius grid[g]."""
        m = self.system.n * (self.system.n + 1) // 2
        g = np.ascontiguousarray(grid, dtype=np.float64).ravel()
        out = np.zeros((g.size, m))
        cc = None if c is None else np.ascontiguousarray(c, dtype=np.complex128)
        _check("ecg_correlation_functions", load().ecg_correlation_functions(
            self._h, cc.ctypes.data_as(_d) if cc is not None else None, g.size, _p(g), _p(out)))
        return out

    def particle_densities(self, grid, c=None, masses=None):
        """ecg_particle_densities: (ngrid, n + 1) array, row g = the
        densities of the n + 1 particles at radius grid[g]; masses default
        to the system's (particle 1 first)."""
        npart = self.system.n + 1
        ms = np.ascontiguousarray(self.system.masses if masses is None else masses, dtype=np.float64).ravel()
        if ms.size != npart:
            raise ValueError("need %d masses" % npart)
        g = np.ascontiguousarray(grid, dtype=np.float64).ravel()
        out = np.zeros((g.size, npart))
        cc = None if c is None else np.ascontiguousarray(c, dtype=np.complex128)
        _check("ecg_particle_densities", load().ecg_particle_densities(
            self._h, cc.ctypes.data_as(_d) if cc is not None else None, _p(ms), g.size, _p(g), _p(out)))
        return out

    def relativistic_corrections(self, c=None, masses=None):
        """ecg_relativistic_corrections: (values, errors) = ([MV, Darwin],
        their rounding estimates); masses default to the system's."""
        npart = self.system.n + 1
        ms = np.ascontiguousarray(self.system.masses if masses is None else masses, dtype=np.float64).ravel()
        if ms.size != npart:
            raise ValueError("need %d masses" % npart)
        vals = np.zeros(2)
        errs = np.zeros(2)
        cc = None if c is None else np.ascontiguousarray(c, dtype=np.complex128)
        _check("ecg_relativistic_corrections", load().ecg_relativistic_corrections(
            self._h, cc.ctypes.data_as(_d) if cc is not None else None, _p(ms), _p(vals), _p(errs)))
        return vals, errs

    def drachmanized_corrections(self, E, c=None, masses=None):
        """ecg_drachmanized_corrections at the current energy E: (values,
        errors), n(n+1)/2 + 3 each (Drachmanized delta of the pairs
        (a, b >= a), orbit-orbit, Drachmanized mass-velocity and Darwin);
        masses default to the system's."""
        n = self.system.n
        ms = np.ascontiguousarray(self.system.masses if masses is None else masses, dtype=np.float64).ravel()
        if ms.size != n + 1:
            raise ValueError("need %d masses" % (n + 1))
        nv = n * (n + 1) // 2 + 3
        vals = np.zeros(nv)
        errs = np.zeros(nv)
        cc = None if c is None else np.ascontiguousarray(c, dtype=np.complex128)
        _check("ecg_drachmanized_corrections", load().ecg_drachmanized_corrections(
            self._h, cc.ctypes.data_as(_d) if cc is not None else None, _p(ms), float(E), _p(vals), _p(errs)))
        return vals, errs

    def trace(self, enable=True):
        _check("ecg_trace", load().ecg_trace(self._h, int(bool(enable))))

    def trace_read(self):
        """{kernel: (launches, total device ms)} since tracing was enabled."""
        n = len(self.TRACE_LABELS)
        cnt = np.zeros(n)
        tot = np.zeros(n)
        m = load().ecg_trace_read(self._h, _p(cnt), _p(tot), n)
        if m < 0:
            raise EcgError("ecg_trace_read", m)
        return {self.TRACE_LABELS[i]: (int(cnt[i]), float(tot[i])) for i in range(m)}

    def get_params(self):
        p = np.zeros((self.K, self.npt))
        _check("ecg_get_params", load().ecg_get_params(self._h, _p(p)))
        return p

    # ---------------------------------------------------------- screening
    def screen_prepare(self, sigma, tol=1e-12, maxiter=40):
        """Fix functions 1..K-1 and the shift; (status, info) with info =
        [current E, negative pivots of H'-sigma S', iterations, measure]."""
        info = np.zeros(4)
        rc = load().ecg_screen_prepare(self._h, float(sigma), float(tol), maxiter, _p(info))
        _check("ecg_screen_prepare", rc, ok=(ECG_OK, ECG_NOT_CONVERGED))
        return rc, info

    def screen(self, cand, iters=4):
        """Energies of candidates (ncand x npt) for function K;
        (E, info) with info rows [measure, negative pivots, pivot, x^H S x]."""
        cand = np.ascontiguousarray(cand, dtype=np.float64)
        nc = cand.shape[0]
        E = np.zeros(nc)
        info = np.zeros((nc, 4))
        _check("ecg_screen", load().ecg_screen(self._h, nc, _p(cand), iters, _p(E), _p(info)))
        return E, info

    def screen_loop(self, cand, iters=4):
        """cand: (nsteps, ncand, npt); (Ebest, ibest) per step."""
        cand = np.ascontiguousarray(cand, dtype=np.float64)
        ns, nc = cand.shape[0], cand.shape[1]
        Eb = np.zeros(ns)
        ib = np.zeros(ns, np.int32)
        _check("ecg_screen_loop", load().ecg_screen_loop(self._h, ns, nc, _p(cand), iters, _p(Eb),
                                                         ib.ctypes.data_as(_i)))
        return Eb, ib

    def eigenvector(self):
        """The last energy call's eigenvector (c^H S c = 1), as EnergyGB
        leaves it in Glob_c; raises EcgError (ECG_ERR_STATE) after a
        parameter change."""
        c = np.zeros(self.K, np.complex128)
        _check("ecg_eigenvector", load().ecg_eigenvector(self._h, c.ctypes.data_as(_d)))
        return c

    def timings(self):
        t = np.zeros(8)
        n = load().ecg_last_timings(self._h, _p(t), 8)
        return t[:n]


class EcgBatch:
    """A packed batch of independent problems (one EcgContext each, with
    parameters set): ecg_batch_create / ecg_batch_energy_grad."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(*[c._h.value for c in self.contexts])
        h = ctypes.c_void_p()
        _check("ecg_batch_create", load().ecg_batch_create(len(self.contexts), arr, ctypes.byref(h)))
        self._h = h

    def set_config(self, packed=1, nstreams=4):
        _check("ecg_batch_set_config", load().ecg_batch_set_config(self._h, int(packed), int(nstreams)))

    def energy_grad(self, sigmas, tol=1e-12, maxiter=20):
        """(status, E per problem, list of gradients, latency ms per problem)."""
        sig = np.ascontiguousarray(sigmas, dtype=np.float64)
        n = len(self.contexts)
        E = np.zeros(n)
        sizes = [c.npt * c.K for c in self.contexts]
        g = np.zeros(int(np.sum(sizes)))
        lat = np.zeros(n)
        rc = load().ecg_batch_energy_grad(self._h, _p(sig), float(tol), maxiter, _p(E), _p(g), _p(lat))
        _check("ecg_batch_energy_grad", rc, ok=(ECG_OK, ECG_NOT_CONVERGED))
        offs = np.concatenate([[0], np.cumsum(sizes)])
        return rc, E, [g[offs[i]:offs[i + 1]] for i in range(n)], lat

    def close(self):
        if self._h is not None and self._h.value:
            load().ecg_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
