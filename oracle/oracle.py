"""ctypes front end of the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as part of the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.oracle_flops_per_pair.restype = ctypes.c_longlong
        L.oracle_flops_per_pair.argtypes = [ctypes.c_int, _i, _d, ctypes.c_double, _d, _d, _d, ctypes.c_int]
        L.oracle_pair.restype = None
        L.oracle_pair.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _d, ctypes.c_double, _d, _d, _d,
                                  ctypes.c_int, _d, _d, _d, _d]
        L.oracle_assemble.restype = ctypes.c_int
        L.oracle_assemble.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _d, ctypes.c_double, _d,
                                      ctypes.c_int, _d, ctypes.c_int, ctypes.c_int, ctypes.c_char,
                                      ctypes.c_double, _d, _d, ctypes.c_int, _d, _d, ctypes.c_int]
        L.oracle_assemble_deriv.restype = ctypes.c_int
        L.oracle_assemble_deriv.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _d, ctypes.c_double, _d,
                                            ctypes.c_int, _d, ctypes.c_int, ctypes.c_char,
                                            ctypes.c_double, _d, _d, ctypes.c_int, _d, _d, _d, ctypes.c_int]
        L.oracle_gradient.restype = None
        L.oracle_gradient.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _d, _d, ctypes.c_double, _d]
        L.oracle_pair_batch.restype = ctypes.c_double
        L.oracle_pair_batch.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _d, ctypes.c_double, _d, _d,
                                        ctypes.c_long, _i, _i, ctypes.c_int]
        L.oracle_ldlhf.restype = ctypes.c_int
        L.oracle_ldlhf.argtypes = [ctypes.c_int, ctypes.c_int, _d, ctypes.c_int, _d]
        L.oracle_ldlhs.restype = None
        L.oracle_ldlhs.argtypes = [ctypes.c_int, _d, ctypes.c_int, _d, _d, _d]
        L.oracle_ghepiis.restype = ctypes.c_int
        L.oracle_ghepiis.argtypes = [ctypes.c_int, ctypes.c_int, _d, ctypes.c_int, _d, _d, ctypes.c_int,
                                     ctypes.c_double, _d, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                     _d, _d]
        L.oracle_overlap_penalty.restype = ctypes.c_double
        L.oracle_overlap_penalty.argtypes = [ctypes.c_int, ctypes.c_int, _d, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double]
        L.oracle_overlap_penalty_grad.restype = ctypes.c_double
        L.oracle_overlap_penalty_grad.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _d, ctypes.c_int, _d,
                                                  _d, ctypes.c_double, ctypes.c_double, _d]
        L.oracle_expc_pair.restype = None
        L.oracle_expc_pair.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _d, ctypes.c_double, _d, _d, _d,
                                       ctypes.c_int, _d]
        L.oracle_expectation_values.restype = ctypes.c_int
        L.oracle_expectation_values.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _d, ctypes.c_double, _d,
                                                ctypes.c_int, _d, _d, _d, _d, _d, ctypes.c_int]
        L.oracle_correlation_functions.restype = ctypes.c_int
        L.oracle_correlation_functions.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _d, ctypes.c_double, _d,
                                                   ctypes.c_int, _d, _d, _d, ctypes.c_int, _d, _d, ctypes.c_int]
        L.oracle_particle_densities.restype = ctypes.c_int
        L.oracle_particle_densities.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _d, ctypes.c_double, _d,
                                                ctypes.c_int, _d, _d, _d, _d, ctypes.c_int, _d, _d, ctypes.c_int]
        L.oracle_relat_pair.restype = None
        L.oracle_relat_pair.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _d, ctypes.c_double, _d, _d, _d,
                                        ctypes.c_int, _d, _d]
        L.oracle_relativistic.restype = ctypes.c_int
        L.oracle_relativistic.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _d, ctypes.c_double, _d,
                                          ctypes.c_int, _d, _d, _d, _d, _d, _d, ctypes.c_int]
        L.oracle_drach_pair.restype = None
        L.oracle_drach_pair.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _d, ctypes.c_double, _d, _d, _d,
                                        ctypes.c_int, _d, ctypes.c_double, _d]
        L.oracle_drachmanized.restype = ctypes.c_int
        L.oracle_drachmanized.argtypes = [ctypes.c_int, ctypes.c_int, _i, _i, _d, ctypes.c_double, _d,
                                          ctypes.c_int, _d, _d, _d, _d, ctypes.c_double, _d, _d, ctypes.c_int]
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(_d)


def _pi(a):
    return a.ctypes.data_as(_i)


class _Sys:
    """Flattened constant inputs of an EcgSystem for the C calls."""

    def __init__(self, system):
        self.n = system.n
        self.T = system.T
        # term t's matrix stored column-major: P[i + n*j + n*n*t] = P_t[i, j]
        self.P = np.ascontiguousarray(np.transpose(system.yhy_mats, (0, 2, 1)).reshape(-1), dtype=np.int32)
        self.coef = np.ascontiguousarray(system.yhy_coef, dtype=np.int32)
        self.M = np.ascontiguousarray(system.mass_matrix.T.reshape(-1), dtype=np.float64)
        self.q0 = float(system.q0)
        self.q = np.ascontiguousarray(system.q, dtype=np.float64)

    def args(self):
        return (self.n, self.T, _pi(self.P), _pi(self.coef), _pd(self.M), self.q0, _pd(self.q))


def pair(system, pk, pl, gradflag=0, term=None):
    """(H, S, Dk, Dl) of one pair summed over the Y^dagger Y terms (or of a
    single term index ``term`` with coefficient 1)."""
    s = _Sys(system)
    n = s.n
    if term is not None:
        s.P = np.ascontiguousarray(s.P[n * n * term:n * n * (term + 1)])
        s.coef = np.ones(1, dtype=np.int32)
        s.T = 1
    pk = np.ascontiguousarray(pk, dtype=np.float64)
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    H = np.zeros(1, np.complex128)
    S = np.zeros(1, np.complex128)
    np4 = 2 * n * (n + 1)
    Dk = np.zeros(np4, np.complex128)
    Dl = np.zeros(np4, np.complex128)
    lib().oracle_pair(*s.args(), _pd(pk), _pd(pl), gradflag, _pd(H.view(np.float64)),
                      _pd(S.view(np.float64)), _pd(Dk.view(np.float64)), _pd(Dl.view(np.float64)))
    return H[0], S[0], Dk, Dl


def flops_per_pair(system, pk, pl, gradflag, term=0):
    s = _Sys(system)
    n = s.n
    P = np.ascontiguousarray(s.P[n * n * term:n * n * (term + 1)])
    pk = np.ascontiguousarray(pk, dtype=np.float64)
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    return int(lib().oracle_flops_per_pair(n, _pi(P), _pd(s.M), s.q0, _pd(s.q), _pd(pk), _pd(pl), gradflag))


def pair_batch(system, params, kidx, lidx, nthreads=0):
    """Evaluate pairs (kidx[i], lidx[i]) with gradients (CPU baseline probe)."""
    s = _Sys(system)
    params = np.ascontiguousarray(params, dtype=np.float64)
    k = np.ascontiguousarray(kidx, dtype=np.int32)
    l = np.ascontiguousarray(lidx, dtype=np.int32)
    return lib().oracle_pair_batch(*s.args(), _pd(params), k.size, _pi(k), _pi(l), nthreads)


def assemble(system, params, method="G", sigma=0.0, nmin=1, nmax=None, H=None, S=None,
             diagH=None, diagS=None, nthreads=0):
    """ComputeMatElem + StoreHS.  params: (K, npt).  Returns (H, S, diagH,
    diagS) as Fortran-ordered complex (K, K) arrays (lower triangles filled
    as the reference stores them)."""
    s = _Sys(system)
    params = np.ascontiguousarray(params, dtype=np.float64)
    K = params.shape[0]
    nmax = K if nmax is None else nmax
    if H is None:
        H = np.zeros((K, K), np.complex128, order="F")
        S = np.zeros((K, K), np.complex128, order="F")
        diagH = np.zeros(K)
        diagS = np.zeros(K)
    rc = lib().oracle_assemble(*s.args(), K, _pd(params), nmin, nmax, method.encode(), sigma,
                               H.ctypes.data_as(_d), S.ctypes.data_as(_d), K, _pd(diagH), _pd(diagS),
                               nthreads)
    if rc != 0:
        raise ValueError("oracle_assemble failed with %d" % rc)
    return H, S, diagH, diagS


def assemble_deriv(system, params, nfru=0, method="G", sigma=0.0, nthreads=0, H=None, S=None, diagH=None,
                   diagS=None):
    """ComputeMatElemAndDeriv(nfru+1, K) + StoreHSD.  Returns (H, S, diagH,
    diagS, D) with D of shape (2*npt, nfo, K) complex, Fortran order.  As in
    the reference, rows <= nfru are not recomputed: pass the arrays of an
    earlier assembly (their diagS is used for the normalisation)."""
    s = _Sys(system)
    params = np.ascontiguousarray(params, dtype=np.float64)
    K = params.shape[0]
    npt = system.npt
    if H is None:
        if nfru > 0:
            H, S, diagH, diagS = assemble(system, params, method, sigma, nthreads=nthreads)
        else:
            H = np.zeros((K, K), np.complex128, order="F")
            S = np.zeros((K, K), np.complex128, order="F")
            diagH = np.zeros(K)
            diagS = np.zeros(K)
    D = np.zeros((2 * npt, K - nfru, K), np.complex128, order="F")
    rc = lib().oracle_assemble_deriv(*s.args(), K, _pd(params), nfru, method.encode(), sigma,
                                     H.ctypes.data_as(_d), S.ctypes.data_as(_d), K, _pd(diagH),
                                     _pd(diagS), D.ctypes.data_as(_d), nthreads)
    if rc != 0:
        raise ValueError("oracle_assemble_deriv failed with %d" % rc)
    return H, S, diagH, diagS, D


def full_hermitian(H, S, diagH):
    """The symmetrisation EnergyGA does before ZHEGVX
    (src/workproc.f90:1737-1748): lower -> upper with conj, diagonals from
    diagH and 1."""
    K = H.shape[0]
    Hf = np.tril(H, -1)
    Hf = Hf + Hf.conj().T + np.diag(diagH.astype(np.complex128))
    Sf = np.tril(S, -1)
    Sf = Sf + Sf.conj().T + np.eye(K, dtype=np.complex128)
    return Hf, Sf


def gradient(D, c, E, nfru=0):
    """EnergyGB gradient contraction.  D: (2*npt, nfo, K) complex."""
    np4, nfo, K = D.shape
    npt = np4 // 2
    D = np.asfortranarray(D)
    c = np.ascontiguousarray(c, dtype=np.complex128)
    g = np.zeros(npt * nfo)
    lib().oracle_gradient(npt, K, nfru, D.ctypes.data_as(_d), c.ctypes.data_as(_d), float(E), _pd(g))
    return g


def overlap_penalty(S, nfru, bmax, t2):
    """ComputeOverlapPenalty on a stored normalised S (Fortran complex
    (K, K), lower triangle read)."""
    assert S.flags.f_contiguous
    K = S.shape[0]
    return lib().oracle_overlap_penalty(K, nfru, S.ctypes.data_as(_d), K, float(bmax), float(t2))


def overlap_penalty_grad(S, diagS, D, nfru, bmax, t2, grad=None):
    """ComputeOverlapPenaltyAndAddGradient: returns (P, grad) with the
    penalty's gradient terms added to grad (zeros if None).  D: (2*npt,
    nfo, K) complex as assemble_deriv returns it."""
    assert S.flags.f_contiguous
    np4, nfo, K = D.shape
    npt = np4 // 2
    D = np.asfortranarray(D)
    g = np.zeros(npt * nfo) if grad is None else np.array(grad, dtype=np.float64)
    P = lib().oracle_overlap_penalty_grad(npt, K, nfru, S.ctypes.data_as(_d), K, _pd(np.ascontiguousarray(diagS)),
                                          D.ctypes.data_as(_d), float(bmax), float(t2), _pd(g))
    return P, g


def n_expectation_values(n):
    """Number of values of expectation_values for n pseudo-particles."""
    return 5 * n * (n + 1) // 2 + 4


def expc_pair(system, pk, pl, term=0):
    """One pair's operators of MatrixElementsForExpcVals for symmetry term
    `term` (complex array of n_expectation_values(n) values: 1/r^2, 1/r, r,
    r^2, delta(r) for (a, b >= a), then H, S, T, V)."""
    s = _Sys(system)
    out = np.zeros(n_expectation_values(system.n), np.complex128)
    lib().oracle_expc_pair(*s.args(), _pd(np.ascontiguousarray(pk, dtype=np.float64)),
                           _pd(np.ascontiguousarray(pl, dtype=np.float64)), int(term), out.ctypes.data_as(_d))
    return out


def expectation_values(system, params, c, diagS, nthreads=0):
    """The reference driver's expectation values in the eigenvector c
    (complex K, c^H S c = 1 for the normalised S) and their rounding-error
    estimates: (values, errors), each n_expectation_values(n) reals."""
    s = _Sys(system)
    params = np.ascontiguousarray(params, dtype=np.float64)
    K = params.shape[0]
    nobs = n_expectation_values(system.n)
    out = np.zeros(nobs)
    err = np.zeros(nobs)
    lib().oracle_expectation_values(*s.args(), K, _pd(params), np.ascontiguousarray(c, dtype=np.complex128).ctypes.data_as(_d),
                                    _pd(np.ascontiguousarray(diagS, dtype=np.float64)), _pd(out), _pd(err), nthreads)
    return out, err


def correlation_functions(system, params, c, diagS, grid, nthreads=0):
    """The reference driver's pair correlation functions at the radii in
    grid, in the vector c: (ngrid, m) reals, row g for grid[g]."""
    s = _Sys(system)
    params = np.ascontiguousarray(params, dtype=np.float64)
    K = params.shape[0]
    m = system.n * (system.n + 1) // 2
    g = np.ascontiguousarray(grid, dtype=np.float64).ravel()
    out = np.zeros((g.size, m))
    lib().oracle_correlation_functions(*s.args(), K, _pd(params), np.ascontiguousarray(c, dtype=np.complex128).ctypes.data_as(_d),
                                       _pd(np.ascontiguousarray(diagS, dtype=np.float64)), g.size, _pd(g), _pd(out),
                                       nthreads)
    return out


def particle_densities(system, params, c, diagS, masses, grid, nthreads=0):
    """The reference driver's particle densities (particles 1..n+1, masses
    in that order) at the radii in grid, in the vector c: (ngrid, n + 1)."""
    s = _Sys(system)
    params = np.ascontiguousarray(params, dtype=np.float64)
    K = params.shape[0]
    g = np.ascontiguousarray(grid, dtype=np.float64).ravel()
    ms = np.ascontiguousarray(masses, dtype=np.float64).ravel()
    assert ms.size == system.n + 1
    out = np.zeros((g.size, system.n + 1))
    lib().oracle_particle_densities(*s.args(), K, _pd(params), np.ascontiguousarray(c, dtype=np.complex128).ctypes.data_as(_d),
                                    _pd(np.ascontiguousarray(diagS, dtype=np.float64)), _pd(ms), g.size, _pd(g),
                                    _pd(out), nthreads)
    return out


def relat_pair(system, pk, pl, term, masses):
    """(MV_kl, Darwin_kl) of one pair and symmetry term (complex)."""
    s = _Sys(system)
    out = np.zeros(4)
    lib().oracle_relat_pair(*s.args(), _pd(np.ascontiguousarray(pk, dtype=np.float64)),
                            _pd(np.ascontiguousarray(pl, dtype=np.float64)), term,
                            _pd(np.ascontiguousarray(masses, dtype=np.float64)), _pd(out))
    return complex(out[0], out[1]), complex(out[2], out[3])


def relativistic(system, params, c, diagS, masses, nthreads=0):
    """The driver's mass-velocity and Darwin values in the vector c and
    their rounding estimates: ([MV, Darwin], [MV_e, Darwin_e])."""
    s = _Sys(system)
    params = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros(2)
    err = np.zeros(2)
    lib().oracle_relativistic(*s.args(), params.shape[0], _pd(params),
                              np.ascontiguousarray(c, dtype=np.complex128).ctypes.data_as(_d),
                              _pd(np.ascontiguousarray(diagS, dtype=np.float64)),
                              _pd(np.ascontiguousarray(masses, dtype=np.float64)), _pd(out), _pd(err), nthreads)
    return out, err


def n_drach_values(n):
    """Number of values of drach_pair / drachmanized for n pseudo-particles."""
    return n * (n + 1) // 2 + 3


def drach_pair(system, pk, pl, term, masses, E):
    """Drachmanized delta(r) of every pair (a, b >= a), then the orbit-orbit
    term, the Drachmanized mass-velocity and Darwin terms of one pair and
    symmetry term, at the current energy E (complex, n_drach_values(n))."""
    s = _Sys(system)
    out = np.zeros(n_drach_values(system.n), np.complex128)
    lib().oracle_drach_pair(*s.args(), _pd(np.ascontiguousarray(pk, dtype=np.float64)),
                            _pd(np.ascontiguousarray(pl, dtype=np.float64)), int(term),
                            _pd(np.ascontiguousarray(masses, dtype=np.float64)), float(E), out.ctypes.data_as(_d))
    return out


def drachmanized(system, params, c, diagS, masses, E, nthreads=0):
    """The driver's Drachmanized values and orbit-orbit term in the vector c
    and their rounding estimates: (values, errors), n_drach_values(n) reals."""
    s = _Sys(system)
    params = np.ascontiguousarray(params, dtype=np.float64)
    nv = n_drach_values(system.n)
    out = np.zeros(nv)
    err = np.zeros(nv)
    lib().oracle_drachmanized(*s.args(), params.shape[0], _pd(params),
                              np.ascontiguousarray(c, dtype=np.complex128).ctypes.data_as(_d),
                              _pd(np.ascontiguousarray(diagS, dtype=np.float64)),
                              _pd(np.ascontiguousarray(masses, dtype=np.float64)), float(E), _pd(out), _pd(err),
                              nthreads)
    return out, err


def ghepiis(k, M, invD, B, apprlambda, v, tol, maxiter=20, specif_norm=0):
    """GHEPIIS restatement.  M, B: Fortran complex (n, n) arrays (M is
    updated in place), invD complex (n,) (updated), v complex (n,).
    Returns (err, lambda, x, relacc, numiter)."""
    n = M.shape[0]
    assert M.flags.f_contiguous and B.flags.f_contiguous
    v = np.ascontiguousarray(v, dtype=np.complex128)
    x = np.zeros(n, np.complex128)
    out3 = np.zeros(3)
    err = lib().oracle_ghepiis(k, n, M.ctypes.data_as(_d), n, invD.ctypes.data_as(_d),
                               B.ctypes.data_as(_d), n, float(apprlambda), v.ctypes.data_as(_d),
                               float(tol), maxiter, specif_norm, x.ctypes.data_as(_d), _pd(out3))
    return err, out3[0], x, out3[1], int(out3[2])
