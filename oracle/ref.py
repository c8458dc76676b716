"""ctypes front end of the reference's own pair routines (oracle/_ref).

TEST INFRASTRUCTURE ONLY: the libraries are the reference's MatrixElements
and MatrixElementsForExpcVals (src/matelem.f90:12-750, 949-1983; with
ref_shim.f90 setting the globals) and its PORT minimiser (DIVSET / DRMNG,
src/dmng.f), compiled from the reference sources by `make -C oracle ref`.
Tests use them to check the oracle's restatement, the device and the
host-side minimiser against the reference's own code; they are absent
(available() / dmng_available() are False) when the reference tree was not
there to build them.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_ref", "libref_matelem.so")
_DMNG = os.path.join(_HERE, "_ref", "libref_dmng.so")
REFSRC = "/root/reference/src"
_lib = None
_dmng = None
_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)


def build():
    """Build _ref/libref_matelem.so and _ref/libref_dmng.so when the
    reference sources are present (None otherwise)."""
    if not os.path.isdir(REFSRC):
        return None
    subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL)
    return _LIB


def available() -> bool:
    return os.path.exists(_LIB)


def dmng_available() -> bool:
    return os.path.exists(_DMNG)


def dmng_lib():
    """The reference's DIVSET and DRMNG (Fortran symbols divset_, drmng_;
    every argument by reference)."""
    global _dmng
    if _dmng is None:
        L = ctypes.CDLL(_DMNG)
        L.divset_.restype = None
        L.divset_.argtypes = [_i, _i, _i, _i, _d]
        L.drmng_.restype = None
        L.drmng_.argtypes = [_d, _d, _d, _i, _i, _i, _i, _d, _d]
        _dmng = L
    return _dmng


def ref_divset(alg, iv, v):
    """DIVSET(ALG, IV, LIV, LV, V) on int32 iv and float64 v, in place."""
    a, li, lv = ctypes.c_int(alg), ctypes.c_int(iv.size), ctypes.c_int(v.size)
    dmng_lib().divset_(ctypes.byref(a), iv.ctypes.data_as(_i), ctypes.byref(li), ctypes.byref(lv),
                       v.ctypes.data_as(_d))


def ref_drmng(d, fx, g, iv, v, x):
    """DRMNG(D, FX, G, IV, LIV, LV, N, V, X) in place; returns IV(1)."""
    f = ctypes.c_double(fx)
    li, lv, n = ctypes.c_int(iv.size), ctypes.c_int(v.size), ctypes.c_int(x.size)
    dmng_lib().drmng_(d.ctypes.data_as(_d), ctypes.byref(f), g.ctypes.data_as(_d), iv.ctypes.data_as(_i),
                      ctypes.byref(li), ctypes.byref(lv), ctypes.byref(n), v.ctypes.data_as(_d),
                      x.ctypes.data_as(_d))
    return int(iv[0])


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(_LIB)
        L.ref_setup.restype = None
        L.ref_setup.argtypes = [ctypes.c_int, _d, ctypes.c_double, _d, ctypes.c_double]
        L.ref_matrix_elements.restype = None
        L.ref_matrix_elements.argtypes = [_d, _d, _i, ctypes.c_int, _d, _d, _d, _d]
        L.ref_expc_pair.restype = None
        L.ref_expc_pair.argtypes = [_d, _d, _i, _i, _d]
        _lib = L
    return _lib


def setup(system, curr_energy=0.0):
    """The reference's globals for `system` (masses, pseudo-charges)."""
    m = np.ascontiguousarray(system.masses, dtype=np.float64)
    q = np.ascontiguousarray(system.q, dtype=np.float64)
    lib().ref_setup(system.n, m.ctypes.data_as(_d), float(system.q0), q.ctypes.data_as(_d), float(curr_energy))


def _pmat(P):
    return np.ascontiguousarray(np.asarray(P, dtype=np.int32).T.reshape(-1))


def matrix_elements(system, pk, pl, P, gradflag=0):
    """(H, S, Dk, Dl) of MatrixElements for one symmetry matrix P (after setup)."""
    n = system.n
    pk = np.ascontiguousarray(pk, dtype=np.float64)
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    Pf = _pmat(P)
    H = np.zeros(1, np.complex128)
    S = np.zeros(1, np.complex128)
    Dk = np.zeros(2 * n * (n + 1), np.complex128)
    Dl = np.zeros(2 * n * (n + 1), np.complex128)
    lib().ref_matrix_elements(pk.ctypes.data_as(_d), pl.ctypes.data_as(_d), Pf.ctypes.data_as(_i), gradflag,
                              H.view(np.float64).ctypes.data_as(_d), S.view(np.float64).ctypes.data_as(_d),
                              Dk.view(np.float64).ctypes.data_as(_d), Dl.view(np.float64).ctypes.data_as(_d))
    return H[0], S[0], Dk, Dl


def pair_sum(system, pk, pl, gradflag=0):
    """MatrixElements summed over the system's Y^dagger Y terms with their
    coefficients, as ComputeMatElem(AndDeriv) sums them."""
    H = S = 0.0
    Dk = Dl = 0.0
    for t in range(system.T):
        h, s, dk, dl = matrix_elements(system, pk, pl, system.yhy_mats[t], gradflag)
        c = float(system.yhy_coef[t])
        H, S, Dk, Dl = H + c * h, S + c * s, Dk + c * dk, Dl + c * dl
    return H, S, Dk, Dl


def expc_pair(system, pk, pl, Pbra, Pket):
    """MatrixElementsForExpcVals for one pair: dict of H, S, T, V, MV,
    drach_MV, Darwin, drach_Darwin, OO and the n x n arrays rm2, rm, r, r2,
    delta, drach_delta (after setup with the current energy)."""
    n = system.n
    out = np.zeros(9 + 6 * n * n, np.complex128)
    pk = np.ascontiguousarray(pk, dtype=np.float64)
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    lib().ref_expc_pair(pk.ctypes.data_as(_d), pl.ctypes.data_as(_d), _pmat(Pbra).ctypes.data_as(_i),
                        _pmat(Pket).ctypes.data_as(_i), out.view(np.float64).ctypes.data_as(_d))
    keys = ("H", "S", "T", "V", "MV", "drach_MV", "Darwin", "drach_Darwin", "OO")
    d = {k: out[i] for i, k in enumerate(keys)}
    for b, k in enumerate(("rm2", "rm", "r", "r2", "delta", "drach_delta")):
        d[k] = out[9 + b * n * n:9 + (b + 1) * n * n].reshape(n, n, order="F")
    return d
