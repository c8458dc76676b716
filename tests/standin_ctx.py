"""A host stand-in for EcgContext in the CPU tests of the optimiser loops.

E(params) = sum_k phi(p_k) with phi(p) = -exp(-|p - t|^2): adding any
function lowers the energy, every function's optimum is t, and the value
does not depend on the order of the functions, so the loops' permutations
can be checked through it.  The "overlap" of two functions is
S_ij = exp(-|p_i - p_j|^2 / 2); the penalty (when switched on) is the
reference's sum of b / (1 - t2) (|S_ij|^2 - t2) over the pairs j > i,
j > nfru, with |S_ij|^2 > t2 (src/workproc.f90:1491-1693), added to E
and, with its gradient, to the gradient.  Statuses can be forced to fail
for chosen calls."""
from __future__ import annotations

import numpy as np


class StandinError(RuntimeError):
    def __init__(self, code):
        super().__init__("stand-in status %d" % code)
        self.code = code


class SeparableContext:
    def __init__(self, target, coeff=1.0):
        self.t = np.asarray(target, dtype=np.float64)
        self.p = np.zeros((0, self.t.size))
        self.coeff = coeff               # value of every linear coefficient (callable: of the call count)
        self.pen = (False, 1.0, 1.0, 0)  # allowed, b, t2, nfru
        self.fail_calls = set()          # energy-call numbers (0-based) that fail with status 1
        self.calls = 0
        self.log = []                    # (kind, K, E) of every energy call
        self.screened = 0
        self.drift = 0.0                 # E gets -drift x (number of calls so far)

    # ------------------------------------------------------------ helpers
    def set_params(self, p):
        self.p = np.array(p, dtype=np.float64).reshape(-1, self.t.size)

    @property
    def K(self):
        return len(self.p)

    def _phi(self, p):
        d = p - self.t
        return -np.exp(-np.sum(d * d, axis=-1))

    def _penalty(self, nfru, b, t2, with_grad):
        K = self.K
        P = 0.0
        g = np.zeros((K, self.t.size))
        for j in range(max(nfru, 1), K):
            for i in range(j):
                d = self.p[j] - self.p[i]
                s2 = np.exp(-np.sum(d * d))       # |S_ij|^2
                if s2 > t2:
                    P += b / (1.0 - t2) * (s2 - t2)
                    if with_grad:
                        gj = b / (1.0 - t2) * s2 * (-2.0 * d)
                        g[j] += gj
                        g[i] -= gj
        return P, g[nfru:].reshape(-1)

    def _status(self):
        k = self.calls
        self.calls += 1
        if k in self.fail_calls:
            raise StandinError(1)

    def _coeffs(self):
        v = self.coeff(self.calls) if callable(self.coeff) else self.coeff
        return np.full(self.K, v, dtype=np.complex128)

    # ------------------------------------------------- EcgContext surface
    def energy(self, which, sigma, tol, maxiter, want_vector=False):
        self._status()
        E = float(np.sum(self._phi(self.p))) - self.drift * self.calls
        info = np.zeros(8)
        info[7] = -1
        if self.pen[0]:
            P, _ = self._penalty(self.pen[3], self.pen[1], self.pen[2], False)
            E += P
            info[6] = P
        self.log.append(("E", self.K, E))
        return 0, E, (self._coeffs() if want_vector else None), info

    def energy_grad(self, nfru, which, sigma, tol, maxiter):
        self._status()
        d = self.p - self.t
        e = np.exp(-np.sum(d * d, axis=-1))
        E = float(-np.sum(e)) - self.drift * self.calls
        g = (2.0 * d * e[:, None])[nfru:].reshape(-1)
        info = np.zeros(8)
        info[7] = -1
        if self.pen[0]:
            P, gp = self._penalty(nfru, self.pen[1], self.pen[2], True)
            E += P
            g = g + gp
            info[6] = P
        self.log.append(("G", self.K, E))
        return 0, E, g, info

    def set_overlap_penalty(self, allowed, max_pair_penalty=1.0, threshold2=1.0, nfru=0):
        self.pen = (bool(allowed), float(max_pair_penalty), float(threshold2), int(nfru))

    def overlap_penalty(self, nfru, max_pair_penalty, threshold2, grad=None):
        P, _ = self._penalty(nfru, max_pair_penalty, threshold2, False)
        return P, grad

    def screen_prepare(self, sigma, tol=1e-12, maxiter=40):
        self.fixed = float(np.sum(self._phi(self.p[:-1]))) - self.drift * self.calls
        return 0, np.array([self.fixed + self._phi(self.p[-1]), 0, 1, 0])

    def screen(self, cand, iters=4):
        cand = np.asarray(cand).reshape(-1, self.t.size)
        self.screened += len(cand)
        return self.fixed + self._phi(cand), np.zeros((len(cand), 4))
