"""High-precision (mpmath) evaluation of the ECG pair closed forms, used to
pin the oracle independently of its FP64 arithmetic.

For phi_k = exp[-r^T (L_k L_k^T + i B_k) r] and a symmetry matrix P acting
on the ket (A_l -> P^T A_l P, B_l -> P^T B_l P), with Ck* = A_k - i B_k,
tCl = P^T (A_l + i B_l) P and C = Ck* + tCl:
  S = 2^(3n/2) z sqrt(z),  z = |det L_k det L_l| / det C        (principal sqrt)
  T = 6 S tr(M tCl C^-1 Ck*)
  V = (2/sqrt(pi)) S [q0 sum_i q_i (C^-1)_ii^-1/2
                      + sum_{i<j} q_i q_j (C^-1_ii + C^-1_jj - 2 C^-1_ij)^-1/2]
The formulas are evaluated with mpmath's own determinant and inverse (LU),
not with the reference's pseudo-Cholesky, and the parameter vector is
unpacked with an index map written independently of the oracle's.
"""
from __future__ import annotations

import mpmath as mp


def _unpack(p, n):
    m = n * (n + 1) // 2
    L = mp.zeros(n, n)
    B = mp.zeros(n, n)
    pos = 0
    for j in range(n):
        for i in range(j, n):
            L[i, j] = mp.mpf(p[pos])
            B[i, j] = mp.mpf(p[m + pos])
            B[j, i] = B[i, j]
            pos += 1
    return L, B


def pair_hs(system, pk, pl, P, dps=40):
    """(H, S) of one pair and one symmetry matrix at `dps` digits."""
    with mp.workdps(dps):
        n = system.n
        Lk, Bk = _unpack(pk, n)
        Ll, Bl = _unpack(pl, n)
        Pm = mp.matrix([[int(P[i][j]) for j in range(n)] for i in range(n)])
        Ak = Lk * Lk.T
        Al = Ll * Ll.T
        Ck = Ak - mp.mpc(0, 1) * Bk
        tCl = Pm.T * (Al + mp.mpc(0, 1) * Bl) * Pm
        C = Ck + tCl
        detLk = mp.fprod([Lk[i, i] for i in range(n)])
        detLl = mp.fprod([Ll[i, i] for i in range(n)])
        Ci = C ** -1
        z = abs(detLk * detLl) / mp.det(C)
        S = mp.power(2, mp.mpf(3 * n) / 2) * z * mp.sqrt(z)
        M = mp.matrix([[mp.mpf(system.mass_matrix[i][j]) for j in range(n)] for i in range(n)])
        X = M * tCl * Ci * Ck
        T = 6 * S * mp.fsum([X[i, i] for i in range(n)])
        acc = mp.mpc(0)
        for i in range(n):
            acc += mp.mpf(system.q0) * mp.mpf(system.q[i]) / mp.sqrt(Ci[i, i])
        for i in range(n):
            for j in range(i + 1, n):
                acc += mp.mpf(system.q[i]) * mp.mpf(system.q[j]) / mp.sqrt(Ci[i, i] + Ci[j, j] - 2 * Ci[i, j])
        V = 2 / mp.sqrt(mp.pi) * S * acc
        return T + V, S


def pair_grad_fd(system, pk, pl, P, side, dps=45, rel_step=mp.mpf("1e-14")):
    """Central differences of (H, S) with respect to the parameters of the
    bra (side='k') or the ket (side='l'), packed like the reference's Dk/Dl:
    (dH/dvechL, dH/dvechB, dS/dvechL, dS/dvechB)."""
    with mp.workdps(dps):
        base = [mp.mpf(v) for v in (pk if side == "k" else pl)]
        npt = len(base)
        dH, dS = [], []
        for q in range(npt):
            h = rel_step * max(mp.mpf(1), abs(base[q]))
            plus = list(base)
            minus = list(base)
            plus[q] += h
            minus[q] -= h
            if side == "k":
                Hp, Sp = pair_hs(system, plus, pl, P, dps)
                Hm, Sm = pair_hs(system, minus, pl, P, dps)
            else:
                Hp, Sp = pair_hs(system, pk, plus, P, dps)
                Hm, Sm = pair_hs(system, pk, minus, P, dps)
            dH.append((Hp - Hm) / (2 * h))
            dS.append((Sp - Sm) / (2 * h))
        return [complex(v) for v in dH] + [complex(v) for v in dS]


def n1_quadrature(m0, m1, q0, q1, pk, pl, dps=30):
    """Two-body (n = 1) matrix elements by radial quadrature of the actual
    integrals, normalised:  <k|l>, <k| -grad^2/(2 mu) |l>, <k| q0 q1 / r |l>
    with phi = exp(-(L^2 + i B) r^2)."""
    with mp.workdps(dps):
        ak = mp.mpf(pk[0]) ** 2 + mp.mpc(0, 1) * mp.mpf(pk[1])
        al = mp.mpf(pl[0]) ** 2 + mp.mpc(0, 1) * mp.mpf(pl[1])
        c = mp.conj(ak) + al

        def radial(f):
            re = mp.quad(lambda r: mp.re(f(r)), [0, mp.inf])
            im = mp.quad(lambda r: mp.im(f(r)), [0, mp.inf])
            return mp.mpc(re, im)

        ov = radial(lambda r: 4 * mp.pi * r ** 2 * mp.exp(-c * r ** 2))
        nk = mp.quad(lambda r: 4 * mp.pi * r ** 2 * mp.exp(-2 * mp.re(ak) * r ** 2), [0, mp.inf])
        nl = mp.quad(lambda r: 4 * mp.pi * r ** 2 * mp.exp(-2 * mp.re(al) * r ** 2), [0, mp.inf])
        inv_mu2 = 1 / (2 * mp.mpf(m0)) + 1 / (2 * mp.mpf(m1))
        # grad conj(phi_k) . grad phi_l = 4 conj(a_k) a_l r^2 conj(phi_k) phi_l
        kin = radial(lambda r: 4 * mp.pi * r ** 2 * 4 * mp.conj(ak) * al * r ** 2 * mp.exp(-c * r ** 2))
        pot = radial(lambda r: 4 * mp.pi * r * mp.exp(-c * r ** 2))
        norm = mp.sqrt(nk * nl)
        return complex(ov / norm), complex(inv_mu2 * kin / norm), complex(mp.mpf(q0) * mp.mpf(q1) * pot / norm)


def overlap_shift(system, pk, pl, P, a, b, s, dps=30):
    """S of the pair with C replaced by C + s J (J = w w^T, w = e_a for
    a == b, e_a - e_b otherwise): the unnormalised <k| exp(-s r_ab^2) |l>
    up to the pair's constant, at `dps` digits (mpmath determinant; the
    working precision grows with log10 s so that C stays resolved beside
    s J at the quadratures' far nodes)."""
    s = mp.mpf(s)
    extra = int(mp.log10(abs(s))) + 5 if abs(s) > 1 else 0
    with mp.workdps(dps + extra):
        n = system.n
        Lk, Bk = _unpack(pk, n)
        Ll, Bl = _unpack(pl, n)
        Pm = mp.matrix([[int(P[i][j]) for j in range(n)] for i in range(n)])
        C = (Lk * Lk.T - mp.mpc(0, 1) * Bk) + Pm.T * (Ll * Ll.T + mp.mpc(0, 1) * Bl) * Pm
        w = mp.zeros(n, 1)
        w[a] = 1
        if b != a:
            w[b] = -1
        Cs = C + s * (w * w.T)
        detLk = mp.fprod([Lk[i, i] for i in range(n)])
        detLl = mp.fprod([Ll[i, i] for i in range(n)])
        z = abs(detLk * detLl) / mp.det(Cs)
        return mp.power(2, mp.mpf(3 * n) / 2) * z * mp.sqrt(z)


def expc_by_integrals(system, pk, pl, P, a, b, dps=30):
    """<1/r^2>, <1/r>, <r>, <r^2>, <delta(r)> of r = |r_a - r_b| (r_a for
    a == b) between the pair's functions, from the shifted overlap f(s):
      1/r^2 = int_0^inf exp(-s r^2) ds,   1/r = (2/sqrt(pi)) int_0^inf exp(-t^2 r^2) dt,
      r = (1/sqrt(pi)) int_0^inf (1 - exp(-t^2 r^2)) / t^2 dt,
      r^2 = -d/ds exp(-s r^2) at 0,  delta(r) = lim (s/pi)^(3/2) exp(-s r^2)."""
    with mp.workdps(dps):
        f = lambda s: overlap_shift(system, pk, pl, P, a, b, s, dps)

        def cquad(g):
            return mp.mpc(mp.quad(lambda t: mp.re(g(t)), [0, 1, mp.inf]), mp.quad(lambda t: mp.im(g(t)), [0, 1, mp.inf]))

        f0 = f(0)
        rm2 = cquad(f)
        rm = 2 / mp.sqrt(mp.pi) * cquad(lambda t: f(t * t))
        r = 1 / mp.sqrt(mp.pi) * cquad(lambda t: (f0 - f(t * t)) / (t * t))
        h = mp.mpf(10) ** (-dps // 3)
        r2 = -(f(h) - f(-h)) / (2 * h)
        s1 = mp.mpf(10) ** (dps // 2)
        g = lambda s: (s / mp.pi) ** mp.mpf(1.5) * f(s)
        delta = 2 * g(2 * s1) - g(s1)
        return [complex(v) for v in (rm2, rm, r, r2, delta)]


def n1_moments(pk, pl, dps=30):
    """Two-body (n = 1) <k| r^-2, r^-1, r, r^2 |l> and <k| delta(r) |l> by
    radial quadrature, normalised by sqrt(<k|k><l|l>)."""
    with mp.workdps(dps):
        ak = mp.mpf(pk[0]) ** 2 + mp.mpc(0, 1) * mp.mpf(pk[1])
        al = mp.mpf(pl[0]) ** 2 + mp.mpc(0, 1) * mp.mpf(pl[1])
        c = mp.conj(ak) + al

        def radial(f):
            return mp.mpc(mp.quad(lambda r: mp.re(f(r)), [0, mp.inf]), mp.quad(lambda r: mp.im(f(r)), [0, mp.inf]))

        nk = mp.quad(lambda r: 4 * mp.pi * r ** 2 * mp.exp(-2 * mp.re(ak) * r ** 2), [0, mp.inf])
        nl = mp.quad(lambda r: 4 * mp.pi * r ** 2 * mp.exp(-2 * mp.re(al) * r ** 2), [0, mp.inf])
        norm = mp.sqrt(nk * nl)
        vals = [radial(lambda r, p=p: 4 * mp.pi * r ** (2 + p) * mp.exp(-c * r ** 2)) for p in (-2, -1, 1, 2)]
        vals.append(mp.mpc(1))   # conj(phi_k(0)) phi_l(0)
        return [complex(v / norm) for v in vals]


def mv_ratio_by_derivatives(system, pk, pl, P, masses, dps=30):
    """MV_kl / S_kl of one pair and symmetry term from Gaussian-moment
    derivatives: for particle i with direction matrix J_i (1 1^T for
    particle 1, e_i e_i^T for particle i + 1), Laplacians of the two
    Gaussians give <D phi_k | D phi_l> / <phi_k | phi_l> =
    16 f_st / f + 24 b f_s / f + 24 a f_t / f + 36 a b with
    f(s, t) = det(C + s P + t Q)^(-3/2), P = Ck J Ck, Q = tCl J tCl,
    a = tr(J Ck), b = tr(J tCl); then MV = -(1/8) sum_i (.) / m_i^3."""
    with mp.workdps(dps):
        n = system.n
        Lk, Bk = _unpack(pk, n)
        Ll, Bl = _unpack(pl, n)
        Pm = mp.matrix([[int(P[i][j]) for j in range(n)] for i in range(n)])
        Ck = Lk * Lk.T - mp.mpc(0, 1) * Bk
        tCl = Pm.T * (Ll * Ll.T + mp.mpc(0, 1) * Bl) * Pm
        C = Ck + tCl
        total = mp.mpc(0)
        for part in range(n + 1):
            if part == 0:
                J = mp.matrix([[1] * n for _ in range(n)])
            else:
                J = mp.zeros(n, n)
                J[part - 1, part - 1] = 1
            Pq = Ck * J * Ck
            Qq = tCl * J * tCl
            a = sum(J[i, j] * Ck[j, i] for i in range(n) for j in range(n))
            b = sum(J[i, j] * tCl[j, i] for i in range(n) for j in range(n))
            f = lambda s, t: mp.det(C + s * Pq + t * Qq) ** mp.mpf(-1.5)
            f0 = f(0, 0)
            fs = mp.diff(lambda s: f(s, 0), 0)
            ft = mp.diff(lambda t: f(0, t), 0)
            fst = mp.diff(f, (0, 0), (1, 1))
            ratio = 16 * fst / f0 + 24 * b * fs / f0 + 24 * a * ft / f0 + 36 * a * b
            total += ratio / mp.mpf(masses[part]) ** 3
        return complex(-total / 8)
