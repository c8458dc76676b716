// Per-pair ECG algebra for CDNA4 (gfx950), FP64.
//
// One pair (k, l) of complex ECGs phi = exp[-r^T (L L^T + i B) r] and one
// symmetry matrix P.  With Ck* = A_k - i B_k and tCl = P^T (A_l + i B_l) P,
// C = Ck* + tCl, the normalised integrals are (src/matelem.f90:244-299)
//   S = 2^(3n/2) z sqrt(z),           z = |det L_k det L_l| / det C
//   T = 6 S tr(M tCl C^-1 Ck*)
//   V = (2/sqrt(pi)) S [q0 sum_i q_i (C^-1_ii)^-1/2
//                       + sum_{i<j} q_i q_j (C^-1_ii + C^-1_jj - 2 C^-1_ij)^-1/2]
// C is factorised as C = L D L^T (unpivoted, same elimination order as the
// reference's pseudo-Cholesky but without square roots); det C = prod D.
//
// Gradients use the matrix adjoints of h = T + V - e S (e = 0 gives dH):
//   dh = (3/2) h (dln|det L_k| + dln|det L_l|) + tr(Gb dCk*) + tr(Gk dtCl)
//   Gb = -(3/2) h C^-1 + S (6 F + Q),  F = Y M Y^T, Y = C^-1 tCl
//   Gk = -(3/2) h C^-1 + S (6 G + Q),  G = Z M Z^T, Z = C^-1 Ck*
//   Q  = (1/sqrt(pi)) sum_w c_w (w^T C^-1 w)^-3/2 (C^-1 w)(C^-1 w)^T
// with w = e_i (c_w = q0 q_i) and e_i - e_j (c_w = q_i q_j).  The chain rule
// to vech L and vech B is applied once per pair (stored mode) or once per
// function after the sums over partners (adjoint mode).
#pragma once
#include <hip/hip_runtime.h>

namespace ecg {

// ----------------------------------------------------------- complex FP64
struct zd {
  double x, y;
};
__host__ __device__ inline zd mkz(double x, double y) { return zd{x, y}; }
__host__ __device__ inline zd operator+(zd a, zd b) { return zd{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline zd operator-(zd a, zd b) { return zd{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline zd operator-(zd a) { return zd{-a.x, -a.y}; }
__host__ __device__ inline zd operator*(zd a, zd b) {
  return zd{fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x)};
}
__host__ __device__ inline zd operator*(double s, zd a) { return zd{s * a.x, s * a.y}; }
__host__ __device__ inline zd operator*(zd a, double s) { return zd{s * a.x, s * a.y}; }
// a + b*c
__host__ __device__ inline zd zfma(zd b, zd c, zd a) {
  return zd{fma(b.x, c.x, fma(-b.y, c.y, a.x)), fma(b.x, c.y, fma(b.y, c.x, a.y))};
}
// a - b*c
__host__ __device__ inline zd zfms(zd b, zd c, zd a) {
  return zd{fma(-b.x, c.x, fma(b.y, c.y, a.x)), fma(-b.x, c.y, fma(-b.y, c.x, a.y))};
}
__host__ __device__ inline zd zconj(zd a) { return zd{a.x, -a.y}; }
__host__ __device__ inline zd zrcp(zd a) {
  const double d = 1.0 / fma(a.x, a.x, a.y * a.y);
  return zd{a.x * d, -a.y * d};
}
// Principal square root: branch cut on the negative real axis, the sign of
// the imaginary part (including a signed zero) picks the side.
__host__ __device__ inline zd zsqrt(zd z) {
  if (z.x == 0.0 && z.y == 0.0) return zd{0.0, z.y};
  const double r = sqrt(fma(z.x, z.x, z.y * z.y));
  const double t = sqrt(0.5 * (fabs(z.x) + r));
  const double h = 0.5 / t;
  if (z.x >= 0.0) return zd{t, z.y * h};
  return zd{fabs(z.y) * h, copysign(t, z.y)};
}

// Reciprocal and square root for the pair kernels, which are bound by
// their VALU issue rate: an IEEE-correct division or square root costs
// 10-12 instructions, these 5-9.  For normal arguments whose results are
// normal (the pair algebra's range) they are within an ulp or so of the
// correctly rounded values.
// 1/x: hardware reciprocal, two Newton steps r += r (1 - x r).
__device__ inline double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
}
// sqrt(x) and 1/sqrt(x), x > 0: hardware reciprocal square root y, then
// two coupled Goldschmidt steps on g = x y -> sqrt(x), h = y / 2 ->
// 1 / (2 sqrt(x)), and a last residual correction of g.
__device__ inline void fsqrt2(double x, double& sq, double& rsq) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  sq = fma(fma(-g, g, x), h, g);
  rsq = 2.0 * h;
}
__device__ inline zd zrcp_f(zd a) {
  const double d = frcp(fma(a.x, a.x, a.y * a.y));
  return zd{a.x * d, -a.y * d};
}
// Principal square root as zsqrt (same branch cut and signed zeros); with
// rinv, also 1 / sqrt(z) = conj(sqrt(z)) / |z|.
__device__ inline zd zsqrt_f(zd z, zd* rinv = nullptr) {
  if (z.x == 0.0 && z.y == 0.0) {
    if (rinv) *rinv = zd{__builtin_inf(), 0.0};
    return zd{0.0, z.y};
  }
  double r, ir;
  fsqrt2(fma(z.x, z.x, z.y * z.y), r, ir);
  double t, it;
  fsqrt2(0.5 * (fabs(z.x) + r), t, it);
  const double h = 0.5 * it;
  const zd w = (z.x >= 0.0) ? zd{t, z.y * h} : zd{fabs(z.y) * h, copysign(t, z.y)};
  if (rinv) *rinv = zd{w.x * ir, -w.y * ir};
  return w;
}
__device__ inline zd zrsqrt_f(zd z) {
  zd ri;
  (void)zsqrt_f(z, &ri);
  return ri;
}

// ------------------------------------------------------ packed symmetric
// Lower triangle packed by rows: (0,0) (1,0) (1,1) (2,0) ...
__host__ __device__ constexpr int sidx(int i, int j) {
  return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}
// vech position of (i, j), i >= j: column-major lower triangle, the
// reference's parameter order (src/matelem.f90:109-122).
__host__ __device__ constexpr int vidx(int n, int i, int j) {
  return j * n - j * (j - 1) / 2 + (i - j);
}

// Constant inputs shared by all pairs of one system.
struct SysConst {
  int n, T;
  double M[5][5];   // mass matrix
  double q0, q[5];  // charges
  double two3n2;    // 2^(3n/2)
  double coef[16];  // Y^dagger Y coefficients
  signed char P[16][5][5];  // Y^dagger Y matrices, P[t][row][col]
};

// Per-pair results of one term.
template <int N> struct PairCore {
  static constexpr int m = N * (N + 1) / 2;
  zd Ci[m];   // C^-1, packed symmetric
  zd S, T, V;
  zd rs[m];   // (w^T C^-1 w)^-1/2 for w = e_i (packed diag) and e_i - e_j
};

// Factor C (packed symmetric), return C^-1 and the reciprocal of det C.
template <int N>
__device__ inline void sym_inverse(const zd* C, zd* Ci, zd& rdet) {
  constexpr int m = N * (N + 1) / 2;
  zd l[m];   // unit lower factor (strict part at sidx(i,j), i>j)
  zd r[N];   // 1 / D_i
  zd u[N];   // u_j = l_ij D_j for the current row
#pragma unroll
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int j = 0; j < i; ++j) {
      zd t = C[sidx(i, j)];
#pragma unroll
      for (int p = 0; p < j; ++p) t = zfms(u[p], l[sidx(j, p)], t);
      u[j] = t;
      l[sidx(i, j)] = t * r[j];
    }
    zd d = C[sidx(i, i)];
#pragma unroll
    for (int p = 0; p < i; ++p) d = zfms(u[p], l[sidx(i, p)], d);
    r[i] = zrcp_f(d);
  }
  rdet = r[0];
#pragma unroll
  for (int i = 1; i < N; ++i) rdet = rdet * r[i];
  // X = L^-1 (unit lower), stored in l's slots.
#pragma unroll
  for (int j = 0; j < N; ++j) {
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      zd t = -l[sidx(i, j)];
#pragma unroll
      for (int p = j + 1; p < i; ++p) t = zfms(l[sidx(i, p)], l[sidx(p, j)], t);
      l[sidx(i, j)] = t;   // columns < j of row i are still L, columns > j already X
    }
  }
  // C^-1_ab = sum_{p >= max(a,b)} X_pa r_p X_pb
#pragma unroll
  for (int a = 0; a < N; ++a) {
#pragma unroll
    for (int b = 0; b <= a; ++b) {
      zd s = (a == b) ? r[a] : r[a] * l[sidx(a, b)];
#pragma unroll
      for (int p = a + 1; p < N; ++p) {
        zd xa = l[sidx(p, a)], xb = (p == b) ? mkz(1.0, 0.0) : l[sidx(p, b)];
        s = zfma(r[p] * xa, xb, s);
      }
      Ci[sidx(a, b)] = s;
    }
  }
}

// Energy part of one (pair, term).  Ckc: Ck* packed; Ctl: tCl packed;
// Nl: M tCl full (row-major N x N); adet = |det L_k det L_l|.
template <int N>
__device__ inline void pair_energy(const SysConst& sc, const zd* Ckc, const zd* Ctl,
                                            const zd* Nl, double adet, PairCore<N>& pc) {
  constexpr int m = N * (N + 1) / 2;
  zd C[m];
#pragma unroll
  for (int e = 0; e < m; ++e) C[e] = Ckc[e] + Ctl[e];
  zd rdet;
  sym_inverse<N>(C, pc.Ci, rdet);
  const zd z = adet * rdet;
  pc.S = (sc.two3n2 * z) * zsqrt_f(z);
  // tr(M tCl C^-1 Ck*) = sum_ab (M tCl)_ab (C^-1 Ck*)_ba
  zd tr = mkz(0.0, 0.0);
#pragma unroll
  for (int b = 0; b < N; ++b)
#pragma unroll
    for (int a = 0; a < N; ++a) {
      zd zba = mkz(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < N; ++c) zba = zfma(pc.Ci[sidx(b, c)], Ckc[sidx(c, a)], zba);
      tr = zfma(Nl[a * N + b], zba, tr);
    }
  pc.T = (6.0 * pc.S) * tr;
  // Coulomb
  zd acc1 = mkz(0.0, 0.0), acc2 = mkz(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const zd rsi = zrsqrt_f(pc.Ci[sidx(i, i)]);
    pc.rs[sidx(i, i)] = rsi;
    acc1 = acc1 + sc.q[i] * rsi;
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = i + 1; j < N; ++j) {
      zd c = pc.Ci[sidx(i, i)] + pc.Ci[sidx(j, j)] - 2.0 * pc.Ci[sidx(j, i)];
      const zd rsij = zrsqrt_f(c);
      pc.rs[sidx(j, i)] = rsij;
      acc2 = acc2 + (sc.q[i] * sc.q[j]) * rsij;
    }
  const double two_over_sqrtpi = 1.12837916709551257389615890312154517;
  pc.V = (two_over_sqrtpi * pc.S) * (sc.q0 * acc1 + acc2);
}

// Matrix adjoints of one (pair, term) for h = H - e S, returned as packed
// symmetric Gb (bra, coefficient of dCk*) and Gk (ket, coefficient of dtCl,
// still in the permuted frame).  Requires pair_energy's results.
template <int N>
__host__ __device__ inline void pair_adjoint(const SysConst& sc, const zd* Ckc, const zd* Ctl,
                                             const PairCore<N>& pc, double e, zd* Gb, zd* Gk) {
  constexpr int m = N * (N + 1) / 2;
  const zd h = pc.T + pc.V - e * pc.S;
  const zd c15h = -1.5 * h;
  // Y = C^-1 tCl, Z = C^-1 Ck* (full)
  zd Y[N][N], Z[N][N];
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = 0; b < N; ++b) {
      zd y = mkz(0.0, 0.0), w = mkz(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < N; ++c) {
        y = zfma(pc.Ci[sidx(a, c)], Ctl[sidx(c, b)], y);
        w = zfma(pc.Ci[sidx(a, c)], Ckc[sidx(c, b)], w);
      }
      Y[a][b] = y;
      Z[a][b] = w;
    }
  // Q (packed): (1/sqrt(pi)) sum_w c_w rs_w^3 u_w u_w^T, u_w = C^-1 w
  const double rsqrtpi = 0.564189583547756286948079451560772586;
  zd Q[m];
#pragma unroll
  for (int e2 = 0; e2 < m; ++e2) Q[e2] = mkz(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      zd rs = pc.rs[sidx(i, j)];
      zd cw = ((i == j) ? (sc.q0 * sc.q[i]) : (sc.q[i] * sc.q[j])) * (rs * rs * rs);
      zd u[N];
#pragma unroll
      for (int a = 0; a < N; ++a) u[a] = (i == j) ? pc.Ci[sidx(a, i)] : pc.Ci[sidx(a, i)] - pc.Ci[sidx(a, j)];
#pragma unroll
      for (int a = 0; a < N; ++a) {
        zd cu = cw * u[a];
#pragma unroll
        for (int b = 0; b <= a; ++b) Q[sidx(a, b)] = zfma(cu, u[b], Q[sidx(a, b)]);
      }
    }
  }
  // F = Y M Y^T and G = Z M Z^T (packed)
  zd YM[N][N], ZM[N][N];
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = 0; b < N; ++b) {
      zd y = mkz(0.0, 0.0), w = mkz(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < N; ++c) {
        y = y + sc.M[c][b] * Y[a][c];
        w = w + sc.M[c][b] * Z[a][c];
      }
      YM[a][b] = y;
      ZM[a][b] = w;
    }
  const zd S = pc.S;
  const zd S6 = 6.0 * S;
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = 0; b <= a; ++b) {
      zd f = mkz(0.0, 0.0), g = mkz(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < N; ++c) {
        f = zfma(YM[a][c], Y[b][c], f);
        g = zfma(ZM[a][c], Z[b][c], g);
      }
      const int s = sidx(a, b);
      const zd base = zfma(c15h, pc.Ci[s], rsqrtpi * (S * Q[s]));
      Gb[s] = zfma(S6, f, base);
      Gk[s] = zfma(S6, g, base);
    }
}

// P G P^T for an integer matrix P (packed symmetric in and out).
template <int N>
__host__ __device__ inline void perm_congruence(const signed char (*P)[5], const zd* G, zd* out) {
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = 0; b <= a; ++b) {
      zd s = mkz(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < N; ++c) {
        if (P[a][c] == 0) continue;
#pragma unroll
        for (int d = 0; d < N; ++d) {
          const int pp = P[a][c] * P[b][d];
          if (pp == 0) continue;
          s = s + double(pp) * G[sidx(c, d)];
        }
      }
      out[sidx(a, b)] = s;
    }
}

// Chain rule from a symmetric adjoint G (coefficient of dA +/- i dB) and
// the det-term scalar alpha to d/dvech L and d/dvech B of one function:
//   d/dL(i,j) = 2 (G L)(i,j) + alpha delta_ij / L_jj   (i >= j)
//   d/dB(i,i) = sgn i G_ii,  d/dB(i,j) = 2 sgn i G_ij     (sgn = -1 bra, +1 ket)
template <int N>
__host__ __device__ inline void chain_to_params(const zd* G, zd alpha, const double* L /*packed lower*/,
                                                double sgn, zd* dL, zd* dB) {
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int i = j; i < N; ++i) {
      zd s = mkz(0.0, 0.0);
#pragma unroll
      for (int c = j; c < N; ++c) s = s + L[sidx(c, j)] * G[sidx(i, c)];
      s = 2.0 * s;
      if (i == j) s = s + alpha * (1.0 / L[sidx(j, j)]);
      dL[vidx(N, i, j)] = s;
      const zd g = G[sidx(i, j)];
      const double f = (i == j) ? sgn : 2.0 * sgn;
      dB[vidx(N, i, j)] = mkz(-f * g.y, f * g.x);
    }
}

}  // namespace ecg
