// CPU restatement of the ECG pair integrals of ATOM-MOL-nonBO.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle: the product path never
// links, loads or calls it.  Only tests/, the smoke check and the CPU
// baseline leg of the benchmark use it, and only as the thing results are
// checked against (or timed beside).
//
// The arithmetic follows the reference step by step (same factorisation,
// same formula shapes, same order of accumulation) so that it reproduces
// the reference's rounding behaviour closely.  It is templated on the real
// type so that an operation-counting type can be substituted to count the
// floating-point work of the reference algorithm.
#pragma once
#include <cmath>
#include <cstdint>

namespace ecgo {

constexpr int NN = 5;                 // max pseudo-particles (src/globvars.f90:43)
constexpr int NNP = NN * (NN + 1) / 2;

// ---------------------------------------------------------------- scalars
template <class R> struct cx {
  R re, im;
  cx() : re(0), im(0) {}
  cx(R r) : re(r), im(0) {}
  cx(R r, R i) : re(r), im(i) {}
};
template <class R> inline cx<R> operator+(const cx<R>& a, const cx<R>& b) { return {a.re + b.re, a.im + b.im}; }
template <class R> inline cx<R> operator-(const cx<R>& a, const cx<R>& b) { return {a.re - b.re, a.im - b.im}; }
template <class R> inline cx<R> operator-(const cx<R>& a) { return {-a.re, -a.im}; }
template <class R> inline cx<R> operator*(const cx<R>& a, const cx<R>& b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R> inline cx<R> operator*(R s, const cx<R>& a) { return {s * a.re, s * a.im}; }
template <class R> inline cx<R> operator*(const cx<R>& a, R s) { return {s * a.re, s * a.im}; }
// Complex division a/b, scaled form (Smith's algorithm) to avoid overflow.
template <class R> inline cx<R> operator/(const cx<R>& a, const cx<R>& b) {
  using std::fabs;
  if (fabs(b.re) >= fabs(b.im)) {
    R r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  } else {
    R r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
}
template <class R> inline cx<R> operator/(const cx<R>& a, R s) { return {a.re / s, a.im / s}; }
template <class R> inline cx<R> conj(const cx<R>& a) { return {a.re, -a.im}; }
template <class R> inline R cabs(const cx<R>& a) { using std::hypot; return hypot(a.re, a.im); }
// Principal square root (branch cut on the negative real axis, sign of the
// imaginary part decides the side, as C csqrt and Fortran SQRT do).
template <class R> inline cx<R> csqrt(const cx<R>& z) {
  using std::fabs; using std::sqrt; using std::copysign;
  if (z.re == R(0) && z.im == R(0)) return {R(0), z.im};
  R t = sqrt((fabs(z.re) + cabs(z)) * R(0.5));
  if (z.re >= R(0)) return {t, z.im / (R(2) * t)};
  return {fabs(z.im) / (R(2) * t), copysign(t, z.im)};
}

// Work that the flop convention does not count (per-function set-up and
// integer permutation products) is wrapped in a Pause guard; it is a no-op
// for ordinary floating types and switches counting off for the counting
// type.
template <class R> struct Pause { Pause() {} };

// ------------------------------------------------------- constant inputs
template <class R> struct System {
  int n;               // pseudo-particles
  int T;               // Y^dagger Y terms
  R M[NN][NN];         // mass matrix (src/workproc.f90:606-610)
  R q0;                // charge of the reference particle
  R q[NN];             // pseudo-charges
  R two3n2;            // 2^(3n/2)
  const int* P;        // n*n*T integer matrices, column-major per term
  const int* coef;     // T integer coefficients
};

// ------------------------------------------------------------- one pair
// Quantities of one pair that MatrixElementsForExpcVals reuses
// (src/matelem.f90:1180-1235 computes them the same way).
template <class R> struct PairParts {
  cx<R> Ci[NN][NN];   // C^{-1}
  cx<R> T, V;
  cx<R> Ck[NN][NN], tCl[NN][NN];   // the bra's Ck^* and the permuted ket's tCl
};

// Restatement of MatrixElements (src/matelem.f90:12-750) for one symmetry
// matrix P (column-major n x n).  Dk, Dl receive 4*np complex values each
// (dH/dvechL, dH/dvechB, dS/dvechL, dS/dvechB) when gradflag asks for them;
// parts (optional) receives C^{-1}, T and V.
template <class R>
void matrix_elements(const System<R>& sy, const R* pk, const R* pl, const int* Pm, int gradflag,
                     cx<R>& Hkl, cx<R>& Skl, cx<R>* Dk, cx<R>* Dl, PairParts<R>* parts = nullptr) {
  const int n = sy.n, np = n * (n + 1) / 2;
  auto P = [&](int i, int j) { return R(Pm[i + n * j]); };
  const bool gk = (gradflag == 1 || gradflag == 3);
  const bool gl = (gradflag == 2 || gradflag == 3);

  // Unpack vech L, vech B (src/matelem.f90:109-122).
  R Lk[NN][NN] = {}, Ll[NN][NN] = {}, Bk[NN][NN], tBl[NN][NN];
  R Ak[NN][NN], tAl[NN][NN];
  R detLk = 1, detLl = 1;
  int idx = 0;
  {
  Pause<R> per_function_setup;
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j, ++idx) {
      Lk[j][i] = pk[idx];
      Ll[j][i] = pl[idx];
      Bk[j][i] = Bk[i][j] = pk[np + idx];
      tBl[j][i] = tBl[i][j] = pl[np + idx];
    }
  // A = L L^T (src/matelem.f90:124-139).
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      R s = 0;
      for (int k = 0; k <= i; ++k) s = s + Lk[i][k] * Lk[j][k];
      Ak[i][j] = Ak[j][i] = s;
      s = 0;
      for (int k = 0; k <= i; ++k) s = s + Ll[i][k] * Ll[j][k];
      tAl[i][j] = tAl[j][i] = s;
    }
  // Ket permutation: tAl = P^T Al P, tBl = P^T Bl P (src/matelem.f90:144-169).
  {
    R W1[NN][NN], W2[NN][NN];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        R a = 0, b = 0;
        for (int k = 0; k < n; ++k) { a = a + P(k, i) * tAl[k][j]; b = b + P(k, i) * tBl[k][j]; }
        W1[i][j] = a; W2[i][j] = b;
      }
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j) {
        R a = 0, b = 0;
        for (int k = 0; k < n; ++k) { a = a + W1[i][k] * P(k, j); b = b + W2[i][k] * P(k, j); }
        tAl[i][j] = tAl[j][i] = a;
        tBl[i][j] = tBl[j][i] = b;
      }
  }
  // det L = product of diagonals (src/matelem.f90:173-178).
  for (int i = 0; i < n; ++i) { detLk = detLk * Lk[i][i]; detLl = detLl * Ll[i][i]; }
  }
  // Ck (holds Ck^*), tCl, tCkl = Ck^* + tCl (src/matelem.f90:182-194).
  cx<R> Ck[NN][NN], tCl[NN][NN], G[NN][NN];
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      Ck[j][i] = Ck[i][j] = cx<R>(Ak[j][i], -Bk[j][i]);
      tCl[j][i] = tCl[i][j] = cx<R>(tAl[j][i], tBl[j][i]);
      G[j][i] = G[i][j] = Ck[j][i] + tCl[j][i];
    }
  // Unpivoted complex-symmetric pseudo-Cholesky C = G G^T, det C = prod of
  // squared pivots (src/matelem.f90:203-217).
  cx<R> detC(R(1), R(0));
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      cx<R> s = G[i][j];
      for (int k = i - 1; k >= 0; --k) s = s - G[i][k] * G[j][k];
      if (i == j) { G[i][i] = csqrt(s); detC = detC * s; }
      else G[j][i] = s / G[i][i];
    }
  // In-place inverse of the lower factor (src/matelem.f90:221-231).
  for (int i = 0; i < n; ++i) {
    G[i][i] = cx<R>(R(1)) / G[i][i];
    for (int j = i + 1; j < n; ++j) {
      cx<R> s;
      for (int k = i; k < j; ++k) s = s - G[j][k] * G[k][i];
      G[j][i] = s / G[j][j];
    }
  }
  // C^{-1} = G^{-T} G^{-1} (src/matelem.f90:233-241).
  cx<R> Ci[NN][NN];
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      cx<R> s;
      for (int k = j; k < n; ++k) s = s + G[k][i] * G[k][j];
      Ci[i][j] = Ci[j][i] = s;
    }
  // Overlap S = 2^(3n/2) z sqrt(z), z = |detLk detLl| / det C
  // (src/matelem.f90:244-245).
  {
    using std::fabs;
    cx<R> z = cx<R>(fabs(detLl * detLk)) / detC;
    Skl = (sy.two3n2 * z) * csqrt(z);
  }
  // Kinetic T = 6 S tr(M tCl C^{-1} Ck^*) (src/matelem.f90:248-274).
  cx<R> W1[NN][NN], W2[NN][NN];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      cx<R> s;
      for (int k = 0; k < n; ++k) s = s + Ci[i][k] * Ck[k][j];
      W1[i][j] = s;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      cx<R> s;
      for (int k = 0; k < n; ++k) s = s + tCl[i][k] * W1[k][j];
      W2[i][j] = s;
    }
  cx<R> tr;
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) tr = tr + sy.M[i][k] * W2[k][i];
  const cx<R> Tkl = R(6) * Skl * tr;
  // Coulomb V (src/matelem.f90:280-299); J32 keeps tr(C^{-1}J_ij)^{-3/2}.
  const R sqrtpi = R(1.77245385090551602729816748334115);
  cx<R> J32[NN][NN];
  cx<R> Vkl;
  {
    cx<R> pref = (R(2) / sqrtpi) * Skl;
    cx<R> acc;
    for (int i = 0; i < n; ++i) {
      cx<R> c = Ci[i][i];
      cx<R> s = csqrt(c);
      J32[i][i] = cx<R>(R(1)) / (s * c);
      acc = acc + sy.q[i] * (pref / s);
    }
    acc = acc * sy.q0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        cx<R> c = Ci[i][i] + Ci[j][j] - Ci[j][i] - Ci[j][i];
        cx<R> s = csqrt(c);
        J32[j][i] = cx<R>(R(1)) / (s * c);
        acc = acc + (sy.q[i] * sy.q[j]) * (pref / s);
      }
    Vkl = acc;
  }
  Hkl = Tkl + Vkl;
  if (parts) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        parts->Ci[i][j] = Ci[i][j];
        parts->Ck[i][j] = Ck[i][j];
        parts->tCl[i][j] = tCl[i][j];
      }
    parts->T = Tkl;
    parts->V = Vkl;
  }
  if (!gk && !gl) return;

  cx<R> dHLk[NNP], dHBk[NNP], dSLk[NNP], dSBk[NNP];
  cx<R> dHLl[NNP], dHBl[NNP], dSLl[NNP], dSBl[NNP];
  cx<R> ttCi[NN][NN];

  // dS wrt bra parameters (src/matelem.f90:305-357).
  if (gk) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cx<R> s;
        for (int k = j; k < n; ++k) s = s + Ci[i][k] * Lk[k][j];
        W1[i][j] = s;
      }
    R iL[NN][NN] = {};
    {
    Pause<R> per_function_inverse;
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j) iL[j][i] = Lk[j][i];
    for (int i = 0; i < n; ++i) {
      iL[i][i] = R(1) / iL[i][i];
      for (int j = i + 1; j < n; ++j) {
        iL[i][j] = 0;
        R s = 0;
        for (int k = i; k < j; ++k) s = s - iL[j][k] * iL[k][i];
        iL[j][i] = s / iL[j][j];
      }
    }
    }
    cx<R> c = Skl * (R(3) / R(2));
    idx = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j, ++idx) dSLk[idx] = (cx<R>(iL[i][j]) - R(2) * W1[j][i]) * c;
    c = Skl * cx<R>(R(0), R(3) / R(2));
    cx<R> c2 = R(2) * c;
    idx = 0;
    for (int i = 0; i < n; ++i) {
      dSBk[idx++] = Ci[i][i] * c;
      for (int j = i + 1; j < n; ++j) dSBk[idx++] = Ci[j][i] * c2;
    }
  }
  // dS wrt ket parameters, through P C^{-1} P^T (src/matelem.f90:359-431).
  if (gl) {
    {
    Pause<R> permutation_product;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + P(i, k) * Ci[k][j];
        W1[i][j] = s;
      }
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + W1[i][k] * P(j, k);
        ttCi[i][j] = ttCi[j][i] = s;
      }
    }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + ttCi[i][k] * Ll[k][j];
        W1[i][j] = s;
      }
    R iL[NN][NN] = {};
    {
    Pause<R> per_function_inverse;
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j) iL[j][i] = Ll[j][i];
    for (int i = 0; i < n; ++i) {
      iL[i][i] = R(1) / iL[i][i];
      for (int j = i + 1; j < n; ++j) {
        iL[i][j] = 0;
        R s = 0;
        for (int k = i; k < j; ++k) s = s - iL[j][k] * iL[k][i];
        iL[j][i] = s / iL[j][j];
      }
    }
    }
    cx<R> c = Skl * (R(3) / R(2));
    idx = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j, ++idx) dSLl[idx] = (cx<R>(iL[i][j]) - R(2) * W1[j][i]) * c;
    c = Skl * cx<R>(R(0), -R(3) / R(2));
    cx<R> c2 = R(2) * c;
    idx = 0;
    for (int i = 0; i < n; ++i) {
      dSBl[idx++] = ttCi[i][i] * c;
      for (int j = i + 1; j < n; ++j) dSBl[idx++] = ttCi[j][i] * c2;
    }
  }
  const cx<R> ToS = Tkl / Skl;
  // dT wrt ket: G = P C^{-1} Ck M Ck C^{-1} P^T (src/matelem.f90:435-519).
  if (gl) {
    cx<R> Gm[NN][NN];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + Ck[i][k] * sy.M[k][j];
        W1[i][j] = s;
      }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + W1[i][k] * Ck[k][j];
        W2[i][j] = s;
      }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + Ci[i][k] * W2[k][j];
        W1[i][j] = s;
      }
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + W1[i][k] * Ci[k][j];
        W2[i][j] = W2[j][i] = s;
      }
    {
    Pause<R> permutation_product;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + P(i, k) * W2[k][j];
        W1[i][j] = s;
      }
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + W1[i][k] * P(j, k);
        Gm[i][j] = Gm[j][i] = s;
      }
    }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cx<R> s;
        for (int k = j; k < n; ++k) s = s + Gm[i][k] * Ll[k][j];
        W1[i][j] = s;
      }
    cx<R> c12 = R(12) * Skl;
    idx = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j, ++idx) dHLl[idx] = ToS * dSLl[idx] + c12 * W1[j][i];
    cx<R> c6 = Skl * cx<R>(R(0), R(6));
    idx = 0;
    for (int i = 0; i < n; ++i) {
      dHBl[idx] = ToS * dSBl[idx] + c6 * Gm[i][i];
      ++idx;
      for (int j = i + 1; j < n; ++j, ++idx) dHBl[idx] = ToS * dSBl[idx] + R(2) * c6 * Gm[j][i];
    }
  }
  // dT wrt bra: F = C^{-1} tCl M tCl C^{-1} (src/matelem.f90:521-598).
  if (gk) {
    cx<R> F[NN][NN];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + tCl[i][k] * sy.M[k][j];
        W1[i][j] = s;
      }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + W1[i][k] * tCl[k][j];
        W2[i][j] = s;
      }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + Ci[i][k] * W2[k][j];
        W1[i][j] = s;
      }
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j) {
        cx<R> s;
        for (int k = 0; k < n; ++k) s = s + W1[i][k] * Ci[k][j];
        F[i][j] = F[j][i] = s;
      }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cx<R> s;
        for (int k = j; k < n; ++k) s = s + F[i][k] * Lk[k][j];
        W1[i][j] = s;
      }
    cx<R> c12 = R(12) * Skl;
    idx = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j, ++idx) dHLk[idx] = ToS * dSLk[idx] + c12 * W1[j][i];
    cx<R> c6 = Skl * cx<R>(R(0), -R(6));
    idx = 0;
    for (int i = 0; i < n; ++i) {
      dHBk[idx] = ToS * dSBk[idx] + c6 * F[i][i];
      ++idx;
      for (int j = i + 1; j < n; ++j, ++idx) dHBk[idx] = ToS * dSBk[idx] + R(2) * c6 * F[j][i];
    }
  }
  // dV (src/matelem.f90:603-734): Q_ij = C^{-1} J_ij C^{-1} weighted by
  // tr(C^{-1} J_ij)^{-3/2}.
  {
    cx<R> WLk[NNP], WBk[NNP], WLl[NNP], WBl[NNP];  // value-initialised to zero
    cx<R> tQ[NN][NN], ttQ[NN][NN];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j <= i; ++j) {
        if (i == j) {
          for (int k = 0; k < n; ++k) {
            tQ[k][k] = Ci[k][i] * Ci[i][k];
            for (int kk = k + 1; kk < n; ++kk) tQ[k][kk] = tQ[kk][k] = Ci[k][i] * Ci[i][kk];
          }
        } else {
          for (int k = 0; k < n; ++k) {
            tQ[k][k] = Ci[k][i] * Ci[i][k] + Ci[k][j] * Ci[j][k] - Ci[k][i] * Ci[j][k] - Ci[k][j] * Ci[i][k];
            for (int kk = k + 1; kk < n; ++kk)
              tQ[k][kk] = tQ[kk][k] =
                  Ci[k][i] * Ci[i][kk] + Ci[k][j] * Ci[j][kk] - Ci[k][i] * Ci[j][kk] - Ci[k][j] * Ci[i][kk];
          }
        }
        if (gk) {
          for (int k = 0; k < n; ++k)
            for (int kk = 0; kk <= k; ++kk) {
              cx<R> s;
              for (int m = kk; m < n; ++m) s = s + tQ[k][m] * Lk[m][kk];
              W1[k][kk] = s;
            }
        }
        if (gl) {
          {
          Pause<R> permutation_product;
          for (int k = 0; k < n; ++k)
            for (int kk = 0; kk < n; ++kk) {
              cx<R> s;
              for (int m = 0; m < n; ++m) s = s + P(k, m) * tQ[m][kk];
              W2[k][kk] = s;
            }
          for (int k = 0; k < n; ++k)
            for (int kk = k; kk < n; ++kk) {
              cx<R> s;
              for (int m = 0; m < n; ++m) s = s + W2[k][m] * P(kk, m);
              ttQ[k][kk] = ttQ[kk][k] = s;
            }
          }
          for (int k = 0; k < n; ++k)
            for (int kk = 0; kk <= k; ++kk) {
              cx<R> s;
              for (int m = kk; m < n; ++m) s = s + ttQ[k][m] * Ll[m][kk];
              W2[k][kk] = s;
            }
        }
        cx<R> w = (i == j) ? (sy.q0 * sy.q[i]) * J32[i][i] : (sy.q[i] * sy.q[j]) * J32[i][j];
        if (gk) {
          int ix = 0;
          for (int k = 0; k < n; ++k) {
            WBk[ix] = WBk[ix] + w * tQ[k][k];
            WLk[ix] = WLk[ix] + w * W1[k][k];
            ++ix;
            for (int kk = k + 1; kk < n; ++kk, ++ix) {
                WBk[ix] = WBk[ix] + R(2) * w * tQ[kk][k];
              WLk[ix] = WLk[ix] + w * W1[kk][k];
            }
          }
        }
        if (gl) {
          int ix = 0;
          for (int k = 0; k < n; ++k) {
            WBl[ix] = WBl[ix] + w * ttQ[k][k];
            WLl[ix] = WLl[ix] + w * W2[k][k];
            ++ix;
            for (int kk = k + 1; kk < n; ++kk, ++ix) {
                WBl[ix] = WBl[ix] + R(2) * w * ttQ[kk][k];
              WLl[ix] = WLl[ix] + w * W2[kk][k];
            }
          }
        }
      }
    const cx<R> c2 = (R(2) / sqrtpi) * Skl;
    const cx<R> ci = cx<R>(R(0), R(1) / sqrtpi) * Skl;
    const cx<R> VoS = Vkl / Skl;
    if (gk)
      for (int m = 0; m < np; ++m) {
        dHLk[m] = dHLk[m] + VoS * dSLk[m] + c2 * WLk[m];
        dHBk[m] = dHBk[m] + VoS * dSBk[m] - ci * WBk[m];
      }
    if (gl)
      for (int m = 0; m < np; ++m) {
        dHLl[m] = dHLl[m] + VoS * dSLl[m] + c2 * WLl[m];
        dHBl[m] = dHBl[m] + VoS * dSBl[m] + ci * WBl[m];
      }
  }
  // Packing (src/matelem.f90:737-748).
  if (gk)
    for (int m = 0; m < np; ++m) {
      Dk[m] = dHLk[m]; Dk[np + m] = dHBk[m]; Dk[2 * np + m] = dSLk[m]; Dk[3 * np + m] = dSBk[m];
    }
  if (gl)
    for (int m = 0; m < np; ++m) {
      Dl[m] = dHLl[m]; Dl[np + m] = dHBl[m]; Dl[2 * np + m] = dSLl[m]; Dl[3 * np + m] = dSBl[m];
    }
}

// Sum over the Y^dagger Y terms of one pair with integer coefficients, in
// term order, as ComputeMatElem(AndDeriv) does for a single rank
// (src/matform.f90:164-176, :294-303).
template <class R>
void pair_sum(const System<R>& sy, const R* pk, const R* pl, int gradflag,
              cx<R>& H, cx<R>& S, cx<R>* Dk, cx<R>* Dl) {
  const int n = sy.n, np4 = 2 * n * (n + 1);
  H = cx<R>(); S = cx<R>();
  const bool gk = (gradflag == 1 || gradflag == 3), gl = (gradflag == 2 || gradflag == 3);
  if (gk) for (int m = 0; m < np4; ++m) Dk[m] = cx<R>();
  if (gl) for (int m = 0; m < np4; ++m) Dl[m] = cx<R>();
  cx<R> h, s, dk[4 * NNP], dl[4 * NNP];
  for (int t = 0; t < sy.T; ++t) {
    matrix_elements(sy, pk, pl, sy.P + t * n * n, gradflag, h, s, dk, dl);
    R c = R(sy.coef[t]);
    H = H + c * h;
    S = S + c * s;
    if (gk) for (int m = 0; m < np4; ++m) Dk[m] = Dk[m] + c * dk[m];
    if (gl) for (int m = 0; m < np4; ++m) Dl[m] = Dl[m] + c * dl[m];
  }
}

}  // namespace ecgo
