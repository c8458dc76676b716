// C entry points of the CPU oracle (test infrastructure only; see the
// header of ecg_oracle.hpp).  Matrices are Fortran column-major, complex
// values are interleaved (re, im) doubles, function indices are 1-based
// where the reference's are.
#include <omp.h>

#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "ecg_oracle.hpp"

namespace ecgo {

// ------------------------------------------------ operation-counting type
// Every +, -, *, / and sqrt on this type adds one to the thread's counter
// while counting is on.  Used only to count the reference algorithm's work.
struct Fl {
  double v;
  Fl() : v(0) {}
  Fl(double x) : v(x) {}
};
thread_local long long g_flops = 0;
thread_local int g_pause = 0;
inline void tick() { if (!g_pause) ++g_flops; }
inline Fl operator+(Fl a, Fl b) { tick(); return Fl(a.v + b.v); }
inline Fl operator-(Fl a, Fl b) { tick(); return Fl(a.v - b.v); }
inline Fl operator*(Fl a, Fl b) { tick(); return Fl(a.v * b.v); }
inline Fl operator/(Fl a, Fl b) { tick(); return Fl(a.v / b.v); }
inline Fl operator-(Fl a) { return Fl(-a.v); }
inline bool operator>=(Fl a, Fl b) { return a.v >= b.v; }
inline bool operator==(Fl a, Fl b) { return a.v == b.v; }
inline Fl fabs(Fl a) { return Fl(std::fabs(a.v)); }
inline Fl sqrt(Fl a) { tick(); return Fl(std::sqrt(a.v)); }
inline Fl hypot(Fl a, Fl b) { tick(); tick(); tick(); tick(); return Fl(std::hypot(a.v, b.v)); }
inline Fl copysign(Fl a, Fl b) { return Fl(std::copysign(a.v, b.v)); }
template <> struct Pause<Fl> {
  Pause() { ++g_pause; }
  ~Pause() { --g_pause; }
};

template <class R>
static System<R> make_sys(int n, int T, const int* P, const int* coef, const double* M, double q0,
                          const double* q) {
  System<R> sy;
  sy.n = n;
  sy.T = T;
  for (int i = 0; i < NN; ++i)
    for (int j = 0; j < NN; ++j) sy.M[i][j] = R((i < n && j < n) ? M[i + n * j] : 0.0);
  sy.q0 = R(q0);
  for (int i = 0; i < NN; ++i) sy.q[i] = R(i < n ? q[i] : 0.0);
  sy.two3n2 = R(std::pow(2.0, (3 * n) / 2.0));
  sy.P = P;
  sy.coef = coef;
  return sy;
}

}  // namespace ecgo

using namespace ecgo;
typedef cx<double> cd;

static inline cd ld(const double* a, long i) { return cd(a[2 * i], a[2 * i + 1]); }
static inline void st(double* a, long i, cd v) { a[2 * i] = v.re; a[2 * i + 1] = v.im; }

extern "C" {

// Number of Y^dagger Y-summed pair elements of MatrixElements-work counted
// for one (pair, term) in the reference algorithm (src/matelem.f90:12-750),
// excluding per-function set-up and integer permutation products.
long long oracle_flops_per_pair(int n, const int* P, const double* M, double q0, const double* q,
                                const double* pk, const double* pl, int gradflag) {
  int coef = 1;
  System<Fl> sy = make_sys<Fl>(n, 1, P, &coef, M, q0, q);
  Fl pkf[2 * NNP], plf[2 * NNP];
  for (int i = 0; i < n * (n + 1); ++i) { pkf[i] = Fl(pk[i]); plf[i] = Fl(pl[i]); }
  cx<Fl> H, S, Dk[4 * NNP], Dl[4 * NNP];
  g_flops = 0;
  g_pause = 0;
  matrix_elements(sy, pkf, plf, P, gradflag, H, S, Dk, Dl);
  return g_flops;
}

// One pair summed over the T terms (MatrixElements x YHY coefficients).
// H, S: 1 complex each; Dk, Dl: 4*np complex each (may be null when the
// gradflag does not ask for them).
void oracle_pair(int n, int T, const int* P, const int* coef, const double* M, double q0,
                 const double* q, const double* pk, const double* pl, int gradflag, double* H,
                 double* S, double* Dk, double* Dl) {
  System<double> sy = make_sys<double>(n, T, P, coef, M, q0, q);
  cd h, s, dk[4 * NNP], dl[4 * NNP];
  pair_sum(sy, pk, pl, gradflag, h, s, dk, dl);
  st(H, 0, h);
  st(S, 0, s);
  const int np4 = 2 * n * (n + 1);
  if (Dk && (gradflag == 1 || gradflag == 3))
    for (int m = 0; m < np4; ++m) st(Dk, m, dk[m]);
  if (Dl && (gradflag == 2 || gradflag == 3))
    for (int m = 0; m < np4; ++m) st(Dl, m, dl[m]);
}

// Throughput probe for the CPU baseline: evaluates the pairs (kidx[i],
// lidx[i]) (0-based, bra k >= ket l) with gradflag 3 (1 on the diagonal),
// summed over the T terms, on nthreads OpenMP threads.  Returns the sum of
// Re(H) over the pairs (so the work cannot be optimised away).
double oracle_pair_batch(int n, int T, const int* P, const int* coef, const double* M, double q0,
                         const double* q, const double* params, long npairs, const int* kidx,
                         const int* lidx, int nthreads) {
  System<double> sy = make_sys<double>(n, T, P, coef, M, q0, q);
  const int npt = n * (n + 1);
  if (nthreads > 0) omp_set_num_threads(nthreads);
  double acc = 0.0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : acc)
  for (long i = 0; i < npairs; ++i) {
    cd h, s, dk[4 * NNP], dl[4 * NNP];
    const int k = kidx[i], l = lidx[i];
    pair_sum(sy, params + (long)npt * k, params + (long)npt * l, k == l ? 1 : 3, h, s, dk, dl);
    acc += h.re + dk[0].re + (k == l ? 0.0 : dl[0].re);
  }
  return acc;
}

// ComputeMatElem(Nmin, Nmax) + StoreHS (src/matform.f90:11-51, 119-230) for
// method 'G' or 'I'.  params: npt x K column-major (column k = function k).
// H, S: ld x ld complex column-major; diagH, diagS: K reals.  The rows
// nmin..nmax (1-based) are (re)computed; diagS of rows < nmin must already
// be present.  Diagonals are stored first, then the off-diagonal elements,
// which reproduces the reference's store order dependence.
int oracle_assemble(int n, int T, const int* P, const int* coef, const double* M, double q0,
                    const double* q, int K, const double* params, int nmin, int nmax, char method,
                    double sigma, double* H, double* S, int ld, double* diagH, double* diagS,
                    int nthreads) {
  if (n < 1 || n > NN || nmin < 1 || nmax > K || nmin > nmax || ld < nmax) return -1;
  if (method != 'G' && method != 'I') return -2;
  System<double> sy = make_sys<double>(n, T, P, coef, M, q0, q);
  const int npt = n * (n + 1);
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
  for (int k = nmin; k <= nmax; ++k) {
    cd h, s;
    pair_sum(sy, params + (long)npt * (k - 1), params + (long)npt * (k - 1), 0, h, s, (cd*)nullptr, (cd*)nullptr);
    const long kk = (long)(k - 1) * ld + (k - 1);
    diagS[k - 1] = s.re;
    if (method == 'G') {
      diagH[k - 1] = h.re / diagS[k - 1];
    } else {
      st(S, kk, cd(1.0, 0.0));
      st(H, kk, cd(h.re / diagS[k - 1] - sigma, 0.0));
    }
  }
  const long npairs_rows = nmax - nmin + 1;
#pragma omp parallel for schedule(dynamic, 1)
  for (long r = 0; r < npairs_rows; ++r) {
    const int k = nmin + (int)r;
    for (int l = k - 1; l >= 1; --l) {
      cd h, s;
      pair_sum(sy, params + (long)npt * (k - 1), params + (long)npt * (l - 1), 0, h, s, (cd*)nullptr, (cd*)nullptr);
      const double f = 1.0 / std::sqrt(diagS[k - 1] * diagS[l - 1]);
      const long kl = (long)(l - 1) * ld + (k - 1), lk = (long)(k - 1) * ld + (l - 1);
      if (method == 'G') {
        st(S, kl, s * f);
        st(H, kl, h * f);
      } else {
        cd sn = s * f;
        st(S, kl, sn);
        st(S, lk, conj(sn));
        st(H, kl, (h - sigma * s) * f);
      }
    }
  }
  return 0;
}

// ComputeMatElemAndDeriv(nfru+1, K) + StoreHSD (src/matform.f90:55-115,
// 234-360).  D: (2*npt) x nfo x ld complex, as Glob_D.
int oracle_assemble_deriv(int n, int T, const int* P, const int* coef, const double* M, double q0,
                          const double* q, int K, const double* params, int nfru, char method,
                          double sigma, double* H, double* S, int ld, double* diagH, double* diagS,
                          double* D, int nthreads) {
  if (n < 1 || n > NN || nfru < 0 || nfru >= K || ld < K) return -1;
  if (method != 'G' && method != 'I') return -2;
  System<double> sy = make_sys<double>(n, T, P, coef, M, q0, q);
  const int npt = n * (n + 1), np4 = 2 * npt, nfo = K - nfru;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  auto Didx = [&](int m, int kf, int j) { return (long)m + (long)np4 * ((long)kf + (long)nfo * j); };
#pragma omp parallel for schedule(dynamic, 4)
  for (int k = nfru + 1; k <= K; ++k) {
    cd h, s, dk[4 * NNP];
    pair_sum(sy, params + (long)npt * (k - 1), params + (long)npt * (k - 1), 1, h, s, dk, (cd*)nullptr);
    const long kk = (long)(k - 1) * ld + (k - 1);
    diagS[k - 1] = s.re;
    if (method == 'G') {
      diagH[k - 1] = h.re / diagS[k - 1];
    } else {
      st(S, kk, cd(1.0, 0.0));
      st(H, kk, cd(h.re / diagS[k - 1] - sigma, 0.0));
    }
    for (int m = 0; m < np4; ++m) st(D, Didx(m, k - 1 - nfru, k - 1), cd(2.0 * dk[m].re / diagS[k - 1], 0.0));
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = nfru + 1; k <= K; ++k) {
    for (int l = k - 1; l >= 1; --l) {
      const int gf = (l > nfru) ? 3 : 1;
      cd h, s, dk[4 * NNP], dl[4 * NNP];
      pair_sum(sy, params + (long)npt * (k - 1), params + (long)npt * (l - 1), gf, h, s, dk, dl);
      const double f = 1.0 / std::sqrt(diagS[k - 1] * diagS[l - 1]);
      const long kl = (long)(l - 1) * ld + (k - 1), lk = (long)(k - 1) * ld + (l - 1);
      if (method == 'G') {
        st(S, kl, s * f);
        st(H, kl, h * f);
      } else {
        cd sn = s * f;
        st(S, kl, sn);
        st(S, lk, conj(sn));
        st(H, kl, (h - sigma * s) * f);
      }
      for (int m = 0; m < np4; ++m) st(D, Didx(m, k - 1 - nfru, l - 1), dk[m] * f);
      if (l > nfru)
        for (int m = 0; m < np4; ++m) st(D, Didx(m, l - 1 - nfru, k - 1), conj(dl[m] * f));
    }
  }
  return 0;
}

// Gradient contraction of EnergyGB/EnergyIB (src/workproc.f90:1925-1940):
// grad((k-1)*npt+m) = 2 Re(conj(c_k') W_m) - |c_k'|^2 (D_H - E D_S)(m,k,k')
// with W_m = sum_l c_l (D_H - E D_S)(m,k,l), k' = k + nfru.
void oracle_gradient(int npt, int nfa, int nfru, const double* D, const double* c, double E,
                     double* grad) {
  const int nfo = nfa - nfru, np4 = 2 * npt;
  auto Dv = [&](int m, int kf, int j) { return ld(D, (long)m + (long)np4 * ((long)kf + (long)nfo * j)); };
  for (int k = 0; k < nfo; ++k) {
    std::vector<cd> W(npt);
    for (int l = 0; l < nfa; ++l) {
      cd t = ld(c, l);
      for (int m = 0; m < npt; ++m) W[m] = W[m] + t * (Dv(m, k, l) - E * Dv(m + npt, k, l));
    }
    cd t = ld(c, k + nfru);
    for (int m = 0; m < npt; ++m) {
      double g = 2.0 * (conj(t) * W[m]).re;
      g -= (t * conj(t) * (Dv(m, k, k + nfru) - E * Dv(m + npt, k, k + nfru))).re;
      grad[(long)k * npt + m] = g;
    }
  }
}

// ComputeOverlapPenalty (src/workproc.f90:1491-1567, single-process
// branch).  S: the stored normalised overlap, read at (j, i) with j > i
// (lower triangle), ldS its leading dimension; 1-based loops as written
// there: first i <= nfru against j > nfru, then i > nfru against j > i.
double oracle_overlap_penalty(int nfa, int nfru, const double* S, int ldS, double bmax, double t2) {
  const double pc = bmax / (1.0 - t2);
  auto Sv = [&](int j, int i) { return ld(S, (long)(i - 1) * ldS + (j - 1)); };
  double tp = 0.0;
  for (int i = 1; i <= nfru; ++i)
    for (int j = nfru + 1; j <= nfa; ++j) {
      const cd s = Sv(j, i);
      const double aux = s.re * s.re + s.im * s.im;
      if (aux > t2) tp += pc * (aux - t2);
    }
  for (int i = nfru + 1; i <= nfa; ++i)
    for (int j = i + 1; j <= nfa; ++j) {
      const cd s = Sv(j, i);
      const double aux = s.re * s.re + s.im * s.im;
      if (aux > t2) tp += pc * (aux - t2);
    }
  return tp;
}

// ComputeOverlapPenaltyAndAddGradient (src/workproc.f90:1571-1693,
// single-process branch): the penalty, and its gradient terms added to
// grad (npt * (nfa - nfru)) exactly as written there, from the stored
// derivatives D (2 npt x nfo x nfa, StoreHSD layout; the S part is rows
// npt..2npt-1) and diagS.  For a hit pair (j, i), j > i:
//   grad_j += 2 pc Re(S(j,i) conj(D(S,j,i) - D(S,j,j) S(j,i) tau / 2))
//   grad_i += 2 pc Re(S(j,i) (D(S,i,j) - D(S,i,i) S(j,i) / (2 tau)))
// with tau = sqrt(diagS_i / diagS_j).
double oracle_overlap_penalty_grad(int npt, int nfa, int nfru, const double* S, int ldS, const double* diagS,
                                   const double* D, double bmax, double t2, double* grad) {
  const int nfo = nfa - nfru, np4 = 2 * npt;
  const double pc = bmax / (1.0 - t2), temp1 = 2.0 * pc;
  auto Sv = [&](int j, int i) { return ld(S, (long)(i - 1) * ldS + (j - 1)); };
  // Glob_D(npt+m, kf, j) with 1-based m, kf, j
  auto DS = [&](int m, int kf, int j) {
    return ld(D, (long)(npt + m - 1) + (long)np4 * ((long)(kf - 1) + (long)nfo * (j - 1)));
  };
  auto add_j = [&](int j, int i, cd s) {
    const double temp2 = std::sqrt(diagS[i - 1] / diagS[j - 1]);
    for (int m = 1; m <= npt; ++m) {
      const cd t = DS(m, j - nfru, i) - 0.5 * DS(m, j - nfru, j) * s * temp2;
      grad[(long)(j - nfru - 1) * npt + m - 1] += temp1 * (s * conj(t)).re;
    }
  };
  double tp = 0.0;
  for (int i = 1; i <= nfru; ++i)
    for (int j = nfru + 1; j <= nfa; ++j) {
      const cd s = Sv(j, i);
      const double aux = s.re * s.re + s.im * s.im;
      if (aux > t2) {
        tp += pc * (aux - t2);
        add_j(j, i, s);
      }
    }
  for (int i = nfru + 1; i <= nfa; ++i)
    for (int j = i + 1; j <= nfa; ++j) {
      const cd s = Sv(j, i);
      const double aux = s.re * s.re + s.im * s.im;
      if (aux > t2) {
        tp += pc * (aux - t2);
        const double temp2 = std::sqrt(diagS[i - 1] / diagS[j - 1]);
        for (int m = 1; m <= npt; ++m) {
          const cd t = DS(m, i - nfru, j) - 0.5 * DS(m, i - nfru, i) * s / temp2;
          grad[(long)(i - nfru - 1) * npt + m - 1] += temp1 * (s * t).re;
        }
        add_j(j, i, s);
      }
    }
  return tp;
}

// LDLHF (src/linalg.f90:316-455, single-process branch): update the LDL^H
// factorisation of the Hermitian matrix held in A's lower triangle from
// row m (1-based) to n.  L^H goes to the strict upper triangle, 1/D to invD.
int oracle_ldlhf(int m, int n, double* A, int nA, double* invD) {
  std::vector<cd> w(n + 1);
  auto a = [&](int i, int j) { return ld(A, (long)(i - 1) + (long)nA * (j - 1)); };
  auto sa = [&](int i, int j, cd v) { st(A, (long)(i - 1) + (long)nA * (j - 1), v); };
  for (int i = m; i <= n; ++i) {
    for (int j = 1; j <= i; ++j) {
      const int jm = j - 1;
      if (i == j) {
        for (int q = 1; q <= jm; ++q) w[q] = a(q, i) * ld(invD, q - 1);
        cd x = a(i, j);
        for (int q = 1; q <= jm; ++q) x = x - conj(a(q, i)) * w[q];
        x = conj(x);
        for (int q = 1; q <= jm; ++q) sa(q, i, w[q]);
        if (x.re == 0.0 && x.im == 0.0) return 1;
        st(invD, j - 1, cd(1.0) / x);
      } else {
        cd x = a(i, j);
        for (int q = 1; q <= jm; ++q) x = x - conj(a(q, i)) * a(q, j);
        sa(j, i, conj(x));
      }
    }
  }
  return 0;
}

// LDLHS (src/linalg.f90:618-775, plain branch): x = A^{-1} b using L^H in
// the strict upper triangle of A and invD.
void oracle_ldlhs(int n, const double* A, int nA, const double* invD, const double* b, double* x) {
  auto a = [&](int i, int j) { return ld(A, (long)(i - 1) + (long)nA * (j - 1)); };
  std::vector<cd> v(n + 1);
  for (int j = 1; j <= n; ++j) v[j] = ld(b, j - 1);
  for (int j = 1; j <= n; ++j) {
    cd t = v[j];
    for (int i = 1; i < j; ++i) t = t - conj(a(i, j)) * v[i];
    v[j] = t;
  }
  for (int j = 1; j <= n; ++j) v[j] = v[j] * ld(invD, j - 1);
  for (int j = n; j >= 1; --j) {
    cd t = v[j];
    for (int i = j - 1; i >= 1; --i) v[i] = v[i] - t * a(i, j);
  }
  for (int j = 1; j <= n; ++j) st(x, j - 1, v[j]);
}

// x^H A x from the lower triangle of A (VMMHMV, src/linalg.f90:1180-1255).
static double vmmhmv(int n, const double* A, int nA, const std::vector<cd>& x) {
  double r = 0.0;
  for (int k = 0; k < n; ++k) {
    cd akk = ld(A, (long)k + (long)nA * k);
    r += (x[k].re * x[k].re + x[k].im * x[k].im) * akk.re;
    double sr = 0.0, si = 0.0;
    for (int j = k + 1; j < n; ++j) {
      cd ajk = ld(A, (long)j + (long)nA * k);
      sr += x[j].re * ajk.re + x[j].im * ajk.im;
      si += x[j].im * ajk.re - x[j].re * ajk.im;
    }
    r += 2.0 * (x[k].re * sr + x[k].im * si);
  }
  return r;
}

// GHEPIIS (src/linalg.f90:1781-1926): inverse iteration for the eigenvalue
// of (A, B) nearest apprlambda, with M holding A - apprlambda*B in its lower
// triangle and the known L^H rows < k in its upper triangle.  B is full.
// Returns ErrorCode (0, 1 singular, 2 not converged).  out3 = (lambda,
// RelAcc, NumIter).
int oracle_ghepiis(int k, int n, double* M, int nM, double* invD, const double* B, int nB,
                   double apprlambda, double* v, double Tol, int MaxIter, int SpecifNorm,
                   double* x_out, double* out3) {
  int err = oracle_ldlhf(k, n, M, nM, invD);
  out3[0] = 0.0; out3[1] = HUGE_VAL; out3[2] = 0.0;
  if (err) return err;
  const double sqrtn = std::sqrt(2.0 * n);
  std::vector<cd> vv(n), w(n), x(n);
  for (int i = 0; i < n; ++i) vv[i] = ld(v, i);
  std::vector<double> wb(2 * n), xb(2 * n);
  bool notconv = true;
  int it = 0;
  double nd = HUGE_VAL, ndprev = HUGE_VAL;
  while (notconv && it < MaxIter) {
    if (it != 0) vv = x;
    // w = B^H v (MHMV)
    for (int i = 0; i < n; ++i) {
      cd t;
      for (int j = 0; j < n; ++j) t = t + conj(ld(B, (long)j + (long)nB * i)) * vv[j];
      w[i] = t;
    }
    for (int i = 0; i < n; ++i) { wb[2 * i] = w[i].re; wb[2 * i + 1] = w[i].im; }
    oracle_ldlhs(n, M, nM, invD, wb.data(), xb.data());
    for (int i = 0; i < n; ++i) x[i] = cd(xb[2 * i], xb[2 * i + 1]);
    double t1 = 0.0;
    for (int i = 0; i < n; ++i) {
      t1 = std::fmax(t1, std::fabs(x[i].re));
      t1 = std::fmax(t1, std::fabs(x[i].im));
    }
    for (int i = 0; i < n; ++i) x[i] = x[i] * cd(1.0 / t1, 0.0);
    cd a;
    double r = 0.0;
    for (int i = 0; i < n; ++i) {
      a = a + conj(x[i]) * vv[i];
      r += vv[i].re * vv[i].re + vv[i].im * vv[i].im;
    }
    cd tc(a.re / r, a.im / r);
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      cd d = x[i] - tc * vv[i];
      s += d.re * d.re + d.im * d.im;
    }
    nd = sqrtn * std::sqrt(s);
    if (Tol > 0.0) {
      if (nd <= Tol) notconv = false;
    } else {
      if (nd > ndprev && nd <= std::fabs(Tol)) notconv = false;
      ndprev = nd;
    }
    ++it;
  }
  if (notconv) err = 2;
  const double t1 = vmmhmv(n, B, nB, x);
  const double t2 = vmmhmv(n, M, nM, x);
  const double lambda = t2 / t1 + apprlambda;
  if (SpecifNorm == 0) {
    for (int i = 0; i < n; ++i) x[i] = x[i] * cd(1.0 / std::sqrt(t1), 0.0);
  } else if (SpecifNorm == 1) {
    double q = 0.0;
    for (int i = 0; i < n; ++i) q += x[i].re * x[i].re + x[i].im * x[i].im;
    for (int i = 0; i < n; ++i) x[i] = x[i] * cd(1.0 / std::sqrt(q), 0.0);
  }
  for (int i = 0; i < n; ++i) st(x_out, i, x[i]);
  out3[0] = lambda;
  out3[1] = nd;
  out3[2] = it;
  return err;
}


// ------------------------------------------------- expectation values
// Pair operators of MatrixElementsForExpcVals (src/matelem.f90:1236-1275,
// 1306-1349) for symmetry term `term`, bra unpermuted: for every
// (a, b >= a) (a = b: pseudo-particle a's distance from particle 1), with
// TrCJ = tr(C^{-1} J_ab) = C^{-1}_aa (+ C^{-1}_bb - 2 C^{-1}_ab):
//   1/r^2 = 2 S / TrCJ,  1/r = (2/sqrt(pi)) S / sqrt(TrCJ),
//   r = (2/sqrt(pi)) S sqrt(TrCJ),  r^2 = (3/2) S TrCJ,
//   delta(r) = S / (pi^(3/2) sqrt(TrCJ) TrCJ);
// then H, S, T, V.  out: 5 n(n+1)/2 + 4 complex values, each block in the
// driver's (a, b >= a) order (src/workproc.f90:8183-8212).
static void expc_pair(const System<double>& sy, const double* pk, const double* pl, const int* Pm, cd* out) {
  cd H, S;
  PairParts<double> pp;
  matrix_elements<double>(sy, pk, pl, Pm, 0, H, S, nullptr, nullptr, &pp);
  const int n = sy.n, m = n * (n + 1) / 2;
  const double sqrtpi = 1.77245385090551602729816748334115, pi = 3.14159265358979323846264338327950;
  const cd cnumb = (2.0 / sqrtpi) * S, caux = 1.5 * S, caux1 = S / (pi * sqrtpi), cmp = 2.0 * S;
  int c = 0;
  for (int a = 0; a < n; ++a)
    for (int b = a; b < n; ++b, ++c) {
      const cd t = (a == b) ? pp.Ci[a][a] : pp.Ci[a][a] + pp.Ci[b][b] - pp.Ci[b][a] - pp.Ci[b][a];
      const cd sq = csqrt(t);
      out[c] = cmp / t;
      out[m + c] = cnumb / sq;
      out[2 * m + c] = cnumb * sq;
      out[3 * m + c] = caux * t;
      out[4 * m + c] = caux1 / (sq * t);
    }
  out[5 * m] = H;
  out[5 * m + 1] = S;
  out[5 * m + 2] = pp.T;
  out[5 * m + 3] = pp.V;
}

void oracle_expc_pair(int n, int T, const int* P, const int* coef, const double* M, double q0, const double* q,
                      const double* pk, const double* pl, int term, double* out) {
  System<double> sy = make_sys<double>(n, T, P, coef, M, q0, q);
  const int nobs = 5 * n * (n + 1) / 2 + 4;
  std::vector<cd> o(nobs);
  expc_pair(sy, pk, pl, P + (size_t)term * n * n, o.data());
  for (int a = 0; a < nobs; ++a) st(out, a, o[a]);
}

// The driver's sums (src/workproc.f90:8158-8230 and :8279, symmetry
// adaptation through the ket's Y^dagger Y terms): over pairs i >= j,
// factor = 1/diagS_i (i = j) or 2/sqrt(diagS_i diagS_j), per term k and
// value a, with O = O_ij of term k:
//   beta = Re(conj(c_i) c_j O),
//   mu = |Re c_i (Re c_j Re O - Im c_j Im O)| + |Re c_j (Re c_i Re O + Im c_i Im O)|
//      + |Im c_i (Im c_j Re O + Re c_j Im O)| + |Im c_j (Im c_i Re O - Re c_i Im O)|,
//   out_a += factor coef_k beta,  err_a += |factor coef_k (mu + 2 |beta|)|,
// and err is finally scaled by epsilon(1.0).  c: K complex.
int oracle_expectation_values(int n, int T, const int* P, const int* coef, const double* M, double q0,
                              const double* q, int K, const double* params, const double* c,
                              const double* diagS, double* out, double* err, int nthreads) {
  System<double> sy = make_sys<double>(n, T, P, coef, M, q0, q);
  const int npt = n * (n + 1), nobs = 5 * n * (n + 1) / 2 + 4;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const int nth = omp_get_max_threads();
  std::vector<double> acc((size_t)nth * 2 * nobs, 0.0);
#pragma omp parallel
  {
    const int tid = omp_get_thread_num();
    double* s_ = acc.data() + (size_t)tid * 2 * nobs;
    double* e_ = s_ + nobs;
    std::vector<cd> o(nobs);
#pragma omp for schedule(dynamic, 4)
    for (int i = 0; i < K; ++i)
      for (int j = 0; j <= i; ++j) {
        const double rci = c[2 * i], ici = c[2 * i + 1], rcj = c[2 * j], icj = c[2 * j + 1];
        const double factor = (i == j) ? 1.0 / std::sqrt(diagS[i] * diagS[i]) : 2.0 / std::sqrt(diagS[i] * diagS[j]);
        for (int k = 0; k < T; ++k) {
          expc_pair(sy, params + (size_t)npt * i, params + (size_t)npt * j, P + (size_t)k * n * n, o.data());
          for (int a = 0; a < nobs; ++a) {
            const double ro = o[a].re, io = o[a].im;
            const double beta = (rci * rcj + ici * icj) * ro + (ici * rcj - rci * icj) * io;
            const double mu = std::fabs(rci * (rcj * ro - icj * io)) + std::fabs(rcj * (rci * ro + ici * io)) +
                              std::fabs(ici * (icj * ro + rcj * io)) + std::fabs(icj * (ici * ro - rci * io));
            s_[a] += factor * coef[k] * beta;
            e_[a] += std::fabs(factor * coef[k] * (mu + 2.0 * std::fabs(beta)));
          }
        }
      }
  }
  for (int a = 0; a < nobs; ++a) {
    double sa = 0.0, ea = 0.0;
    for (int t = 0; t < nth; ++t) {
      sa += acc[(size_t)t * 2 * nobs + a];
      ea += acc[(size_t)t * 2 * nobs + nobs + a];
    }
    out[a] = sa;
    if (err) err[a] = ea * 2.220446049250313e-16;
  }
  return 0;
}

// Correlation functions (src/matelem.f90:1934-1946): per pair (a, b >= a)
// and radius r, CF = S exp(-r^2 / TrCJ) / (sqrt(TrCJ) TrCJ) / (pi sqrt(pi)),
// TrCJ = w^T C^-1 w; contracted as the driver's CFDMEkl_s
// (src/workproc.f90:8234-8243): out[g m + q] += factor coef_k beta.
int oracle_correlation_functions(int n, int T, const int* P, const int* coef, const double* M, double q0,
                                 const double* q, int K, const double* params, const double* c,
                                 const double* diagS, int ngrid, const double* grid, double* out, int nthreads) {
  System<double> sy = make_sys<double>(n, T, P, coef, M, q0, q);
  const int npt = n * (n + 1), m = n * (n + 1) / 2, nv = m * ngrid;
  const double sqrtpi = 1.77245385090551602729816748334115, pi = 3.14159265358979323846264338327950;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const int nth = omp_get_max_threads();
  std::vector<double> acc((size_t)nth * nv, 0.0);
#pragma omp parallel
  {
    const int tid = omp_get_thread_num();
    double* s_ = acc.data() + (size_t)tid * nv;
#pragma omp for schedule(dynamic, 4)
    for (int i = 0; i < K; ++i)
      for (int j = 0; j <= i; ++j) {
        const double rci = c[2 * i], ici = c[2 * i + 1], rcj = c[2 * j], icj = c[2 * j + 1];
        const double factor = (i == j) ? 1.0 / std::sqrt(diagS[i] * diagS[i]) : 2.0 / std::sqrt(diagS[i] * diagS[j]);
        for (int k = 0; k < T; ++k) {
          cd H, S;
          PairParts<double> pp;
          matrix_elements<double>(sy, params + (size_t)npt * i, params + (size_t)npt * j, P + (size_t)k * n * n, 0,
                                  H, S, nullptr, nullptr, &pp);
          const cd ctemp = S / (pi * sqrtpi);
          int qq = 0;
          for (int a = 0; a < n; ++a)
            for (int b = a; b < n; ++b, ++qq) {
              const cd t = (a == b) ? pp.Ci[a][a] : pp.Ci[a][a] + pp.Ci[b][b] - pp.Ci[b][a] - pp.Ci[b][a];
              const cd den = csqrt(t) * t;
              for (int g = 0; g < ngrid; ++g) {
                const cd e = cd{-grid[g] * grid[g], 0.0} / t;
                const double mag = std::exp(e.re);
                const cd ex = {mag * std::cos(e.im), mag * std::sin(e.im)};
                const cd o = (ctemp * ex) / den;
                const double beta = (rci * rcj + ici * icj) * o.re + (ici * rcj - rci * icj) * o.im;
                s_[(size_t)g * m + qq] += factor * coef[k] * beta;
              }
            }
        }
      }
  }
  for (int a = 0; a < nv; ++a) {
    double sa = 0.0;
    for (int t = 0; t < nth; ++t) sa += acc[(size_t)t * nv + a];
    out[a] = sa;
  }
  return 0;
}

// Particle densities (src/matelem.f90:1948-1963) with Glob_bvc
// (src/workproc.f90:612-621) built from the n + 1 masses: per particle i and
// radius r, Dens = S exp(-r^2 / caux) / (sqrt(caux) caux) / (pi sqrt(pi)),
// caux = sum_p [bvc(p,i)^2 Ci(p,p) + sum_{q>p} 2 bvc(q,i) bvc(p,i) Ci(q,p)];
// contracted as the driver's CFDMEkl_s (src/workproc.f90:8245-8254).
int oracle_particle_densities(int n, int T, const int* P, const int* coef, const double* M, double q0,
                              const double* q, int K, const double* params, const double* c, const double* diagS,
                              const double* masses, int ngrid, const double* grid, double* out, int nthreads) {
  System<double> sy = make_sys<double>(n, T, P, coef, M, q0, q);
  const int npt = n * (n + 1), npart = n + 1, nv = npart * ngrid;
  const double sqrtpi = 1.77245385090551602729816748334115, pi = 3.14159265358979323846264338327950;
  double mtot = 0.0;
  for (int i = 0; i < npart; ++i) mtot += masses[i];
  std::vector<double> bvc((size_t)n * npart);   // bvc[p + n i]
  for (int i = 0; i < npart; ++i)
    for (int p = 0; p < n; ++p) bvc[p + (size_t)n * i] = -masses[p + 1] / mtot;
  for (int i = 1; i < npart; ++i) bvc[(i - 1) + (size_t)n * i] += 1.0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const int nth = omp_get_max_threads();
  std::vector<double> acc((size_t)nth * nv, 0.0);
#pragma omp parallel
  {
    const int tid = omp_get_thread_num();
    double* s_ = acc.data() + (size_t)tid * nv;
#pragma omp for schedule(dynamic, 4)
    for (int i = 0; i < K; ++i)
      for (int j = 0; j <= i; ++j) {
        const double rci = c[2 * i], ici = c[2 * i + 1], rcj = c[2 * j], icj = c[2 * j + 1];
        const double factor = (i == j) ? 1.0 / std::sqrt(diagS[i] * diagS[i]) : 2.0 / std::sqrt(diagS[i] * diagS[j]);
        for (int k = 0; k < T; ++k) {
          cd H, S;
          PairParts<double> pp;
          matrix_elements<double>(sy, params + (size_t)npt * i, params + (size_t)npt * j, P + (size_t)k * n * n, 0,
                                  H, S, nullptr, nullptr, &pp);
          const cd ctemp = S / (pi * sqrtpi);
          for (int pa = 0; pa < npart; ++pa) {
            const double* b = bvc.data() + (size_t)n * pa;
            cd caux = {0.0, 0.0};
            for (int p = 0; p < n; ++p) {
              caux = caux + (b[p] * b[p]) * pp.Ci[p][p];
              for (int qq = p + 1; qq < n; ++qq) caux = caux + (2.0 * b[qq] * b[p]) * pp.Ci[qq][p];
            }
            const cd den = csqrt(caux) * caux;
            for (int g = 0; g < ngrid; ++g) {
              const cd e = cd{-grid[g] * grid[g], 0.0} / caux;
              const double mag = std::exp(e.re);
              const cd ex = {mag * std::cos(e.im), mag * std::sin(e.im)};
              const cd o = (ctemp * ex) / den;
              const double beta = (rci * rcj + ici * icj) * o.re + (ici * rcj - rci * icj) * o.im;
              s_[(size_t)g * npart + pa] += factor * coef[k] * beta;
            }
          }
        }
      }
  }
  for (int a = 0; a < nv; ++a) {
    double sa = 0.0;
    for (int t = 0; t < nth; ++t) sa += acc[(size_t)t * nv + a];
    out[a] = sa;
  }
  return 0;
}

// Mass-velocity and Darwin pair values (src/matelem.f90:1479-1667 and
// 1866-1884), restated with the reference's loops: for particle 1 the
// column sums of tCk (and tCl) stand for J = 1 1^T, for particle i + 1 the
// column i (J = e_i e_i^T);
//   Wc1 = tCk J tCk, tr1 = tr(J tCk), Wc2 = tCl J tCl, tr2 = tr(J tCl),
//   MV_i = 36 S (tr(C^-1 Wc1) tr(C^-1 Wc2) + 2/3 tr(C^-1 Wc1 C^-1 Wc2)
//                - tr(C^-1 (tr1 Wc2 + tr2 Wc1)) + tr1 tr2) / m_i^3,
//   MV = -sum_i MV_i / 8;
//   Darwin = -pi/2 [sum_i (1/m_0^2 + 1/m_i^2) q0 q_i delta(i, i)
//                   + sum_{i != j} q_i q_j delta(i, j) / m_i^2].
static void relat_pair(const System<double>& sy, const double* pk, const double* pl, const int* Pm,
                       const double* masses, cd* out) {
  cd H, S;
  PairParts<double> pp;
  matrix_elements<double>(sy, pk, pl, Pm, 0, H, S, nullptr, nullptr, &pp);
  const int n = sy.n;
  const double sqrtpi = 1.77245385090551602729816748334115, pi = 3.14159265358979323846264338327950;
  cd MV = {0.0, 0.0};
  for (int part = 0; part <= n; ++part) {
    cd Wc1[NN][NN], Wc2[NN][NN], tr1, tr2;
    if (part == 0) {
      cd uk[NN], ul[NN];
      for (int p = 0; p < n; ++p) {
        uk[p] = {0.0, 0.0};
        ul[p] = {0.0, 0.0};
        for (int q = 0; q < n; ++q) {
          uk[p] = uk[p] + pp.Ck[q][p];
          ul[p] = ul[p] + pp.tCl[q][p];
        }
      }
      tr1 = {0.0, 0.0};
      tr2 = {0.0, 0.0};
      for (int p = 0; p < n; ++p) {
        tr1 = tr1 + uk[p];
        tr2 = tr2 + ul[p];
      }
      for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) {
          cd a = {0.0, 0.0}, b = {0.0, 0.0};
          for (int i = 0; i < n; ++i) {
            a = a + pp.Ck[p][i] * uk[q];
            b = b + pp.tCl[p][i] * ul[q];
          }
          Wc1[p][q] = a;
          Wc2[p][q] = b;
        }
    } else {
      const int i = part - 1;
      for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) {
          Wc1[q][p] = pp.Ck[q][i] * pp.Ck[i][p];
          Wc2[q][p] = pp.tCl[q][i] * pp.tCl[i][p];
        }
      tr1 = pp.Ck[i][i];
      tr2 = pp.tCl[i][i];
    }
    cd Wc4[NN][NN], Wc5[NN][NN];
    for (int q = 0; q < n; ++q)
      for (int p = 0; p < n; ++p) {
        cd a = {0.0, 0.0}, b = {0.0, 0.0};
        for (int j = 0; j < n; ++j) {
          a = a + pp.Ci[q][j] * Wc1[j][p];
          b = b + pp.Ci[q][j] * Wc2[j][p];
        }
        Wc4[q][p] = a;
        Wc5[q][p] = b;
      }
    cd tr4 = {0.0, 0.0}, tr5 = {0.0, 0.0}, cnumb = {0.0, 0.0}, caux = {0.0, 0.0};
    for (int p = 0; p < n; ++p) {
      tr4 = tr4 + Wc4[p][p];
      tr5 = tr5 + Wc5[p][p];
      for (int q = 0; q < n; ++q) {
        cnumb = cnumb + Wc4[p][q] * Wc5[q][p];
        caux = caux + pp.Ci[p][q] * (tr1 * Wc2[q][p] + tr2 * Wc1[q][p]);
      }
    }
    const double m3 = masses[part] * masses[part] * masses[part];
    MV = MV + (36.0 * S * (tr4 * tr5 + (2.0 / 3.0) * cnumb - caux + tr1 * tr2)) / m3;
  }
  out[0] = -0.125 * MV;
  // Darwin from the pair delta functions
  cd darwin = {0.0, 0.0};
  auto delta = [&](int a, int b) {
    const cd t = (a == b) ? pp.Ci[a][a] : pp.Ci[a][a] + pp.Ci[b][b] - pp.Ci[b][a] - pp.Ci[b][a];
    return (S / (pi * sqrtpi)) / (csqrt(t) * t);
  };
  for (int i = 0; i < n; ++i)
    darwin = darwin + ((1.0 / (masses[0] * masses[0]) + 1.0 / (masses[i + 1] * masses[i + 1])) * sy.q0 * sy.q[i]) *
                          delta(i, i);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (j != i) darwin = darwin + (sy.q[i] * sy.q[j] / (masses[i + 1] * masses[i + 1])) * delta(i, j);
  out[1] = (-pi / 2.0) * darwin;
}

void oracle_relat_pair(int n, int T, const int* P, const int* coef, const double* M, double q0, const double* q,
                       const double* pk, const double* pl, int term, const double* masses, double* out) {
  System<double> sy = make_sys<double>(n, T, P, coef, M, q0, q);
  cd o[2];
  relat_pair(sy, pk, pl, P + (size_t)term * n * n, masses, o);
  st(out, 0, o[0]);
  st(out, 1, o[1]);
}

// ------------------------------------ Drachmanized forms and orbit-orbit
// Restatement of MatrixElementsForExpcVals' remaining pair values
// (src/matelem.f90:1276-1478 and 1669-1933) and its helper integrals
// ME_rXr_over_rij, ME_rXr_rYr_over_rij (:753-945), ME_dXd_dYd, ME_dXd and
// ME_dXd_over_rij (:1998-2044), loop for loop, for a pair whose bra is
// unpermuted and ket permuted by Pm (the driver's form).  Matrices are
// A[row][column].  Inputs beyond the pair: the particle masses, the
// charges, the mass matrix and the current energy E (Glob_CurrEnergy).
//
// tr[C^-1 J_ij C^-1 J_pq] is kept, as the reference keeps it, in a 4-index
// table filled by its loops (:1275-1302).  Those loops never write the
// entries (i, i, p, q) with i < p < q (n >= 3), and the reference reads
// them in the 1/(r_ij r_pq) table (:1305-1343) used by the Drachmanized
// delta.  Its build (ifort, no -auto: local arrays are static, zero at
// start) reads zero there, which sends those entries to the uncorrelated
// branch (4 S / pi) / sqrt(TrCJ_ii TrCJ_pq); the table below starts at
// zero so that the same values result.
typedef cd Mat[NN][NN];

static cd clog_(const cd& z) { return {std::log(std::hypot(z.re, z.im)), std::atan2(z.im, z.re)}; }

// <phi_k| r'Xr / r_ij |phi_l> (:753-818)
static cd me_rxr_over_rij(int n, const Mat& X, int i, int j, const Mat& Ci, const cd& rm, const cd& trcj) {
  cd trcx = {0.0, 0.0};
  for (int m = 0; m < n; ++m)
    for (int p = 0; p < n; ++p) trcx = trcx + Ci[m][p] * X[p][m];
  cd cxi[NN], cxj[NN];
  for (int m = 0; m < n; ++m) {
    cxi[m] = {0.0, 0.0};
    for (int p = 0; p < n; ++p) cxi[m] = cxi[m] + Ci[i][p] * X[p][m];
  }
  if (j != i)
    for (int m = 0; m < n; ++m) {
      cxj[m] = {0.0, 0.0};
      for (int p = 0; p < n; ++p) cxj[m] = cxj[m] + Ci[j][p] * X[p][m];
    }
  cd t = {0.0, 0.0};
  if (i == j)
    for (int m = 0; m < n; ++m) t = t + cxi[m] * Ci[m][i];
  else
    for (int m = 0; m < n; ++m) t = t + (cxi[m] - cxj[m]) * (Ci[m][i] - Ci[m][j]);
  return rm * (3.0 * trcx - t / trcj) / 2.0;
}

// <phi_k| (r'Xr)(r'Yr) / r_ij |phi_l> (:821-945)
static cd me_rxr_ryr_over_rij(int n, const Mat& X, const Mat& Y, int i, int j, const Mat& Ci, const cd& rm,
                              const cd& trcj) {
  Mat Xs, Ys, CX, CY;
  for (int p = 0; p < n; ++p) {
    Ys[p][p] = Y[p][p];
    for (int q = p + 1; q < n; ++q) Ys[q][p] = Ys[p][q] = (Y[p][q] + Y[q][p]) / 2.0;
  }
  for (int p = 0; p < n; ++p) {
    Xs[p][p] = X[p][p];
    for (int q = p + 1; q < n; ++q) Xs[q][p] = Xs[p][q] = (X[p][q] + X[q][p]) / 2.0;
  }
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q) {
      CY[p][q] = {0.0, 0.0};
      for (int m = 0; m < n; ++m) CY[p][q] = CY[p][q] + Ci[p][m] * Ys[m][q];
    }
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q) {
      CX[p][q] = {0.0, 0.0};
      for (int m = 0; m < n; ++m) CX[p][q] = CX[p][q] + Ci[p][m] * Xs[m][q];
    }
  cd ci_[NN], cj_[NN];
  for (int m = 0; m < n; ++m) {
    ci_[m] = {0.0, 0.0};
    for (int p = 0; p < n; ++p) ci_[m] = ci_[m] + CX[i][p] * CY[p][m];
  }
  if (j != i)
    for (int m = 0; m < n; ++m) {
      cj_[m] = {0.0, 0.0};
      for (int p = 0; p < n; ++p) cj_[m] = cj_[m] + CX[j][p] * CY[p][m];
    }
  cd trcy = {0.0, 0.0}, trcx = {0.0, 0.0}, trcxcy = {0.0, 0.0};
  for (int m = 0; m < n; ++m) {
    trcy = trcy + CY[m][m];
    trcx = trcx + CX[m][m];
  }
  for (int m = 0; m < n; ++m)
    for (int p = 0; p < n; ++p) trcxcy = trcxcy + CX[m][p] * CY[p][m];
  cd tx = {0.0, 0.0}, ty = {0.0, 0.0}, txy = {0.0, 0.0};
  if (i == j) {
    for (int m = 0; m < n; ++m) tx = tx + CX[i][m] * Ci[m][i];
    for (int m = 0; m < n; ++m) ty = ty + CY[i][m] * Ci[m][i];
    for (int m = 0; m < n; ++m) txy = txy + ci_[m] * Ci[m][i];
  } else {
    for (int m = 0; m < n; ++m) tx = tx + (CX[i][m] - CX[j][m]) * (Ci[m][i] - Ci[m][j]);
    for (int m = 0; m < n; ++m) ty = ty + (CY[i][m] - CY[j][m]) * (Ci[m][i] - Ci[m][j]);
    for (int m = 0; m < n; ++m) txy = txy + (ci_[m] - cj_[m]) * (Ci[m][i] - Ci[m][j]);
  }
  return rm * (9.0 * trcy * trcx + 6.0 * trcxcy - (3.0 * (trcy * tx + trcx * ty) + 4.0 * txy) / trcj +
               3.0 * tx * ty / (trcj * trcj)) / 4.0;
}

static void matmul_(int n, const Mat& A, const Mat& B, Mat& C) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      cd s = {0.0, 0.0};
      for (int k = 0; k < n; ++k) s = s + A[i][k] * B[k][j];
      C[i][j] = s;
    }
}
static cd trace_(int n, const Mat& A) {
  cd s = {0.0, 0.0};
  for (int i = 0; i < n; ++i) s = s + A[i][i];
  return s;
}

// ME_dXd_dYd (:1998-2016)
static cd me_dxd_dyd(int n, const Mat& X, const Mat& Y, const Mat& Ci, const Mat& Ck, const Mat& Cl, const cd& S) {
  Mat M, MM, Z, t1, t2;
  matmul_(n, X, Ck, M);
  const cd tr1 = trace_(n, M);
  matmul_(n, Cl, Y, M);
  const cd tr2 = trace_(n, M);
  matmul_(n, X, Ck, t1);
  matmul_(n, Ck, t1, t2);
  matmul_(n, Ci, t2, M);
  const cd tr3 = trace_(n, M);
  matmul_(n, Y, Cl, t1);
  matmul_(n, Cl, t1, t2);
  matmul_(n, Ci, t2, MM);
  const cd tr4 = trace_(n, MM);
  matmul_(n, M, MM, Z);
  const cd tr5 = trace_(n, Z);
  return (36.0 * tr3 * tr4 + 24.0 * tr5 - 36.0 * tr1 * tr4 - 36.0 * tr2 * tr3 + 36.0 * tr1 * tr2) * S;
}

// ME_dXd (:2018-2030)
static cd me_dxd(int n, const Mat& X, const Mat& Ci, const Mat& Cl, const cd& S) {
  Mat M, t1, t2;
  matmul_(n, Cl, X, M);
  const cd tr1 = trace_(n, M);
  matmul_(n, X, Cl, t1);
  matmul_(n, Cl, t1, t2);
  matmul_(n, Ci, t2, M);
  const cd tr2 = trace_(n, M);
  return 6.0 * (tr2 - tr1) * S;
}

// ME_dXd_over_rij (:2032-2044)
static cd me_dxd_over_rij(int n, const Mat& X, int i, int j, const Mat& Ci, const Mat& Cl, const cd& rm,
                          const cd& trcj) {
  Mat M, Z, t1;
  matmul_(n, X, Cl, t1);
  matmul_(n, Cl, t1, M);
  matmul_(n, Cl, X, Z);
  const cd tr1 = trace_(n, Z);
  return 4.0 * me_rxr_over_rij(n, M, i, j, Ci, rm, trcj) - 6.0 * tr1 * rm;
}

// out: the Drachmanized delta of (a, b >= a) (n(n+1)/2 values), then OO,
// the Drachmanized mass-velocity and the Drachmanized Darwin term.
static void drach_pair(const System<double>& sy, const double* pk, const double* pl, const int* Pm,
                       const double* masses, double E, cd* out) {
  cd H, S;
  PairParts<double> pp;
  matrix_elements<double>(sy, pk, pl, Pm, 0, H, S, nullptr, nullptr, &pp);
  const int n = sy.n;
  const double pi = 3.14159265358979323846264338327950, sqrtpi = 1.77245385090551602729816748334115;
  Mat Ci, Ck, Cl, Ckl;
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      Ci[a][b] = pp.Ci[a][b];
      Ck[a][b] = pp.Ck[a][b];
      Cl[a][b] = pp.tCl[a][b];
      Ckl[a][b] = Ck[a][b] + Cl[a][b];
    }
  auto mass = [&](int part) { return masses[part]; };   // particle 1 = 0, pseudo-particle a = a + 1
  // TrCJ, <1/r_ij> (:1236-1273)
  Mat trcj, rm;
  const cd cn = (2.0 / sqrtpi) * S;
  for (int a = 0; a < n; ++a) {
    trcj[a][a] = Ci[a][a];
    rm[a][a] = cn / csqrt(trcj[a][a]);
  }
  for (int a = 0; a < n; ++a)
    for (int b = a + 1; b < n; ++b) {
      trcj[b][a] = trcj[a][b] = Ci[a][a] + Ci[b][b] - Ci[b][a] - Ci[b][a];
      rm[b][a] = rm[a][b] = cn / csqrt(trcj[a][b]);
    }
  // tr[C^-1 J_ij C^-1 J_pq] (:1275-1302)
  static thread_local cd tc[NN][NN][NN][NN];
  for (int a = 0; a < NN; ++a)
    for (int b = 0; b < NN; ++b)
      for (int c = 0; c < NN; ++c)
        for (int d = 0; d < NN; ++d) tc[a][b][c][d] = {0.0, 0.0};
  for (int i = 0; i < n; ++i) {
    tc[i][i][i][i] = Ci[i][i] * Ci[i][i];
    for (int j = i + 1; j < n; ++j) {
      const cd c0 = Ci[j][i] * Ci[j][i];
      tc[j][j][i][i] = c0;
      tc[i][i][j][j] = c0;
      for (int p = i; p < n; ++p) {
        const cd t = Ci[p][i] - Ci[p][j];
        const cd c1 = t * t;
        tc[i][j][p][p] = c1;
        tc[j][i][p][p] = c1;
        tc[p][p][i][j] = c1;
        tc[p][p][j][i] = c1;
        for (int q = p + 1; q < n; ++q) {
          const cd u = Ci[q][i] - Ci[p][i] - Ci[q][j] + Ci[p][j];
          const cd c2 = u * u;
          tc[i][j][p][q] = c2;
          tc[j][i][p][q] = c2;
          tc[i][j][q][p] = c2;
          tc[j][i][q][p] = c2;
          tc[p][q][i][j] = c2;
          tc[p][q][j][i] = c2;
          tc[q][p][i][j] = c2;
          tc[q][p][j][i] = c2;
        }
      }
    }
  }
  // <1/(r_ij r_pq)> (:1305-1343)
  static thread_local cd rr[NN][NN][NN][NN];
  const cd c4 = (4.0 * S) / pi, c2s = 2.0 * S;
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j)
      for (int p = i; p < n; ++p)
        for (int q = p; q < n; ++q) {
          cd v;
          if ((p == i && q == j) || (p == j && q == i)) {
            v = c2s / trcj[i][j];
          } else {
            const cd ca = csqrt(tc[i][j][p][q]);
            if (ca.re != 0.0 || ca.im != 0.0) {
              const cd ca1 = ca / csqrt(trcj[i][j] * trcj[p][q]);
              const cd w = cd{0.0, 1.0} * ca1 + csqrt(cd{1.0, 0.0} - ca1 * ca1);
              const cd as = cd{0.0, -1.0} * clog_(w);
              v = c4 * as / ca;
            } else {
              v = c4 / csqrt(trcj[i][j] * trcj[p][q]);
            }
          }
          rr[i][j][p][q] = v;
          rr[j][i][p][q] = v;
          rr[i][j][q][p] = v;
          rr[j][i][q][p] = v;
          rr[p][q][i][j] = v;
          rr[p][q][j][i] = v;
          rr[q][p][i][j] = v;
          rr[q][p][j][i] = v;
        }
  // Drachmanized delta(r_pq) (:1353-1392)
  Mat W1, W2, W3, W4, W5;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      cd t = {0.0, 0.0};
      for (int k = 0; k < n; ++k) t = t + sy.M[i][k] * Cl[k][j];
      W1[i][j] = t;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      cd t = {0.0, 0.0};
      for (int k = 0; k < n; ++k) t = t + Ck[i][k] * W1[k][j];
      W2[i][j] = t;
    }
  int o = 0;
  for (int p = 0; p < n; ++p)
    for (int q = p; q < n; ++q, ++o) {
      cd ct = {0.0, 0.0};
      for (int i = 0; i < n; ++i) {
        ct = ct + (sy.q0 * sy.q[i]) * rr[p][q][i][i];
        for (int j = i + 1; j < n; ++j) ct = ct + (sy.q[i] * sy.q[j]) * rr[p][q][i][j];
      }
      const double ca = (p == q) ? 2.0 * pi * sy.M[p][p]
                                 : 2.0 * pi * (sy.M[p][p] + sy.M[q][q] - sy.M[p][q] - sy.M[p][q]);
      const cd cnm = me_rxr_over_rij(n, W2, p, q, Ci, rm[p][q], trcj[p][q]);
      out[o] = (E * rm[p][q] - ct - 4.0 * cnm) / ca;
    }
  const int m = o;
  // orbit-orbit (:1394-1477)
  cd OO = {0.0, 0.0};
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const cd tr1 = Cl[j][i], tr2 = tr1, tr3 = 3.0 * Cl[j][j];
      for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) W1[q][p] = Cl[q][i] * Cl[p][j] + Cl[q][j] * Cl[p][i];
      for (int p = 0; p < n; ++p) W1[p][j] = W1[p][j] + Ckl[p][j] * Cl[j][i];
      for (int p = 0; p < n; ++p) W1[j][p] = W1[j][p] + Cl[i][j] * Cl[p][j] + tr3 * Cl[p][i];
      for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) W2[p][q] = Ckl[p][j] * Cl[q][j];
      for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) W3[p][q] = {0.0, 0.0};
      for (int p = 0; p < n; ++p) W3[j][p] = Cl[p][i];
      const cd cnm = me_rxr_over_rij(n, W1, j, j, Ci, rm[j][j], trcj[j][j]);
      const cd ct = me_rxr_ryr_over_rij(n, W2, W3, j, j, Ci, rm[j][j], trcj[j][j]);
      const cd ca = -6.0 * (tr1 + tr2) * rm[j][j] + 4.0 * cnm - 8.0 * ct;
      OO = OO - ca * (sy.q[j] / mass(j + 1));
    }
  OO = OO * (sy.q0 / mass(0));
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const cd tr1 = Cl[j][i], tr2 = tr1, tr3 = 3.0 * Cl[j][j];
      for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) W1[q][p] = Cl[q][i] * Cl[p][j] + Cl[q][j] * Cl[p][i];
      for (int p = 0; p < n; ++p) {
        W1[p][j] = W1[p][j] + Ckl[p][j] * Cl[j][i];
        W1[p][i] = W1[p][i] - Ckl[p][j] * Cl[j][i];
      }
      for (int p = 0; p < n; ++p) {
        const cd cnm = Cl[i][j] * Cl[p][j] + tr3 * Cl[p][i];
        W1[j][p] = W1[j][p] + cnm;
        W1[i][p] = W1[i][p] - cnm;
      }
      for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) W2[p][q] = Ckl[p][j] * Cl[q][j];
      for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) W3[p][q] = {0.0, 0.0};
      for (int p = 0; p < n; ++p) {
        W3[j][p] = Cl[p][i];
        W3[i][p] = -Cl[p][i];
      }
      const cd cnm = me_rxr_over_rij(n, W1, i, j, Ci, rm[i][j], trcj[i][j]);
      const cd ct = me_rxr_ryr_over_rij(n, W2, W3, i, j, Ci, rm[i][j], trcj[i][j]);
      const cd ca = -6.0 * (tr1 + tr2) * rm[i][j] + 4.0 * cnm - 8.0 * ct;
      OO = OO + ca * (sy.q[i] * sy.q[j] / (mass(i + 1) * mass(j + 1)));
    }
  out[m] = OO / 2.0;
  // Drachmanized mass-velocity (:1669-1865)
  {
    const double c0 = 1.0 / (mass(0) * mass(0));
    for (int p = 0; p < n; ++p) {
      W1[p][p] = cd{c0 + 1.0 / (mass(p + 1) * mass(p + 1)), 0.0};
      for (int q = p + 1; q < n; ++q) W1[p][q] = W1[q][p] = cd{c0, 0.0};
    }
  }
  cd tr1 = {0.0, 0.0};
  for (int p = 0; p < n; ++p) {
    for (int q = 0; q < n; ++q) {
      cd t = {0.0, 0.0};
      for (int j = 0; j < n; ++j) t = t + Cl[j][q] * W1[j][p];
      W2[q][p] = t;
    }
    tr1 = tr1 + W2[p][p];
  }
  for (int p = 0; p < n; ++p)
    for (int q = p; q < n; ++q) {
      cd t = {0.0, 0.0};
      for (int j = 0; j < n; ++j) t = t + W2[q][j] * Cl[j][p];
      W3[q][p] = W3[p][q] = t;
    }
  cd tr2 = {0.0, 0.0};
  for (int p = 0; p < n; ++p)
    for (int j = 0; j < n; ++j) tr2 = tr2 + Ci[j][p] * W3[j][p];
  cd dmv = S * E * 6.0 * (tr2 - tr1) / 4.0;
  {
    cd cnm = {0.0, 0.0};
    for (int i = 0; i < n; ++i)
      cnm = cnm + (sy.q0 * sy.q[i]) *
                      (4.0 * me_rxr_over_rij(n, W3, i, i, Ci, rm[i][i], trcj[i][i]) - 6.0 * tr1 * rm[i][i]);
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j)
        cnm = cnm + (sy.q[i] * sy.q[j]) *
                        (4.0 * me_rxr_over_rij(n, W3, i, j, Ci, rm[i][j], trcj[i][j]) - 6.0 * tr1 * rm[i][j]);
    dmv = dmv - cnm / 4.0;
  }
  auto four_traces = [&](const Mat& A3, const Mat& A4, const cd& t1, const cd& t2, const cd& t3, const cd& t4) {
    cd t5 = {0.0, 0.0};
    for (int p = 0; p < n; ++p)
      for (int j = 0; j < n; ++j) t5 = t5 + A3[p][j] * A4[j][p];
    return S * (36.0 * t3 * t4 + 24.0 * t5 - 36.0 * t1 * t4 - 36.0 * t2 * t3 + 36.0 * t1 * t2) / 8.0;
  };
  {
    tr1 = {0.0, 0.0};
    for (int p = 0; p < n; ++p) {
      cd cs = {0.0, 0.0};
      for (int q = 0; q < n; ++q) cs = cs + Ck[q][p];
      for (int q = 0; q < n; ++q) W1[q][p] = cs;
      tr1 = tr1 + cs;
    }
    tr2 = {0.0, 0.0};
    for (int p = 0; p < n; ++p) {
      const double cnp = (mass(0) + mass(p + 1)) / (mass(0) * mass(0) * mass(p + 1) * mass(p + 1));
      for (int q = 0; q < n; ++q) W2[q][p] = cnp * Cl[q][p];
      tr2 = tr2 + W2[p][p];
    }
    matmul_(n, Ci, Ck, W5);
    cd tr3 = {0.0, 0.0};
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q) {
        cd t = {0.0, 0.0};
        for (int j = 0; j < n; ++j) t = t + W5[q][j] * W1[j][p];
        W3[q][p] = t;
      }
    for (int p = 0; p < n; ++p) tr3 = tr3 + W3[p][p];
    matmul_(n, Ci, W2, W5);
    matmul_(n, W5, Cl, W4);
    const cd tr4 = trace_(n, W4);
    dmv = dmv + four_traces(W3, W4, tr1, tr2, tr3, tr4);
  }
  for (int k = 0; k < n - 1; ++k) {
    tr1 = Ck[k][k];
    for (int p = 0; p <= k; ++p)
      for (int q = 0; q < n; ++q) W2[q][p] = {0.0, 0.0};
    tr2 = {0.0, 0.0};
    for (int p = k + 1; p < n; ++p) {
      const double cnp =
          (mass(k + 1) + mass(p + 1)) / (mass(k + 1) * mass(k + 1) * mass(p + 1) * mass(p + 1));
      for (int q = 0; q < n; ++q) W2[q][p] = cnp * Cl[q][p];
      tr2 = tr2 + W2[p][p];
    }
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q) W5[q][p] = Ck[q][k] * Ck[p][k];
    matmul_(n, Ci, W5, W3);
    const cd tr3 = trace_(n, W3);
    matmul_(n, Ci, W2, W5);
    matmul_(n, W5, Cl, W4);
    const cd tr4 = trace_(n, W4);
    dmv = dmv + four_traces(W3, W4, tr1, tr2, tr3, tr4);
  }
  out[m + 1] = dmv;
  // Drachmanized Darwin (:1890-1933)
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) W1[a][b] = cd{1.0 / (mass(0) * mass(0)), 0.0};
  for (int i = 0; i < n; ++i) W1[i][i] = W1[i][i] + cd{1.0 / (mass(i + 1) * mass(i + 1)), 0.0};
  cd dd = E * me_dxd(n, W1, Ci, Cl, S) / 4.0;
  for (int i = 0; i < n; ++i)
    dd = dd - (sy.q0 * sy.q[i]) * me_dxd_over_rij(n, W1, i, i, Ci, Cl, rm[i][i], trcj[i][i]) / 4.0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      dd = dd - (sy.q[i] * sy.q[j]) * me_dxd_over_rij(n, W1, i, j, Ci, Cl, rm[i][j], trcj[i][j]) / 4.0;
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      W1[a][b] = {1.0, 0.0};
      W2[a][b] = {0.0, 0.0};
    }
  for (int j = 0; j < n; ++j)
    W2[j][j] = W2[j][j] + cd{(mass(0) + mass(j + 1)) / (mass(0) * mass(0) * mass(j + 1) * mass(j + 1)), 0.0};
  dd = dd + me_dxd_dyd(n, W1, W2, Ci, Ck, Cl, S) / 8.0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) W1[a][b] = W2[a][b] = {0.0, 0.0};
      W1[i][i] = {1.0, 0.0};
      W2[j][j] = {1.0, 0.0};
      const double cnm = (mass(i + 1) + mass(j + 1)) / (mass(i + 1) * mass(i + 1) * mass(j + 1) * mass(j + 1));
      dd = dd + cnm * me_dxd_dyd(n, W1, W2, Ci, Ck, Cl, S) / 8.0;
    }
  out[m + 2] = dd;
}

void oracle_drach_pair(int n, int T, const int* P, const int* coef, const double* M, double q0, const double* q,
                       const double* pk, const double* pl, int term, const double* masses, double E, double* out) {
  System<double> sy = make_sys<double>(n, T, P, coef, M, q0, q);
  const int m = n * (n + 1) / 2;
  std::vector<cd> o(m + 3);
  drach_pair(sy, pk, pl, P + (size_t)term * n * n, masses, E, o.data());
  for (int a = 0; a < m + 3; ++a) st(out, a, o[a]);
}

// The Drachmanized values and OO in the vector c, with the driver's
// contraction and rounding estimate (as oracle_expectation_values):
// out, err: n(n+1)/2 + 3 reals each.
int oracle_drachmanized(int n, int T, const int* P, const int* coef, const double* M, double q0, const double* q,
                        int K, const double* params, const double* c, const double* diagS, const double* masses,
                        double E, double* out, double* err, int nthreads) {
  System<double> sy = make_sys<double>(n, T, P, coef, M, q0, q);
  const int npt = n * (n + 1), nv = n * (n + 1) / 2 + 3;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const int nth = omp_get_max_threads();
  std::vector<double> acc((size_t)nth * 2 * nv, 0.0);
#pragma omp parallel
  {
    const int tid = omp_get_thread_num();
    double* s_ = acc.data() + (size_t)tid * 2 * nv;
    double* e_ = s_ + nv;
    std::vector<cd> o(nv);
#pragma omp for schedule(dynamic, 4)
    for (int i = 0; i < K; ++i)
      for (int j = 0; j <= i; ++j) {
        const double rci = c[2 * i], ici = c[2 * i + 1], rcj = c[2 * j], icj = c[2 * j + 1];
        const double factor = (i == j) ? 1.0 / std::sqrt(diagS[i] * diagS[i]) : 2.0 / std::sqrt(diagS[i] * diagS[j]);
        for (int k = 0; k < T; ++k) {
          drach_pair(sy, params + (size_t)npt * i, params + (size_t)npt * j, P + (size_t)k * n * n, masses, E,
                     o.data());
          for (int a = 0; a < nv; ++a) {
            const double ro = o[a].re, io = o[a].im;
            const double beta = (rci * rcj + ici * icj) * ro + (ici * rcj - rci * icj) * io;
            const double mu = std::fabs(rci * (rcj * ro - icj * io)) + std::fabs(rcj * (rci * ro + ici * io)) +
                              std::fabs(ici * (icj * ro + rcj * io)) + std::fabs(icj * (ici * ro - rci * io));
            s_[a] += factor * coef[k] * beta;
            e_[a] += std::fabs(factor * coef[k] * (mu + 2.0 * std::fabs(beta)));
          }
        }
      }
  }
  for (int a = 0; a < nv; ++a) {
    double sa = 0.0, ea = 0.0;
    for (int t = 0; t < nth; ++t) {
      sa += acc[(size_t)t * 2 * nv + a];
      ea += acc[(size_t)t * 2 * nv + nv + a];
    }
    out[a] = sa;
    if (err) err[a] = ea * 2.220446049250313e-16;
  }
  return 0;
}

// MV and Darwin in the vector c with the driver's contraction and rounding
// estimate (as oracle_expectation_values): out[2], err[2].
int oracle_relativistic(int n, int T, const int* P, const int* coef, const double* M, double q0, const double* q,
                        int K, const double* params, const double* c, const double* diagS, const double* masses,
                        double* out, double* err, int nthreads) {
  System<double> sy = make_sys<double>(n, T, P, coef, M, q0, q);
  const int npt = n * (n + 1);
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const int nth = omp_get_max_threads();
  std::vector<double> acc((size_t)nth * 4, 0.0);
#pragma omp parallel
  {
    const int tid = omp_get_thread_num();
    double* s_ = acc.data() + (size_t)tid * 4;
#pragma omp for schedule(dynamic, 4)
    for (int i = 0; i < K; ++i)
      for (int j = 0; j <= i; ++j) {
        const double rci = c[2 * i], ici = c[2 * i + 1], rcj = c[2 * j], icj = c[2 * j + 1];
        const double factor = (i == j) ? 1.0 / std::sqrt(diagS[i] * diagS[i]) : 2.0 / std::sqrt(diagS[i] * diagS[j]);
        for (int k = 0; k < T; ++k) {
          cd o[2];
          relat_pair(sy, params + (size_t)npt * i, params + (size_t)npt * j, P + (size_t)k * n * n, masses, o);
          for (int a = 0; a < 2; ++a) {
            const double ro = o[a].re, io = o[a].im;
            const double beta = (rci * rcj + ici * icj) * ro + (ici * rcj - rci * icj) * io;
            const double mu = std::fabs(rci * (rcj * ro - icj * io)) + std::fabs(rcj * (rci * ro + ici * io)) +
                              std::fabs(ici * (icj * ro + rcj * io)) + std::fabs(icj * (ici * ro - rci * io));
            s_[a] += factor * coef[k] * beta;
            s_[2 + a] += std::fabs(factor * coef[k] * (mu + 2.0 * std::fabs(beta)));
          }
        }
      }
  }
  for (int a = 0; a < 2; ++a) {
    double sa = 0.0, ea = 0.0;
    for (int t = 0; t < nth; ++t) {
      sa += acc[(size_t)t * 4 + a];
      ea += acc[(size_t)t * 4 + 2 + a];
    }
    out[a] = sa;
    if (err) err[a] = ea * 2.220446049250313e-16;
  }
  return 0;
}

}  // extern "C"
