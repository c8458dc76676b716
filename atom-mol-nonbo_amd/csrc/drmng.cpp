// Reverse-communication secant minimiser with a double-dogleg trust
// region: the routine the reference's optimisation drivers call to move
// the nonlinear parameters (PORT's DRMNG with the helpers it calls,
// src/dmng.f; the drivers are BasisEnlG/I, OptCycleG/I and FullOpt1G/I in
// src/workproc.f90, e.g. :4515-4534 and :4648-4731 for OptCycleG).
//
// The caller owns IV (integer workspace, LIV >= 59) and V (double
// workspace, LV >= 71 + n (n + 13) / 2); ecg_divset fills them with the
// defaults, the caller may change some, then calls ecg_drmng in a loop:
// on return IV(1) = 1 asks for f(x) (pass it as fx on the next call),
// IV(1) = 2 asks for the gradient at x, 3..8 are convergence codes, 9 and
// 10 the evaluation and iteration limits, larger values errors.  The
// workspace layout (the index of every stored quantity in IV and V) and
// the arithmetic follow src/dmng.f step by step, in the same order, so
// that a caller that saved IV and V from the reference (or passes the same
// settings) sees the same sequence of points bit for bit; the test suite
// drives this code and the reference's own routine side by side.
//
// Arrays are addressed 1-based below (IV(i) is iv[i - 1]) to keep the
// stored indices readable against the reference's layout.  Nothing here
// prints: the reference's summary printing (DITSUM) only writes its report,
// and the few workspace entries it changes are changed here too.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "../../include/ecg_drmng.h"

namespace {

#define IV_(i) iv[(i) - 1]
#define V_(i) v[(i) - 1]

// IV subscripts
constexpr int TOOBIG = 2, IVNEED = 3, VNEED = 4, MODEL = 5, NFCALL = 6, NFGCAL = 7, RADINC = 8, RESTOR = 9,
              STAGE = 10, STGLIM = 11, SWITCH = 12, XIRC = 13, COVPRT = 14, COVREQ = 15, DTYPE = 16, MXFCAL = 17,
              MXITER = 18, OUTLEV = 19, PARPRT = 20, PRUNIT = 21, SOLPRT = 22, STATPR = 23, X0PRT = 24, INITH = 25,
              INITS = 25, IRC = 29, NGCALL = 30, NITER = 31, MLSTGD = 32, KAGQT = 33, NWTSTP = 34, MODE = 35,
              NEEDHD = 36, DG = 37, OLDN = 38, PRNTIT = 39, STEP = 40, STLSTG = 41, LMAT = 42, X0 = 43,
              LASTIV = 44, LASTV = 45, NEXTIV = 46, NEXTV = 47, G0 = 48, PARSAV = 49, NVDFLT = 50, ALGSAV = 51,
              NFCOV = 52, NGCOV = 53, DTYPE0 = 54, CNVCOD = 55, RDREQ = 57, PERM = 58, VSAVE = 60, HC = 71,
              IERR = 75, IPIVOT = 76, RMAT = 78, QRTYP = 80, DRADPR = 101;
constexpr int NVSAVE = 9;
// V subscripts
constexpr int DGNORM = 1, DSTNRM = 2, DST0 = 3, GTSTEP = 4, STPPAR = 5, NREDUC = 6, PREDUC = 7, RADIUS = 8,
              RAD0 = 9, F = 10, FDIF = 11, FLSTGD = 12, F0 = 13, GTSLST = 14, PLSTGD = 15, RADFAC = 16,
              RELDX = 17, DSTSAV = 18, EPSLON = 19, PHMNFC = 20, PHMXFC = 21, DECFAC = 22, INCFAC = 23,
              RDFCMN = 24, RDFCMX = 25, TUNER1 = 26, TUNER2 = 27, TUNER3 = 28, TUNER4 = 29, TUNER5 = 30,
              AFCTOL = 31, RFCTOL = 32, XCTOL = 33, XFTOL = 34, LMAX0 = 35, LMAXS = 36, SCTOL = 37, DINIT = 38,
              DTINIT = 39, D0INIT = 40, DFAC = 41, DLTFDC = 42, ETA0 = 42, BIAS = 43, DLTFDJ = 43, GTHG = 44,
              DELTA0 = 44, FUZZ = 45, GRDFAC = 45, NWTFAC = 46, RLIMIT = 46, COSMIN = 47, HUBERC = 48,
              RSPTOL = 49, SIGMIN = 50;

// IEEE double machine constants (the reference's D1MACH(1), (2), (4)):
// smallest normal, largest finite, spacing of 1.
constexpr double ETA_MIN = 2.2250738585072014e-308;
constexpr double BIG_MAX = 1.7976931348623157e308;
constexpr double MACHEP = 2.220446049250313e-16;

double r7mdc(int k) {   // dmng.f:2425
  switch (k) {
    case 1: return ETA_MIN;
    case 2: return std::sqrt(256.0 * ETA_MIN) / 16.0;
    case 3: return MACHEP;
    case 4: return std::sqrt(MACHEP);
    case 5: return std::sqrt(BIG_MAX / 256.0) * 16.0;
    default: return BIG_MAX;
  }
}

// x . y skipping products that would underflow (dmng.f:929)
double d7tpr(int p, const double* x, const double* y) {
  double s = 0.0;
  if (p <= 0) return s;
  const double sqteta = r7mdc(2);
  for (int i = 0; i < p; ++i) {
    double t = std::fmax(std::fabs(x[i]), std::fabs(y[i]));
    if (!(t > 1.0)) {
      if (t < sqteta) continue;
      t = (x[i] / sqteta) * y[i];
      if (std::fabs(t) < sqteta) continue;
    }
    s = s + x[i] * y[i];
  }
  return s;
}

// scaled 2-norm (dmng.f:2970)
double v2nrm(int p, const double* x) {
  if (p <= 0) return 0.0;
  int i = 0;
  while (i < p && x[i] == 0.0) ++i;
  if (i == p) return 0.0;
  double scale = std::fabs(x[i]);
  if (i == p - 1) return scale;
  double t = 1.0;
  const double sqteta = r7mdc(2);
  for (int j = i + 1; j < p; ++j) {
    const double xi = std::fabs(x[j]);
    if (!(xi > scale)) {
      const double r = xi / scale;
      if (r > sqteta) t = t + r * r;
    } else {
      double r = scale / xi;
      if (r <= sqteta) r = 0.0;
      t = 1.0 + t * r * r;
      scale = xi;
    }
  }
  return scale * std::sqrt(t);
}

void v2axy(int p, double* w, double a, const double* x, const double* y) {   // w = a x + y
  for (int i = 0; i < p; ++i) w[i] = a * x[i] + y[i];
}
void v7cpy(int p, double* y, const double* x) {
  for (int i = 0; i < p; ++i) y[i] = x[i];
}
void v7scp(int p, double* y, double s) {
  for (int i = 0; i < p; ++i) y[i] = s;
}
void v7vmp(int n, double* x, const double* y, const double* z, int k) {   // x = y * z or y / z
  if (k >= 0)
    for (int i = 0; i < n; ++i) x[i] = y[i] * z[i];
  else
    for (int i = 0; i < n; ++i) x[i] = y[i] / z[i];
}

// relative difference of x and x0 in the d-scaled max norm (dmng.f:2479)
double rldst(int p, const double* d, const double* x, const double* x0) {
  double emax = 0.0, xmax = 0.0;
  for (int i = 0; i < p; ++i) {
    double t = std::fabs(d[i] * (x[i] - x0[i]));
    if (emax < t) emax = t;
    t = d[i] * (std::fabs(x[i]) + std::fabs(x0[i]));
    if (xmax < t) xmax = t;
  }
  return xmax > 0.0 ? emax / xmax : 0.0;
}

// Packed lower-triangular L (row by row, L(i,j) at i (i - 1) / 2 + j, 1-based).
// x = L^{-T} y (dmng.f:1336); x may alias y.
void l7itv(int n, double* x, const double* l, const double* y) {
  for (int i = 0; i < n; ++i) x[i] = y[i];
  int i0 = n * (n + 1) / 2;
  for (int ii = 1; ii <= n; ++ii) {
    const int i = n + 1 - ii;
    const double xi = x[i - 1] / l[i0 - 1];
    x[i - 1] = xi;
    if (i <= 1) return;
    i0 = i0 - i;
    if (xi == 0.0) continue;
    for (int j = 1; j <= i - 1; ++j) x[j - 1] = x[j - 1] - xi * l[i0 + j - 1];
  }
}

// x = L^{-1} y (dmng.f:1372)
void l7ivm(int n, double* x, const double* l, const double* y) {
  int k = 1;
  for (; k <= n; ++k) {
    if (y[k - 1] != 0.0) break;
    x[k - 1] = 0.0;
  }
  if (k > n) return;
  int j = k * (k + 1) / 2;
  x[k - 1] = y[k - 1] / l[j - 1];
  if (k >= n) return;
  for (int i = k + 1; i <= n; ++i) {
    const double t = d7tpr(i - 1, l + j, x);
    j = j + i;
    x[i - 1] = (y[i - 1] - t) / l[j - 1];
  }
}

// x = L^T y (dmng.f:1407)
void l7tvm(int n, double* x, const double* l, const double* y) {
  int i0 = 0;
  for (int i = 1; i <= n; ++i) {
    const double yi = y[i - 1];
    x[i - 1] = 0.0;
    for (int j = 1; j <= i; ++j) x[j - 1] = x[j - 1] + yi * l[i0 + j - 1];
    i0 = i0 + i;
  }
}

// x = L y (dmng.f:1579); x may alias y
void l7vml(int n, double* x, const double* l, const double* y) {
  int i0 = n * (n + 1) / 2;
  for (int ii = 1; ii <= n; ++ii) {
    const int i = n + 1 - ii;
    i0 = i0 - i;
    double t = 0.0;
    for (int j = 1; j <= i; ++j) t = t + l[i0 + j - 1] * y[j - 1];
    x[i - 1] = t;
  }
}

// Cholesky factor of (I + z w^T)(L L^T)(I + w z^T): lplus (may be l) from
// l, w, z; beta, gamma, lambda are scratch (dmng.f:1437, Goldfarb's
// recurrence 3)
void l7upd(double* beta, double* gamma, const double* l, double* lambda, double* lplus, int n, double* w, double* z) {
  double nu = 1.0, eta = 0.0;
  if (n > 1) {
    double s = 0.0;
    for (int i = 1; i <= n - 1; ++i) {
      const int j = n - i;
      s = s + w[j] * w[j];
      lambda[j - 1] = s;
    }
    for (int j = 1; j <= n - 1; ++j) {
      const double wj = w[j - 1];
      const double a = nu * z[j - 1] - eta * wj;
      const double theta = 1.0 + a * wj;
      s = a * lambda[j - 1];
      double lj = std::sqrt(theta * theta + a * s);
      if (theta > 0.0) lj = -lj;
      lambda[j - 1] = lj;
      const double b = theta * wj + s;
      gamma[j - 1] = b * nu / lj;
      beta[j - 1] = (a - b * eta) / lj;
      nu = -nu / lj;
      eta = -(eta + (a * a) / (theta - lj)) / lj;
    }
  }
  lambda[n - 1] = 1.0 + (nu * z[n - 1] - eta * w[n - 1]) * w[n - 1];
  int jj = n * (n + 1) / 2;
  for (int k = 1; k <= n; ++k) {
    const int j = n + 1 - k;
    const double lj = lambda[j - 1];
    const double ljj = l[jj - 1];
    lplus[jj - 1] = lj * ljj;
    const double wj = w[j - 1];
    w[j - 1] = ljj * wj;
    const double zj = z[j - 1];
    z[j - 1] = ljj * zj;
    if (k != 1) {
      const double bj = beta[j - 1], gj = gamma[j - 1];
      int ij = jj + j;
      for (int i = j + 1; i <= n; ++i) {
        const double lij = l[ij - 1];
        lplus[ij - 1] = lj * lij + bj * w[i - 1] + gj * z[i - 1];
        w[i - 1] = w[i - 1] + lij * wj;
        z[i - 1] = z[i - 1] + lij * zj;
        ij = ij + i;
      }
    }
    jj = jj - j;
  }
}

// w, z of the BFGS update of L L^T for step s and gradient change y,
// damped (Powell) when y.s < 0.1 s^T L L^T s (dmng.f:3179)
void w7zbf(const double* l, int n, const double* s, double* w, const double* y, double* z) {
  const double eps = 0.1;
  l7tvm(n, w, l, s);
  const double shs = d7tpr(n, w, w);
  const double ys = d7tpr(n, y, s);
  double cy, cs;
  if (ys >= eps * shs) {
    cy = 1.0 / (std::sqrt(ys) * std::sqrt(shs));
    cs = 1.0 / shs;
  } else {
    const double theta = (1.0 - eps) * shs / (shs - ys);
    const double epsrt = std::sqrt(eps);
    cy = theta / (shs * epsrt);
    cs = (1.0 + (theta - 1.0) / epsrt) / shs;
  }
  l7ivm(n, z, l, y);
  for (int i = 0; i < n; ++i) z[i] = cy * z[i] - cs * w[i];
}

// double-dogleg step of length V(RADIUS) from the scaled gradient dig and
// the Newton step (dmng.f:728)
void d7dog(const double* dig, int n, const double* nwtstp, double* step, double* v) {
  const double nwtnrm = V_(DST0);
  double rlambd = 1.0;
  if (nwtnrm > 0.0) rlambd = V_(RADIUS) / nwtnrm;
  const double gnorm = V_(DGNORM);
  const double ghinvg = 2.0 * V_(NREDUC);
  V_(GRDFAC) = 0.0;
  V_(NWTFAC) = 0.0;
  if (!(rlambd < 1.0)) {   // full Newton step
    V_(STPPAR) = 0.0;
    V_(DSTNRM) = nwtnrm;
    V_(GTSTEP) = -ghinvg;
    V_(PREDUC) = V_(NREDUC);
    V_(NWTFAC) = -1.0;
    for (int i = 0; i < n; ++i) step[i] = -nwtstp[i];
    return;
  }
  V_(DSTNRM) = V_(RADIUS);
  double cfact = gnorm / V_(GTHG);
  cfact = cfact * cfact;
  const double cnorm = gnorm * cfact;
  const double relax = 1.0 - V_(BIAS) * (1.0 - gnorm * cnorm / ghinvg);
  if (!(rlambd < relax)) {   // scaled Newton step
    V_(STPPAR) = 1.0 - (rlambd - relax) / (1.0 - relax);
    const double t = -rlambd;
    V_(GTSTEP) = t * ghinvg;
    V_(PREDUC) = rlambd * (1.0 - 0.5 * rlambd) * ghinvg;
    V_(NWTFAC) = t;
    for (int i = 0; i < n; ++i) step[i] = t * nwtstp[i];
    return;
  }
  if (!(cnorm < V_(RADIUS))) {   // scaled Cauchy step
    const double t = -V_(RADIUS) / gnorm;
    V_(GRDFAC) = t;
    V_(STPPAR) = 1.0 + cnorm / V_(RADIUS);
    V_(GTSTEP) = -V_(RADIUS) * gnorm;
    double q = V_(GTHG) / gnorm;
    q = q * q;
    V_(PREDUC) = V_(RADIUS) * (gnorm - 0.5 * V_(RADIUS) * q);
    for (int i = 0; i < n; ++i) step[i] = t * dig[i];
    return;
  }
  // dogleg between the relaxed Newton and Cauchy points
  const double ctrnwt = cfact * relax * ghinvg / gnorm;
  const double cf2 = cfact * cfact;
  double t1 = ctrnwt - gnorm * cf2;
  double t2 = V_(RADIUS) * (V_(RADIUS) / gnorm) - gnorm * cf2;
  double t = relax * nwtnrm;
  const double femnsq = (t / gnorm) * t - ctrnwt - t1;
  t = t2 / (t1 + std::sqrt(t1 * t1 + femnsq * t2));
  t1 = (t - 1.0) * cfact;
  V_(GRDFAC) = t1;
  t2 = -t * relax;
  V_(NWTFAC) = t2;
  V_(STPPAR) = 2.0 - t;
  V_(GTSTEP) = t1 * (gnorm * gnorm) + t2 * ghinvg;
  const double q = V_(GTHG) * t1;
  V_(PREDUC) = -t1 * gnorm * ((t2 + 1.0) * gnorm) - t2 * (1.0 + 0.5 * t2) * ghinvg - 0.5 * (q * q);
  for (int i = 0; i < n; ++i) step[i] = t1 * dig[i] + t2 * nwtstp[i];
}

// Default tolerances and tuning constants (dmng.f:3044), written at v
// (1-based V subscripts relative to v).
void v7dfl(int alg, double* v) {
  const double machep = r7mdc(3);
  V_(AFCTOL) = 1.0e-20;
  if (machep > 1.0e-10) V_(AFCTOL) = machep * machep;
  V_(DECFAC) = 0.5;
  const double sqteps = r7mdc(4);
  V_(DFAC) = 0.6;
  V_(DTINIT) = 1.0e-6;
  const double mepcrt = std::pow(machep, 1.0 / 3.0);
  V_(D0INIT) = 1.0;
  V_(EPSLON) = 0.1;
  V_(INCFAC) = 2.0;
  V_(LMAX0) = 1.0;
  V_(LMAXS) = 1.0;
  V_(PHMNFC) = -0.1;
  V_(PHMXFC) = 0.1;
  V_(RDFCMN) = 0.1;
  V_(RDFCMX) = 4.0;
  V_(RFCTOL) = std::fmax(1.0e-10, mepcrt * mepcrt);
  V_(SCTOL) = V_(RFCTOL);
  V_(TUNER1) = 0.1;
  V_(TUNER2) = 1.0e-4;
  V_(TUNER3) = 0.75;
  V_(TUNER4) = 0.5;
  V_(TUNER5) = 0.75;
  V_(XCTOL) = sqteps;
  V_(XFTOL) = 1.0e2 * machep;
  if (alg < 2) {
    V_(COSMIN) = std::fmax(1.0e-6, 1.0e2 * machep);
    V_(DINIT) = 0.0;
    V_(DELTA0) = sqteps;
    V_(DLTFDC) = mepcrt;
    V_(DLTFDJ) = sqteps;
    V_(FUZZ) = 1.5;
    V_(HUBERC) = 0.7;
    V_(RLIMIT) = r7mdc(5);
    V_(RSPTOL) = 1.0e-3;
    V_(SIGMIN) = 1.0e-4;
  } else {
    V_(BIAS) = 0.8;
    V_(DINIT) = -1.0;
    V_(ETA0) = 1.0e3 * machep;
  }
}

constexpr int MINIV[4] = {82, 59, 103, 103};
constexpr int MINV[4] = {98, 71, 101, 85};

// State changes of the reference's iteration summary (dmng.f:967); its
// printed report is not produced.
void itsum(int* iv, double* v) {
  if (IV_(PRUNIT) == 0) return;
  int iv1 = IV_(1);
  if (iv1 > 62) iv1 -= 51;
  const int ol = IV_(OUTLEV);
  if (iv1 < 2 || iv1 > 15) return;
  bool zero_block = false;
  if (iv1 < 12) {
    if (iv1 == 2 && IV_(NITER) == 0) {
      zero_block = true;
    } else if (ol != 0 && !(iv1 >= 10 && IV_(PRNTIT) == 0)) {
      bool report = true;
      if (iv1 <= 2) {
        IV_(PRNTIT) = IV_(PRNTIT) + 1;
        if (IV_(PRNTIT) < std::abs(ol)) return;
      }
      if (report) {
        IV_(PRNTIT) = 0;
        IV_(NEEDHD) = 0;
      }
    }
  }
  if (!zero_block) {
    if (iv1 <= 2) return;
    const int i = IV_(STATPR);
    bool to_sol = (i == -1) || (i + iv1 < 0);
    if (!to_sol) {
      if (iv1 <= 2) return;
      if (iv1 <= 11) {
        IV_(NEEDHD) = 1;
        to_sol = true;
      } else if (iv1 == 12) {
        zero_block = true;
      } else if (iv1 == 13 || iv1 == 15) {
        return;
      } else {   // 14
        if (IV_(NITER) > 0) to_sol = true;
        else zero_block = true;
      }
    }
    if (to_sol && !zero_block) {
      if (IV_(SOLPRT) == 0) return;
      IV_(NEEDHD) = 1;
      return;
    }
  }
  // summary of the starting point: the step statistics are cleared
  V_(DSTNRM) = 0.0;
  V_(FDIF) = 0.0;
  V_(NREDUC) = 0.0;
  V_(PREDUC) = 0.0;
  V_(RELDX) = 0.0;
  if (iv1 >= 12) return;
  IV_(NEEDHD) = 0;
  IV_(PRNTIT) = 0;
}

// Range table of the tunable V entries checked by the parameter check
// (dmng.f:2092, VM / VX): entry i covers V(EPSLON + i) for the first 23,
// then V(ETA0) and V(BIAS) for the secant algorithms.
struct Range {
  double lo, hi;
};
void ranges(Range* r) {
  const double tiny = r7mdc(1), machep = r7mdc(3), big = r7mdc(6);
  const Range tab[34] = {
      {1.0e-3, 0.9},   {-0.99, -1.0e-3}, {1.0e-3, 1.0e1}, {1.0e-2, 0.8}, {1.2, 1.0e2},   {1.0e-2, 0.8},
      {1.2, 1.0e2},    {0.0, 0.5},       {0.0, 0.5},      {1.0e-3, 1.0}, {-1.0, 1.0},   {machep, big},
      {0.0, big},      {machep, 0.1},    {0.0, 1.0},      {0.0, 1.0},    {tiny, big},   {tiny, big},
      {0.0, 1.0},      {-10.0, big},     {0.0, big},      {0.0, big},    {0.0, 1.0},    {machep, 1.0},
      {machep, 1.0},   {machep, 1.0},    {1.01, 1.0e10},  {1.0e10, r7mdc(5)}, {machep, 1.0}, {0.0, big},
      {0.0, 1.0},      {0.0, 1.0},       {machep, 1.0},   {0.0, 1.0}};
  for (int i = 0; i < 34; ++i) r[i] = tab[i];
}

// Check of the workspace sizes, the tunables' ranges and the scale vector
// (dmng.f:2092); sets IV(1) to an error code on failure.
void parck(int alg, const double* d, int* iv, int liv, int lv, int n, double* v) {
  if (ALGSAV <= liv && alg != IV_(ALGSAV)) {
    IV_(1) = 67;
    return;
  }
  if (alg < 1 || alg > 4) {
    IV_(1) = 67;
    return;
  }
  int miv1 = MINIV[alg - 1];
  if (IV_(1) == 15) {
    if (LASTIV <= liv) IV_(LASTIV) = miv1;
    if (LASTV <= liv) IV_(LASTV) = 0;
    return;
  }
  const int alg1 = (alg - 1) % 2 + 1;
  if (IV_(1) == 0) ecg_divset(alg, iv, liv, lv, v);
  int iv1 = IV_(1);
  if (iv1 == 13 || iv1 == 12) {
    if (PERM <= liv) miv1 = std::max(miv1, IV_(PERM) - 1);
    int miv2 = miv1;
    if (IVNEED <= liv) miv2 = miv1 + std::max(IV_(IVNEED), 0);
    if (LASTIV <= liv) IV_(LASTIV) = miv2;
    if (liv < miv1) {
      IV_(1) = 15;
      return;
    }
    IV_(IVNEED) = 0;
    IV_(LASTV) = std::max(IV_(VNEED), 0) + IV_(LMAT) - 1;
    IV_(VNEED) = 0;
    if (liv < miv2) {
      IV_(1) = 15;
      if (liv >= miv1 && lv < IV_(LASTV)) IV_(1) = 16;
      return;
    }
    if (lv < IV_(LASTV)) {
      IV_(1) = 16;
      return;
    }
  }
  if (iv1 >= 12 && iv1 <= 14) {
    if (n < 1) {
      IV_(1) = 81;
      return;
    }
    if (iv1 != 14) {
      IV_(NEXTIV) = IV_(PERM);
      IV_(NEXTV) = IV_(LMAT);
    }
    if (iv1 == 13) return;
    const int k = IV_(PARSAV) - EPSLON;
    v7dfl(alg1, v + k);
    IV_(DTYPE0) = 2 - alg1;
    IV_(OLDN) = n;
  } else {
    if (n != IV_(OLDN)) {
      IV_(1) = 17;
      return;
    }
    if (!(iv1 <= 11 && iv1 >= 1)) {
      IV_(1) = 80;
      return;
    }
  }
  if (iv1 == 14) iv1 = 12;
  Range r[34];
  ranges(r);
  int m = 0;
  const int jlim = alg1 == 1 ? 0 : 24;
  const int ndfalt = alg1 == 1 ? 32 : 25;
  for (int l = 1, i = 1, k = EPSLON; l <= ndfalt; ++l) {
    const double vk = V_(k);
    if (!(vk >= r[i - 1].lo && vk <= r[i - 1].hi)) m = k;
    ++k;
    ++i;
    if (i == jlim) i = 33;
  }
  if (IV_(NVDFLT) != ndfalt) {
    IV_(1) = 51;
    return;
  }
  if (!((IV_(DTYPE) > 0 || V_(DINIT) > 0.0) && iv1 == 12))
    for (int i = 0; i < n; ++i)
      if (!(d[i] > 0.0)) m = 18;
  if (m != 0) {
    IV_(1) = m;
    return;
  }
  if (IV_(PRUNIT) == 0 || IV_(PARPRT) == 0) return;
  IV_(DTYPE0) = IV_(DTYPE);
  v7cpy(IV_(NVDFLT), &V_(IV_(PARSAV)), &V_(EPSLON));
}

// Step acceptance and trust-region radius update (dmng.f:210).
void a7sst(int* iv, double* v) {
  int nfc = IV_(NFCALL);
  IV_(SWITCH) = 0;
  IV_(RESTOR) = 0;
  double rfac1 = 1.0;
  bool goodx = true;
  int i = IV_(IRC);
  enum {
    L10, L20, L30, L40, L50, L60, L70, L80, L110, L120, L130, L140, L150, L160, L170, L180, L190, L200, L210,
    L220, L230, L240, L250, L260, L270, L280, L290, DONE
  } lab;
  static const decltype(L10) entry[12] = {L20, L30, L10, L10, L40, L280, L220, L220, L220, L220, L220, L170};
  if (i >= 1 && i <= 12) {
    lab = entry[i - 1];
  } else {
    IV_(IRC) = 13;
    return;
  }
  double emax = 0.0, emaxs = 0.0;
  for (;;) {
    switch (lab) {
      case L10:
        IV_(STAGE) = 1;
        IV_(RADINC) = 0;
        V_(FLSTGD) = V_(F0);
        if (IV_(TOOBIG) == 0) { lab = L110; break; }
        IV_(STAGE) = -1;
        IV_(XIRC) = i;
        lab = L60;
        break;
      case L20:
        if (IV_(MODEL) != IV_(MLSTGD)) { lab = L30; break; }
        IV_(STAGE) = IV_(STGLIM);
        IV_(RADINC) = -1;
        lab = L110;
        break;
      case L30:
        IV_(STAGE) = IV_(STAGE) + 1;
        lab = L40;
        break;
      case L40:
        if (IV_(STAGE) > 0) { lab = L50; break; }
        if (IV_(TOOBIG) != 0) { lab = L60; break; }
        IV_(STAGE) = -IV_(STAGE);
        i = IV_(XIRC);
        switch (i) {
          case 1: lab = L20; break;
          case 2: lab = L30; break;
          case 3: case 4: lab = L110; break;
          case 5: lab = L70; break;
          default: lab = L50; break;
        }
        break;
      case L50:
        if (IV_(TOOBIG) == 0) { lab = L70; break; }
        IV_(TOOBIG) = 0;
        if (IV_(RADINC) > 0) { lab = L80; break; }
        IV_(STAGE) = -IV_(STAGE);
        IV_(XIRC) = IV_(IRC);
        lab = L60;
        break;
      case L60:
        IV_(TOOBIG) = 0;
        V_(RADFAC) = V_(DECFAC);
        IV_(RADINC) = IV_(RADINC) - 1;
        IV_(IRC) = 5;
        IV_(RESTOR) = 1;
        V_(F) = V_(FLSTGD);
        return;
      case L70:
        if (V_(F) < V_(FLSTGD)) { lab = L110; break; }
        if (IV_(MODEL) != IV_(MLSTGD)) {
          IV_(MODEL) = IV_(MLSTGD);
          IV_(SWITCH) = 1;
        }
        lab = L80;
        break;
      case L80:
        if (V_(FLSTGD) >= V_(F0)) { lab = L110; break; }
        IV_(RESTOR) = 1;
        V_(F) = V_(FLSTGD);
        V_(PREDUC) = V_(PLSTGD);
        V_(GTSTEP) = V_(GTSLST);
        if (IV_(SWITCH) == 0) rfac1 = V_(DSTNRM) / V_(DSTSAV);
        V_(DSTNRM) = V_(DSTSAV);
        nfc = IV_(NFGCAL);
        goodx = false;
        lab = L110;
        break;
      case L110:
        V_(FDIF) = V_(F0) - V_(F);
        if (V_(FDIF) > V_(TUNER2) * V_(PREDUC)) { lab = L140; break; }
        if (IV_(RADINC) > 0) { lab = L140; break; }
        if (V_(F) < V_(F0)) { lab = L120; break; }
        IV_(MLSTGD) = IV_(MODEL);
        V_(FLSTGD) = V_(F);
        V_(F) = V_(F0);
        IV_(RESTOR) = 1;
        lab = L130;
        break;
      case L120:
        IV_(NFGCAL) = nfc;
        lab = L130;
        break;
      case L130:
        IV_(IRC) = 1;
        if (IV_(STAGE) < IV_(STGLIM)) { lab = L160; break; }
        IV_(IRC) = 5;
        IV_(RADINC) = IV_(RADINC) - 1;
        lab = L160;
        break;
      case L140:
        IV_(NFGCAL) = nfc;
        rfac1 = 1.0;
        V_(DSTSAV) = V_(DSTNRM);
        if (V_(FDIF) > V_(PREDUC) * V_(TUNER1)) { lab = L190; break; }
        if (IV_(STAGE) >= IV_(STGLIM)) { lab = L150; break; }
        IV_(IRC) = 2;
        lab = L160;
        break;
      case L150:
        IV_(IRC) = 4;
        lab = L160;
        break;
      case L160: {
        IV_(XIRC) = IV_(IRC);
        const double em = V_(GTSTEP) + V_(FDIF);
        V_(RADFAC) = 0.5 * rfac1;
        if (em < V_(GTSTEP)) V_(RADFAC) = rfac1 * std::fmax(V_(RDFCMN), 0.5 * V_(GTSTEP) / em);
        lab = L170;
        break;
      }
      case L170:
        if (V_(RELDX) <= V_(XFTOL)) { lab = L180; break; }
        IV_(IRC) = IV_(XIRC);
        if (V_(F) < V_(F0)) { lab = L200; break; }
        lab = L230;
        break;
      case L180:
        IV_(IRC) = 12;
        lab = L240;
        break;
      case L190: {
        if (V_(FDIF) < (-V_(TUNER3) * V_(GTSTEP))) { lab = L210; break; }
        if (IV_(RADINC) < 0) { lab = L210; break; }
        if (IV_(RESTOR) == 1) { lab = L210; break; }
        V_(RADFAC) = V_(RDFCMX);
        const double gts = V_(GTSTEP);
        if (V_(FDIF) < (0.5 / V_(RADFAC) - 1.0) * gts) V_(RADFAC) = std::fmax(V_(INCFAC), 0.5 * gts / (gts + V_(FDIF)));
        IV_(IRC) = 4;
        if (V_(STPPAR) == 0.0) { lab = L230; break; }
        if (V_(DST0) >= 0.0 && (V_(DST0) < 2.0 * V_(DSTNRM) || V_(NREDUC) < 1.2 * V_(FDIF))) { lab = L230; break; }
        IV_(IRC) = 5;
        IV_(RADINC) = IV_(RADINC) + 1;
        lab = L200;
        break;
      }
      case L200:
        V_(FLSTGD) = V_(F);
        IV_(MLSTGD) = IV_(MODEL);
        if (IV_(RESTOR) != 1) IV_(RESTOR) = 2;
        V_(DSTSAV) = V_(DSTNRM);
        IV_(NFGCAL) = nfc;
        V_(PLSTGD) = V_(PREDUC);
        V_(GTSLST) = V_(GTSTEP);
        lab = L230;
        break;
      case L210:
        V_(RADFAC) = 1.0;
        IV_(IRC) = 3;
        lab = L230;
        break;
      case L220:
        IV_(IRC) = IV_(XIRC);
        if (V_(DSTSAV) >= 0.0) { lab = L240; break; }
        IV_(IRC) = 12;
        lab = L240;
        break;
      case L230:
        IV_(XIRC) = IV_(IRC);
        lab = L240;
        break;
      case L240:
        if (IV_(RESTOR) == 1 && V_(FLSTGD) < V_(F0)) IV_(RESTOR) = 3;
        if (std::fabs(V_(F)) < V_(AFCTOL)) IV_(IRC) = 10;
        if (0.5 * V_(FDIF) > V_(PREDUC)) return;
        emax = V_(RFCTOL) * std::fabs(V_(F0));
        emaxs = V_(SCTOL) * std::fabs(V_(F0));
        if (V_(PREDUC) <= emaxs && (V_(DSTNRM) > V_(LMAXS) || V_(STPPAR) == 0.0)) IV_(IRC) = 11;
        if (!(V_(DST0) < 0.0)) {
          int c = 0;
          if ((V_(NREDUC) > 0.0 && V_(NREDUC) <= emax) || (V_(NREDUC) == 0.0 && V_(PREDUC) == 0.0)) c = 2;
          if (V_(STPPAR) == 0.0 && V_(RELDX) <= V_(XCTOL) && goodx) c = c + 1;
          if (c > 0) IV_(IRC) = c + 6;
        }
        lab = L250;
        break;
      case L250:
        if (IV_(IRC) > 5 && IV_(IRC) != 12) return;
        if (V_(STPPAR) == 0.0) return;
        if (V_(DSTNRM) > V_(LMAXS)) { lab = L260; break; }
        if (V_(PREDUC) >= emaxs) return;
        if (V_(DST0) <= 0.0) { lab = L270; break; }
        if (0.5 * V_(DST0) <= V_(LMAXS)) return;
        lab = L270;
        break;
      case L260: {
        if (0.5 * V_(DSTNRM) <= V_(LMAXS)) return;
        const double xmax = V_(LMAXS) / V_(DSTNRM);
        if (xmax * (2.0 - xmax) * V_(PREDUC) >= emaxs) return;
        lab = L270;
        break;
      }
      case L270: {
        if (V_(NREDUC) < 0.0) { lab = L290; break; }
        V_(GTSLST) = V_(GTSTEP);
        V_(DSTSAV) = V_(DSTNRM);
        if (IV_(IRC) == 12) V_(DSTSAV) = -V_(DSTSAV);
        V_(PLSTGD) = V_(PREDUC);
        const int r = IV_(RESTOR);
        IV_(RESTOR) = 2;
        if (r == 3) IV_(RESTOR) = 0;
        IV_(IRC) = 6;
        return;
      }
      case L280:
        V_(GTSTEP) = V_(GTSLST);
        V_(DSTNRM) = std::fabs(V_(DSTSAV));
        IV_(IRC) = IV_(XIRC);
        if (V_(DSTSAV) <= 0.0) IV_(IRC) = 12;
        V_(NREDUC) = -V_(PREDUC);
        V_(PREDUC) = V_(PLSTGD);
        IV_(RESTOR) = 3;
        lab = L290;
        break;
      case L290:
        if (-V_(NREDUC) <= V_(SCTOL) * std::fabs(V_(F0))) IV_(IRC) = 11;
        return;
      case DONE:
        return;
    }
  }
}

}  // namespace

extern "C" {

void ecg_divset(int alg, int* iv, int liv, int lv, double* v) {
  if (PRUNIT <= liv) IV_(PRUNIT) = 6;   // the reference's standard output unit (I1MACH(2))
  if (ALGSAV <= liv) IV_(ALGSAV) = alg;
  if (alg < 1 || alg > 4) {
    IV_(1) = 67;
    return;
  }
  const int miv = MINIV[alg - 1];
  if (liv < miv) {
    IV_(1) = 15;
    return;
  }
  const int mv = MINV[alg - 1];
  if (lv < mv) {
    IV_(1) = 16;
    return;
  }
  const int alg1 = (alg - 1) % 2 + 1;
  v7dfl(alg1, v);
  IV_(1) = 12;
  if (alg > 2) IV_(DRADPR) = 1;
  IV_(IVNEED) = 0;
  IV_(LASTIV) = miv;
  IV_(LASTV) = mv;
  IV_(LMAT) = mv + 1;
  IV_(MXFCAL) = 200;
  IV_(MXITER) = 150;
  IV_(OUTLEV) = 1;
  IV_(PARPRT) = 1;
  IV_(PERM) = miv + 1;
  IV_(SOLPRT) = 1;
  IV_(STATPR) = 1;
  IV_(VNEED) = 0;
  IV_(X0PRT) = 1;
  if (alg1 < 2) {
    IV_(COVPRT) = 3;
    IV_(COVREQ) = 1;
    IV_(DTYPE) = 1;
    IV_(HC) = 0;
    IV_(IERR) = 0;
    IV_(INITS) = 0;
    IV_(IPIVOT) = 0;
    IV_(NVDFLT) = 32;
    IV_(VSAVE) = 58;
    if (alg > 2) IV_(VSAVE) = IV_(VSAVE) + 3;
    IV_(PARSAV) = IV_(VSAVE) + NVSAVE;
    IV_(QRTYP) = 1;
    IV_(RDREQ) = 3;
    IV_(RMAT) = 0;
    return;
  }
  IV_(DTYPE) = 0;
  IV_(INITH) = 1;
  IV_(NFCOV) = 0;
  IV_(NGCOV) = 0;
  IV_(NVDFLT) = 25;
  IV_(PARSAV) = 47;
  if (alg > 2) IV_(PARSAV) = 61;
}

void ecg_drmng(double* d, double fx, const double* g, int* iv, int liv, int lv, int n, double* v, double* x) {
  enum {
    L10, L20, L50, L60, L70, L80, L90, L100, L110, L120, L130, L140, L150, L160, L170, L180, L190, L200, L210,
    L220, L230, L240, L250, L260, L270, L280, L290, L300, L310, L320, L330, L340, L350
  } lab;
  int rstrst = 0;
  const int i1 = IV_(1);
  if (i1 == 1) {
    lab = L50;
  } else if (i1 == 2) {
    lab = L60;
  } else {
    if (IV_(1) == 0) ecg_divset(2, iv, liv, lv, v);
    if (IV_(1) == 12 || IV_(1) == 13) IV_(VNEED) = IV_(VNEED) + n * (n + 13) / 2;
    parck(2, d, iv, liv, lv, n, v);
    const int i = IV_(1) - 2;
    if (i > 12) return;
    switch (i) {
      case 1: case 2: case 3: case 4: case 5: case 6: lab = L190; break;
      case 7: case 9: lab = L120; break;
      case 8: lab = L90; break;
      case 12: lab = L20; break;
      default: lab = L10; break;   // 10, 11 (and out of range: the next statement)
    }
  }
  for (;;) {
    switch (lab) {
      case L10: {
        const int l = IV_(LMAT);
        IV_(X0) = l + n * (n + 1) / 2;
        IV_(STEP) = IV_(X0) + n;
        IV_(STLSTG) = IV_(STEP) + n;
        IV_(G0) = IV_(STLSTG) + n;
        IV_(NWTSTP) = IV_(G0) + n;
        IV_(DG) = IV_(NWTSTP) + n;
        IV_(NEXTV) = IV_(DG) + n;
        if (IV_(1) != 13) { lab = L20; break; }
        IV_(1) = 14;
        return;
      }
      case L20:
        IV_(NITER) = 0;
        IV_(NFCALL) = 1;
        IV_(NGCALL) = 1;
        IV_(NFGCAL) = 1;
        IV_(MODE) = -1;
        IV_(MODEL) = 1;
        IV_(STGLIM) = 1;
        IV_(TOOBIG) = 0;
        IV_(CNVCOD) = 0;
        IV_(RADINC) = 0;
        V_(RAD0) = 0.0;
        if (V_(DINIT) >= 0.0) v7scp(n, d, V_(DINIT));
        if (IV_(INITH) == 1) {   // initial L = diag(d)
          const int l = IV_(LMAT);
          v7scp(n * (n + 1) / 2, &V_(l), 0.0);
          int k = l - 1;
          for (int i = 1; i <= n; ++i) {
            k = k + i;
            double t = d[i - 1];
            if (t <= 0.0) t = 1.0;
            V_(k) = t;
          }
        }
        IV_(1) = 1;
        return;
      case L50:   // f(x) returned
        V_(F) = fx;
        if (IV_(MODE) >= 0) { lab = L190; break; }
        V_(F0) = fx;
        IV_(1) = 2;
        if (IV_(TOOBIG) == 0) return;
        IV_(1) = 63;
        lab = L350;
        break;
      case L60:   // gradient returned
        if (IV_(TOOBIG) != 0) {
          IV_(1) = 65;
          lab = L350;
          break;
        }
        lab = L70;
        break;
      case L70: {
        const int dg1 = IV_(DG);
        v7vmp(n, &V_(dg1), g, d, -1);
        V_(DGNORM) = v2nrm(n, &V_(dg1));
        if (IV_(CNVCOD) != 0) { lab = L340; break; }
        if (IV_(MODE) == 0) { lab = L300; break; }
        V_(RADIUS) = V_(LMAX0);
        IV_(MODE) = 0;
        lab = L80;
        break;
      }
      case L80:
        itsum(iv, v);
        lab = L90;
        break;
      case L90: {
        const int k = IV_(NITER);
        if (k < IV_(MXITER)) {
          IV_(NITER) = k + 1;
          if (k > 0) V_(RADIUS) = V_(RADFAC) * V_(DSTNRM);
          lab = L100;
          break;
        }
        IV_(1) = 10;
        lab = L350;
        break;
      }
      case L100: {   // start of an iteration: save x and g
        const int g01 = IV_(G0), x01 = IV_(X0);
        V_(F0) = V_(F);
        IV_(IRC) = 4;
        IV_(KAGQT) = -1;
        v7cpy(n, &V_(x01), x);
        v7cpy(n, &V_(g01), g);
        lab = L110;
        break;
      }
      case L110:   // no interruption requested by the caller
        lab = L130;
        break;
      case L120: {
        if (V_(F) >= V_(F0)) { lab = L130; break; }
        V_(RADFAC) = 1.0;
        const int k = IV_(NITER);
        IV_(NITER) = k + 1;
        if (k > 0) V_(RADIUS) = V_(RADFAC) * V_(DSTNRM);
        lab = L100;
        break;
      }
      case L130:
        if (IV_(NFCALL) < IV_(MXFCAL)) { lab = L150; break; }
        IV_(1) = 9;
        lab = L140;
        break;
      case L140:
        if (V_(F) >= V_(F0)) { lab = L350; break; }
        IV_(CNVCOD) = IV_(1);
        lab = L290;
        break;
      case L150: {   // compute the step
        const int step1 = IV_(STEP), dg1 = IV_(DG), nwtst1 = IV_(NWTSTP);
        if (IV_(KAGQT) < 0) {
          const int l = IV_(LMAT);
          l7ivm(n, &V_(nwtst1), &V_(l), g);
          V_(NREDUC) = 0.5 * d7tpr(n, &V_(nwtst1), &V_(nwtst1));
          l7itv(n, &V_(nwtst1), &V_(l), &V_(nwtst1));
          v7vmp(n, &V_(step1), &V_(nwtst1), d, 1);
          V_(DST0) = v2nrm(n, &V_(step1));
          v7vmp(n, &V_(dg1), &V_(dg1), d, -1);
          l7tvm(n, &V_(step1), &V_(l), &V_(dg1));
          V_(GTHG) = v2nrm(n, &V_(step1));
          IV_(KAGQT) = 0;
        }
        lab = L160;
        break;
      }
      case L160: {
        const int step1 = IV_(STEP), dg1 = IV_(DG), nwtst1 = IV_(NWTSTP);
        d7dog(&V_(dg1), n, &V_(nwtst1), &V_(step1), v);
        if (IV_(IRC) != 6) { lab = L170; break; }
        if (IV_(RESTOR) != 2) { lab = L190; break; }
        rstrst = 2;
        lab = L200;
        break;
      }
      case L170:
        IV_(TOOBIG) = 0;
        if (V_(DSTNRM) <= 0.0) { lab = L190; break; }
        if (IV_(IRC) != 5) { lab = L180; break; }
        if (V_(RADFAC) <= 1.0) { lab = L180; break; }
        if (V_(PREDUC) > 1.2 * V_(FDIF)) { lab = L180; break; }
        if (IV_(RESTOR) != 2) { lab = L190; break; }
        rstrst = 0;
        lab = L200;
        break;
      case L180: {   // ask for f at x0 + step
        const int x01 = IV_(X0), step1 = IV_(STEP);
        v2axy(n, x, 1.0, &V_(step1), &V_(x01));
        IV_(NFCALL) = IV_(NFCALL) + 1;
        IV_(1) = 1;
        return;
      }
      case L190:
        rstrst = 3;
        lab = L200;
        break;
      case L200: {   // assess the step
        const int x01 = IV_(X0);
        V_(RELDX) = rldst(n, d, x, &V_(x01));
        a7sst(iv, v);
        switch (IV_(RESTOR) + 1) {
          case 1: lab = L240; break;
          case 3: lab = L220; break;
          case 4: lab = L230; break;
          default: lab = L210; break;
        }
        break;
      }
      case L210: {
        const int x01 = IV_(X0);
        v7cpy(n, x, &V_(x01));
        lab = L240;
        break;
      }
      case L220: {
        const int step1 = IV_(STEP), lstgst = IV_(STLSTG);
        v7cpy(n, &V_(lstgst), &V_(step1));
        lab = L240;
        break;
      }
      case L230: {
        const int x01 = IV_(X0), step1 = IV_(STEP), lstgst = IV_(STLSTG);
        v7cpy(n, &V_(step1), &V_(lstgst));
        v2axy(n, x, 1.0, &V_(step1), &V_(x01));
        V_(RELDX) = rldst(n, d, x, &V_(x01));
        IV_(RESTOR) = rstrst;
        lab = L240;
        break;
      }
      case L240:
        switch (IV_(IRC)) {
          case 2: case 3: case 4: lab = L280; break;
          case 6: lab = L260; break;
          case 7: case 8: case 9: case 10: case 11: case 12: lab = L270; break;
          case 13: lab = L330; break;
          case 14: lab = L300; break;
          default: lab = L250; break;   // 1, 5
        }
        break;
      case L250:
        V_(RADIUS) = V_(RADFAC) * V_(DSTNRM);
        lab = L110;
        break;
      case L260:
        V_(RADIUS) = V_(LMAXS);
        lab = L150;
        break;
      case L270:
        IV_(CNVCOD) = IV_(IRC) - 4;
        if (V_(F) >= V_(F0)) { lab = L340; break; }
        if (IV_(XIRC) == 14) { lab = L340; break; }
        IV_(XIRC) = 14;
        lab = L280;
        break;
      case L280:
        if (IV_(IRC) == 3) {   // L L^T step for the radius test after the update
          const int step1 = IV_(STEP), temp1 = IV_(STLSTG), l = IV_(LMAT);
          l7tvm(n, &V_(temp1), &V_(l), &V_(step1));
          l7vml(n, &V_(temp1), &V_(l), &V_(temp1));
        }
        lab = L290;
        break;
      case L290:
        IV_(NGCALL) = IV_(NGCALL) + 1;
        IV_(1) = 2;
        return;
      case L300: {   // secant update of L after a new gradient
        const int g01 = IV_(G0), step1 = IV_(STEP), temp1 = IV_(STLSTG);
        v2axy(n, &V_(g01), -1.0, &V_(g01), g);
        bool inc = false;
        if (IV_(IRC) == 3) {
          v2axy(n, &V_(temp1), -1.0, &V_(g01), &V_(temp1));
          v7vmp(n, &V_(temp1), &V_(temp1), d, -1);
          if (v2nrm(n, &V_(temp1)) <= V_(DGNORM) * V_(TUNER4)) inc = true;
          else if (!(d7tpr(n, g, &V_(step1)) >= V_(GTSTEP) * V_(TUNER5))) inc = true;
        }
        if (inc) V_(RADFAC) = V_(INCFAC);
        const int w = IV_(NWTSTP), z = IV_(X0), l = IV_(LMAT);
        w7zbf(&V_(l), n, &V_(step1), &V_(w), &V_(g01), &V_(z));
        l7upd(&V_(temp1), &V_(step1), &V_(l), &V_(g01), &V_(l), n, &V_(w), &V_(z));
        IV_(1) = 2;
        lab = L80;
        break;
      }
      case L310:
      case L320:
        lab = L80;
        break;
      case L330:
        IV_(1) = 64;
        lab = L350;
        break;
      case L340:
        IV_(1) = IV_(CNVCOD);
        IV_(CNVCOD) = 0;
        lab = L350;
        break;
      case L350:
        itsum(iv, v);
        return;
    }
  }
}

}  // extern "C"
