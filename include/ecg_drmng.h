/*
 * C-ABI of the nonlinear-parameter minimiser (libecg_drmng.so, host code
 * only): the reverse-communication secant method with a double-dogleg
 * trust region that the reference's optimisation drivers call (PORT's
 * DIVSET / DRMNG, src/dmng.f:1218 and :2510; drivers BasisEnlG/I,
 * OptCycleG/I, FullOpt1G/I in src/workproc.f90).
 *
 * Arguments mirror the Fortran routines one for one (arrays by pointer,
 * scalars by value), with the same workspace layout, so a Fortran host can
 * bind them through ISO_C_BINDING in place of its own DIVSET / DRMNG:
 *   iv[liv], v[lv]  the integer and double workspaces; IV(k) is iv[k - 1];
 *   d[n]            positive scale vector;
 *   x[n]            the point, overwritten with the next point to evaluate;
 *   fx, g[n]        f(x) and its gradient at the point of the last request.
 * Protocol: ecg_divset(2, iv, liv, lv, v) loads the defaults (IV(1) = 12);
 * the caller may change IV / V entries; then ecg_drmng is called in a loop
 * and its IV(1) on return says what it needs: 1 = f(x) into fx, 2 = the
 * gradient into g, 3..8 = converged (x is the best point), 9 / 10 =
 * evaluation / iteration limit, other values = the reference's error codes
 * (e.g. 18: a d(i) <= 0; 15 / 16: liv / lv too small).  A caller that
 * cannot evaluate f at x sets IV(2) = 1 before the next call; the
 * minimiser then shortens the step.  Sizes: liv >= 59, lv >= 71 + n (n + 13)
 * / 2.  Nothing is allocated and nothing is printed; calls on different
 * workspaces may run concurrently.
 */
#ifndef ECG_DRMNG_H
#define ECG_DRMNG_H

#ifdef __cplusplus
extern "C" {
#endif

/* DIVSET (src/dmng.f:1218): defaults of iv / v for algorithm alg (2: the
 * general unconstrained minimiser with analytic gradient). */
void ecg_divset(int alg, int* iv, int liv, int lv, double* v);

/* DRMNG (src/dmng.f:2510): one reverse-communication step. */
void ecg_drmng(double* d, double fx, const double* g, int* iv, int liv, int lv, int n, double* v, double* x);

#ifdef __cplusplus
}
#endif

#endif
