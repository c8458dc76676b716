/*
 * C-ABI of the MI355X ECG pair path (libecg_hip.so).
 *
 * Conventions (the reference's Fortran storage, so a Fortran host can pass
 * its arrays through ISO_C_BINDING unchanged):
 *   - matrices are column-major with an explicit leading dimension;
 *   - complex values are two interleaved doubles (re, im), i.e.
 *     complex(8) / double _Complex / hipDoubleComplex, passed as double*;
 *   - function indices (nmin, nmax, nfru, which, k) are 1-based;
 *   - the parameter column of function k is (vech L_k, vech B_k), npt =
 *     n(n+1) doubles, vech running down the columns of the lower triangle
 *     (src/matelem.f90:109-122);
 *   - every entry point returns an int status (ECG_OK = 0) and never exits.
 *
 * The caller owns every host array; the library owns the device state of a
 * context (parameters, H, S, factor, work buffers) so repeated calls do not
 * re-upload.  A context is not re-entrant: calls on one context must not
 * overlap.  All entry points are synchronous: they return after the results
 * are in the caller's arrays.
 */
#ifndef ECG_H
#define ECG_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ECG_OK 0
#define ECG_ERR_ARG (-1)         /* invalid argument */
#define ECG_ERR_HIP (-2)         /* HIP runtime error (no device, launch failure) */
#define ECG_ERR_NOMEM (-3)       /* device allocation failed */
#define ECG_ERR_NONFINITE (-4)   /* a pair element is Inf/NaN (pseudo-Cholesky breakdown);
                                    the reference does not flag this (src/matelem.f90:211-214) */
#define ECG_ERR_STATE (-5)       /* call order violated (e.g. no parameters set) */
#define ECG_ERR_SYNC (-6)        /* a device-side wait between workgroups exceeded its time limit */
/* positive codes mirror the reference:
 *   GHEPIIS / EnergyIA: 1 = H - sigma S singular (zero pivot), 2 = inverse
 *   iteration not converged to Tol within MaxIter (src/linalg.f90:1911);
 *   energy with root search: 3 = the requested root could not be isolated. */
#define ECG_SINGULAR 1
#define ECG_NOT_CONVERGED 2
#define ECG_ROOT_NOT_FOUND 3

typedef struct ecg_ctx ecg_ctx;

/* Create a context for one physical system.  Replaces the set-up the
 * reference does in ProgramDataInit (src/workproc.f90:575-1202):
 *   n          pseudo-particles (1..5)
 *   T          number of simplified Y^dagger Y terms (1..16)
 *   yhy_mats   T integer n x n matrices, column-major, term after term
 *              (Glob_YHYMatr, src/globvars.f90)
 *   yhy_coef   T integer coefficients (Glob_YHYCoeff)
 *   mass       n x n mass matrix, column-major (Glob_MassMatrix)
 *   q0, q      charge of particle 1 and the n pseudo-charges
 *              (Glob_PseudoCharge0, Glob_PseudoCharge) */
int ecg_ctx_create(int n, int T, const int* yhy_mats, const int* yhy_coef, const double* mass, double q0,
                   const double* q, ecg_ctx** out);
int ecg_ctx_destroy(ecg_ctx* ctx);

/* Upload the nonlinear parameters of K functions (Glob_NonlinParam(npt, K),
 * src/globvars.f90:141-146).  With the basis size unchanged, only the
 * functions from the first to the last one that differ from the previous
 * upload are sent (for up to 240 doubles as a kernel argument, without
 * waiting) and their records recomputed; the next energy call reassembles
 * the rows from the first changed function. */
int ecg_set_params(ecg_ctx* ctx, int K, const double* params);
/* The same from device memory (dparams: npt x K doubles in device memory of
 * the context's GPU, for callers that keep the parameters resident).  The
 * copy is queued on the context's stream without waiting: dparams must hold
 * its values when the call is made (work that produced them completed or
 * synchronised) and stay unchanged until the next energy or assembly call on
 * ctx has returned.  Values are not checked for finiteness here; non-finite
 * parameters surface as ECG_ERR_NONFINITE from that call.  No counterpart in
 * the reference, whose parameters live in host memory. */
int ecg_set_params_device(ecg_ctx* ctx, int K, const double* dparams);
/* Replace functions first..first+count-1 (1-based) of the current basis
 * (the one-function steps of OptCycleG/I, src/workproc.f90:5202-5240):
 * params (npt x count) is read before the call returns; for up to 240
 * doubles (20 functions at n = 3) the values travel as a kernel argument
 * and the call queues one launch on the context's stream without waiting,
 * larger updates are copied and waited for. */
int ecg_update_params(ecg_ctx* ctx, int first, int count, const double* params);

/* One pair and one symmetry matrix P on the device: MatrixElements
 * (src/matelem.f90:12-750).  Dk, Dl: 4*np complex each
 * (dH/dvechL, dH/dvechB, dS/dvechL, dS/dvechB), written when gradflag asks
 * for them (1: k, 2: l, 3: both).  The system constants of ctx are used. */
int ecg_matrix_elements(ecg_ctx* ctx, const double* paramk, const double* paraml, const int* P, int gradflag,
                        double* Hkl, double* Skl, double* Dk, double* Dl);

/* ecg_matrix_elements without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture set-up, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results (bit for bit) as
 * ecg_matrix_elements; the pair's staging (parameters, records, output:
 * at most 640 doubles for n <= 5) lives in a library-owned 8 KiB device
 * array per GPU, held until the process ends.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (from several threads,
 * or contexts on the same device) race and must be serialised by the
 * caller.  ECG_ERR_ARG if the staging would exceed the array. */
int ecg_matrix_elements_noalloc(ecg_ctx* ctx, const double* paramk, const double* paraml, const int* P,
                                int gradflag, double* Hkl, double* Skl, double* Dk, double* Dl);

/* ComputeMatElem(Nmin, Nmax) + StoreHS (src/matform.f90:11-51, 119-230):
 * rows nmin..nmax, all l <= k, summed over the Y^dagger Y terms, written
 * with the reference's normalisation and layout:
 *   method 'G': diagS(k) = Re S_kk, diagH(k) = Re H_kk / diagS(k); lower
 *               triangles of H, S hold H_kl f, S_kl f, f = 1/sqrt(diagS_k diagS_l)
 *   method 'I': S(k,k) = 1, H(k,k) = diagH(k) - approx_energy; S in both
 *               triangles; lower H holds (H_kl - approx_energy S_kl) f
 * Elements of rows < nmin are neither computed nor written, but their
 * diagS must be the values of an earlier call (as in the reference). */
int ecg_compute_mat_elem(ecg_ctx* ctx, int nmin, int nmax, char method, double approx_energy, double* H,
                         double* S, int ld, double* diagH, double* diagS);

/* ComputeMatElemAndDeriv(nfru+1, K) + StoreHSD (src/matform.f90:55-115,
 * 234-360).  D: Glob_D(2*npt, ldD, K) complex (src/workproc.f90:4992
 * allocates Glob_D(2*npt, NumOfFuncToOpt, cbs)); ldD >= K - nfru is its
 * allocated second extent, so that the last chunk of a cycle, with fewer
 * functions than NumOfFuncToOpt, lands where the reference stores it.
 * Only D(:, 1:K-nfru, 1:K) is written.  H, S, diagH, diagS as
 * ecg_compute_mat_elem for rows nfru+1..K. */
int ecg_compute_mat_elem_and_deriv(ecg_ctx* ctx, int nfru, char method, double approx_energy, double* H,
                                   double* S, int ld, double* diagH, double* diagS, double* D, int ldD);

/* GHEPIIS (src/linalg.f90:1781-1926) on host arrays: inverse iteration for
 * the eigenvalue of (A, B) nearest apprlambda.  M: lower triangle holds
 * A - apprlambda B (rows >= k; rows < k as well, the factor is recomputed
 * from row 1 on the device), the strict upper triangle receives L^H;
 * invD receives 1/D (complex, imaginary part zero); B full Hermitian.
 * v: start vector; x: eigenvector (normalised x^H B x = 1 when
 * specif_norm = 0, x^H x = 1 when 1, max |Re|,|Im| = 1 otherwise).
 * Returns 0, ECG_SINGULAR or ECG_NOT_CONVERGED like the reference's
 * ErrorCode.  tol > 0: converged when the measure is <= tol; tol <= 0
 * (the reference's other mode, src/linalg.f90:1902-1906): converged at the
 * first iteration whose measure rises above the previous one and is
 * <= |tol|; numiter counts that iteration.  k is accepted for the
 * reference's argument list only: the factor is always recomputed from
 * row 1, so M's upper triangle need not hold a factor.  M is factored as
 * given (the shift is added to the Rayleigh quotient afterwards).
 * The device H, S of ctx then hold M and B: the context's basis must be
 * reassembled, and when n differs from its K the parameters must be set
 * again (energy calls return ECG_ERR_STATE until ecg_set_params).
 * The energy paths below take tol > 0 the same way; with tol <= 0 they
 * run the iteration to its rounding floor (the measure stops halving below
 * 1e-9) and succeed when the measure there is <= |tol|. */
int ecg_ghepiis(ecg_ctx* ctx, int k, int n, double* M, int nM, double* invD, const double* B, int nB,
                double apprlambda, const double* v, double tol, int maxiter, int specif_norm, double* lambda,
                double* x, double* relacc, int* numiter);

/* Device-resident energy: assembles H, S for all K functions, then finds
 * root number `which` (1 = lowest) of H c = E S c by LDL^H inverse
 * iteration.  If sigma is finite it is used as the shift (the reference's
 * method 'I', Glob_ApproxEnergy); otherwise the shift is found from
 * Sylvester inertia counts of LDL^H(H - sigma S) (replacing ZHEGVX's
 * RANGE='I' selection, src/workproc.f90:1750).  c (nullable): K complex,
 * c^H S c = 1 for the normalised S.  info (nullable, 8 doubles):
 *   [0] iterations, [1] convergence measure, [2] final shift,
 *   [3] factorisations, [4] negative pivots at the final shift,
 *   [5] x^H S x before normalisation, [7] rows refactored by a tail update
 *   (-1: full factorisation).
 * Like EnergyIA(Nmin, ...) (src/workproc.f90:1954-2061), work starts at
 * the first function whose parameters changed since the previous call
 * (ecg_set_params compares with the previous upload, ecg_update_params
 * names it): only rows from there on are reassembled, and with a finite
 * sigma equal to the previous call's and at most 64 changed rows, the
 * LDL^H factor and L^{-1} are completed from that row (O(r K^2) for r
 * rows, the reference's LDLHF from Nmin) instead of recomputed. */
int ecg_energy(ecg_ctx* ctx, int which, double sigma, double tol, int maxiter, double* E, double* c, double* info);

/* Device-resident energy and analytic gradient (EnergyGB/EnergyIB,
 * src/workproc.f90:1841-1950, 2063-2154): the gradient with respect to the
 * parameters of functions nfru+1..K, ordered (dE/dvechL_k, dE/dvechB_k)
 * per function (npt*(K-nfru) doubles).  The per-pair derivatives are
 * recomputed inside one fused adjoint pass and contracted with the
 * eigenvector on the device; nothing of size K^2*npt is stored.  Rows are
 * reassembled and refactored from the first changed function as in
 * ecg_energy. */
int ecg_energy_grad(ecg_ctx* ctx, int nfru, int which, double sigma, double tol, int maxiter, double* E,
                    double* grad, double* info);

/* The same energy and gradient in the reference's stored-derivative
 * scheme (ComputeMatElemAndDeriv + StoreHSD + EnergyGB's contraction,
 * src/matform.f90:234-360, src/workproc.f90:1925-1940): the per-pair
 * derivatives of rows nfru+1..K are stored on the device in Glob_D's layout
 * (2 npt x (K-nfru) x K complex: 1.5 GB at K=2000, n=3) and contracted with
 * the eigenvector in a second pass.  Timings: [0] assembly with stored
 * derivatives, [1] eigen step, [2] contraction, [3] whole call. */
int ecg_energy_grad_stored(ecg_ctx* ctx, int nfru, int which, double sigma, double tol, int maxiter, double* E,
                           double* grad, double* info);

/* Packed batches of independent problems (uneven K, mixed n), no
 * counterpart in the reference (which runs one problem per MPI job):
 * ecg_batch_create binds nprob contexts whose parameters are set.
 * ecg_batch_energy_grad computes, for every problem i, the root nearest
 * sigma[i] and the gradient of all its functions (as ecg_energy_grad with
 * nfru = 0): the pair phases of all problems with the same n run as one
 * launch each (assembly, adjoint), the eigen steps of up to nstreams
 * problems run concurrently (largest first).  grad: the problems' gradients
 * concatenated in order; lat_ms (nullable): time from the call's start
 * until each problem's results were on the host.
 * ecg_batch_set_config selects the mode: 0 every problem's whole chain on
 * its own (up to nstreams concurrently); 1 pair phases packed, eigen steps
 * per problem; 2 pair phases packed and the factorisations of H - sigma S
 * of all problems with a finite sigma[i] in shared launches (problem-major
 * workgroups), then their inverse iterations per problem.  nstreams (1..16)
 * bounds the problems whose per-problem steps run concurrently.  An
 * autotuner times the combinations (bench.py). */
typedef struct ecg_batch ecg_batch;
int ecg_batch_create(int nprob, ecg_ctx** ctxs, ecg_batch** out);
int ecg_batch_destroy(ecg_batch* b);
int ecg_batch_set_config(ecg_batch* b, int packed, int nstreams);
int ecg_batch_energy_grad(ecg_batch* b, const double* sigma, double tol, int maxiter, double* E, double* grad,
                          double* lat_ms);

/* Per-kernel tracing of one context: with tracing on, every launch of the
 * kernels below is bracketed by HIP events on the context's stream and its
 * device time is accumulated per kernel (the reference's own timing is
 * MPI_WTIME around whole routines).  ecg_trace(ctx, 1) enables and resets;
 * ecg_trace_read writes launch counts and total milliseconds per label and
 * returns the number of labels written. */
#define ECG_TRACE_OFFDIAG 0       /* off-diagonal pair assembly */
#define ECG_TRACE_ADJOINT 1       /* fused adjoint pair pass */
#define ECG_TRACE_LDL_UPDATE 2    /* trailing update + look-ahead diagonal block + next panel */
#define ECG_TRACE_LDL_DIAG 3      /* first diagonal block */
#define ECG_TRACE_LDL_PANEL 4     /* first panel */
#define ECG_TRACE_TRSV_FWD 5      /* forward triangular solve */
#define ECG_TRACE_TRSV_BWD 6      /* backward triangular solve */
#define ECG_TRACE_GEMV 7          /* (reserved) */
#define ECG_TRACE_FORM_SHIFTED 8  /* no launches: the factorisation reads H - sigma S in its first pass */
#define ECG_TRACE_LDL_TAIL 9      /* incremental refactorisation from a changed row */
#define ECG_TRACE_COUNT 10
int ecg_trace(ecg_ctx* ctx, int enable);
int ecg_trace_read(ecg_ctx* ctx, double* count, double* total_ms, int nmax);

/* Current parameters of the K functions (npt x K), e.g. after
 * ecg_screen_loop accepted candidates into function K. */
int ecg_get_params(ecg_ctx* ctx, double* params);

/* Candidate screening for the last basis position: the batched
 * replacement of the reference's trial loops (GenerateTrialParam and the
 * trial loops of BasisEnl / OptCycle, src/workproc.f90:1375-1487,
 * 3397-3413, 4054-4070), each trial of which recomputes row K of H, S,
 * updates the LDL^H factor from row K (LDLHF, src/linalg.f90:316) and runs
 * GHEPIIS (src/linalg.f90:1781) at the shift Glob_ApproxEnergy.
 *
 * ecg_screen_prepare fixes functions 1..K-1 of the current parameters and
 * the shift sigma: factor of H - sigma S, G = (H' - sigma S')^{-1} and
 * G S' of the leading K-1 block on the device, and the energy and
 * eigenvector of the current basis at sigma (inverse iteration, tol and
 * maxiter as GHEPIIS).  info (nullable, 4): [0] current energy, [1]
 * negative pivots of H' - sigma S', [2] iterations, [3] convergence
 * measure.  Returns ECG_SINGULAR if H' - sigma S' has a zero pivot. */
int ecg_screen_prepare(ecg_ctx* ctx, double sigma, double tol, int maxiter, double* info);

/* Energies of ncand candidates for function K (cand: npt x ncand; each
 * replaces function K alone): `iters` inverse iterations at sigma from the
 * current eigenvector on the bordered matrices, Rayleigh quotients.
 * info (nullable, 4 per candidate): [0] convergence measure of the last
 * iteration (GHEPIIS's), [1] negative pivots of H - sigma S with the
 * candidate, [2] the candidate's pivot (Schur complement), [3] x^H S x. */
int ecg_screen(ecg_ctx* ctx, int ncand, const double* cand, int iters, double* E, double* info);

/* nsteps optimisation steps on the device (cand: npt x ncand x nsteps):
 * each step screens its ncand candidates and keeps the best one as
 * function K when it lowers the current energy.  One step is captured in
 * a hipGraph and replayed nsteps times with no host synchronisation.
 * Ebest[s]: best energy of step s; ibest[s]: accepted candidate index, or
 * -1 - index of the best candidate when it was not accepted. */
int ecg_screen_loop(ecg_ctx* ctx, int nsteps, int ncand, const double* cand, int iters, double* Ebest, int* ibest);

/* The screening step's shapes, no counterpart in the reference: the
 * per-step GEMMs' row tile (32 or 64) and the border-row kernel's
 * candidates per workgroup (1, 2, 4 or 8); 0 leaves a shape to the timing
 * that ecg_screen / ecg_screen_loop do on their first call for a candidate
 * count (the default).  Results do not depend on the shapes; with at most
 * 8 candidates the products run as a GEMV and the row tile is unused.  ECG_ERR_ARG
 * for other values or a candidates-per-workgroup count whose staged
 * records exceed 64 KB of LDS for this system. */
int ecg_screen_set_shapes(ecg_ctx* ctx, int row_tile, int cand_per_block);

/* Overlap penalty settings: the reference's Glob_OverlapPenaltyAllowed,
 * Glob_MaxOverlapPenalty (b) and Glob_OverlapPenaltyThreshold2 (t2)
 * (src/globvars.f90:309-326), plus the Glob_nfru that energy-only calls
 * use.  While allowed, ecg_energy returns E + P (EnergyGA/EnergyIA,
 * src/workproc.f90:1759-1764, 1996-2001), and ecg_energy_grad,
 * ecg_energy_grad_stored and ecg_batch_energy_grad return E + P and add
 * the penalty's gradient (EnergyGB/EnergyIB, src/workproc.f90:1942-1945,
 * 2144-2147; P over the pairs of their own nfru, 0 for batches; nothing
 * for K = 1).  info[6] of those calls receives P (0 when not allowed).
 * Screening energies do not include it. */
int ecg_set_overlap_penalty(ecg_ctx* ctx, int allowed, double max_pair_penalty, double threshold2, int nfru);

/* ComputeOverlapPenalty / ComputeOverlapPenaltyAndAddGradient
 * (src/workproc.f90:1491-1693) on the device S of the last assembly of the
 * current parameters (any energy call, or ecg_compute_mat_elem up to row
 * K): total = sum of b/(1-t2) (|S_ji|^2 - t2) over the pairs j > i,
 * j > nfru, with |S_ji|^2 > t2.  grad (nullable, npt*(K-nfru) doubles,
 * ordered as ecg_energy_grad's) has the penalty's gradient terms added,
 * as the reference adds them to Glob_WkGR.  ECG_ERR_STATE if no such
 * assembly is on the device. */
int ecg_overlap_penalty(ecg_ctx* ctx, int nfru, double max_pair_penalty, double threshold2, double* total,
                        double* grad);

/* Expectation values (MatrixElementsForExpcVals, src/matelem.f90:949-1983,
 * summed as the reference's driver sums them, src/workproc.f90:8158-8300,
 * with the bra unpermuted and the ket's Y^dagger Y terms) in the
 * eigenvector c: K complex, normalised c^H S c = 1 for the normalised S;
 * NULL uses the eigenvector of the last ecg_energy / ecg_energy_grad call,
 * which must belong to the current parameters (ECG_ERR_STATE after any
 * call that changed them, or after a failed energy call).
 * values (5 m + 4 doubles, m = n(n+1)/2): <1/r^2>, <1/r>, <r>, <r^2>,
 * <delta(r)>, each for the pseudo-particle pairs (a, b >= a) in the order
 * (1,1), (1,2), .., (1,n), (2,2), .. (a = b: pseudo-particle a's distance
 * from particle 1), then <H>, <S>, <T>, <V>.  errors (nullable, same
 * layout): the driver's estimate of each value's rounding error.  The
 * reference's Drachmanized delta, mass-velocity, Darwin and orbit-orbit
 * terms, correlation functions and densities are not computed. */
int ecg_expectation_values(ecg_ctx* ctx, const double* c, double* values, double* errors);

/* Pair correlation functions on a radial grid: the AreCorrFuncNeeded part
 * of MatrixElementsForExpcVals (src/matelem.f90:1934-1946),
 *   CF_kl(q, r) = S_kl exp(-r^2 / TrCJ_q) / (sqrt(TrCJ_q) TrCJ_q) / pi^(3/2),
 * contracted with the eigenvector of the last energy call (c == NULL) or
 * with c (K complex) as the driver's CFDMEkl_s (src/workproc.f90:
 * 8234-8243).  grid: ngrid radii; values: n(n+1)/2 x ngrid doubles, pair
 * index q fastest (pairs (a, b >= a) in the order of the other pair
 * values).  CF_q(0) equals <delta(r_q)>; 4 pi int r^2 CF_q(r) dr = <S>. */
int ecg_correlation_functions(ecg_ctx* ctx, const double* c, int ngrid, const double* grid, double* values);

/* Particle densities on a radial grid: the ArePartDensNeeded part of
 * MatrixElementsForExpcVals (src/matelem.f90:1948-1963), the form above with
 * TrCJ replaced by b_i^T C^-1 b_i, where b_i (Glob_bvc, src/workproc.f90:
 * 612-621) comes from the n + 1 particle masses (Glob_Mass: masses[0] is
 * particle 1, the reference particle).  values: (n + 1) x ngrid doubles,
 * particle index fastest, as the driver's Dens(i, a). */
int ecg_particle_densities(ecg_ctx* ctx, const double* c, const double* masses, int ngrid, const double* grid,
                           double* values);

/* Mass-velocity and Darwin corrections (without the alpha^2 factor and
 * without the Drachman forms): the MVkl and Darwinkl pair values of
 * MatrixElementsForExpcVals (src/matelem.f90:1479-1667, 1866-1884)
 * contracted with the vector (c == NULL: the last energy call's) as the
 * driver does, with its rounding estimates (src/workproc.f90:8158-8300).
 * masses: the n + 1 particle masses (Glob_Mass).  values, errors: 2 each
 * (MV, Darwin); errors nullable. */
int ecg_relativistic_corrections(ecg_ctx* ctx, const double* c, const double* masses, double* values,
                                 double* errors);

/* Drachmanized delta functions, orbit-orbit term and the Drachmanized
 * mass-velocity and Darwin corrections (without the alpha^2 factor): the
 * drach_deltarkl, OOkl, drach_MVkl and drach_Darwinkl pair values of
 * MatrixElementsForExpcVals (src/matelem.f90:1276-1478, 1669-1933)
 * contracted with the vector (c == NULL: the last energy call's) as the
 * driver does (src/workproc.f90:8158-8300), with its rounding estimates.
 * masses: the n + 1 particle masses (Glob_Mass); E: the current energy
 * (Glob_CurrEnergy), which the Drachmanized forms take.  values, errors:
 * n(n+1)/2 + 3 each: the Drachmanized delta of the pairs (a, b >= a) in the
 * order of ecg_expectation_values, then OO, drach MV, drach Darwin; errors
 * nullable.  For n >= 3 the Drachmanized delta takes, as the reference's
 * build does, zero for the tr[C^-1 J_aa C^-1 J_bc] (a < b < c) that its
 * table leaves unwritten. */
int ecg_drachmanized_corrections(ecg_ctx* ctx, const double* c, const double* masses, double E, double* values,
                                 double* errors);

/* The eigenvector (K complex, c^H S c = 1 for the normalised S) of the last
 * energy call on the current parameters: what EnergyGB / EnergyIB leave
 * in Glob_c (src/workproc.f90:1885-1890, 2114-2119).  ECG_ERR_STATE when
 * the parameters changed since, or the last energy call failed. */
int ecg_eigenvector(ecg_ctx* ctx, double* c);

/* Device time (ms, HIP events on the context's stream) of the phases of
 * the last ecg_energy / ecg_energy_grad call: [0] assembly, [1] eigen step
 * (all factorisations and iterations), [2] adjoint pass, [3] whole call,
 * [4] adjoint kernel alone, [5] off-diagonal assembly kernel alone.
 * After ecg_screen / ecg_screen_loop, [0] is the device time of the call,
 * [1] the row tile (64 or 32) its GEMMs ran with and [2] the candidates per
 * workgroup (1, 2, 4 or 8) of its border-row kernel, both picked by timing
 * the variants on the first call for this candidate count and basis-size
 * band.  Returns the number of values written. */
int ecg_last_timings(ecg_ctx* ctx, double* ms, int nmax);

#ifdef __cplusplus
}
#endif

#endif /* ECG_H */
