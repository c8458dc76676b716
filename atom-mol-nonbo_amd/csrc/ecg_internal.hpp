// Internal declarations shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "ecg_math.hpp"

namespace ecg {

// p[i] if ok, else zero, with an unconditional load (of p[0], always valid,
// when !ok) and a select of the value.  Written as `ok ? p[i] : 0` the load
// becomes a branch, and the compiler waits for each such load at the join,
// so a run of them costs one memory latency each instead of one in total.
__device__ inline zd ld_or0(const zd* p, size_t i, bool ok) {
  const zd v = p[ok ? i : 0];
  return mkz(ok ? v.x : 0.0, ok ? v.y : 0.0);
}
__device__ inline double ld_or0(const double* p, size_t i, bool ok) {
  const double v = p[ok ? i : 0];
  return ok ? v : 0.0;
}

constexpr int TILE = 16;   // pair tile edge (TILE x TILE pairs per workgroup)
constexpr int NB = 64;     // block size of the LDL^H factorisation

// Per-function bra record (doubles): Ck* packed complex, |det L|, L packed.
template <int N> struct BraRec {
  static constexpr int m = N * (N + 1) / 2;
  static constexpr int ck = 0;
  static constexpr int adet = 2 * m;
  static constexpr int lp = 2 * m + 1;
  static constexpr int stride = ((3 * m + 1) + 1) / 2 * 2;
};
// Per-(function, term) ket record: P^T (A + iB) P packed complex, M tCl.
template <int N> struct KetRec {
  static constexpr int m = N * (N + 1) / 2;
  static constexpr int ct = 0;
  static constexpr int nl = 2 * m;
  static constexpr int stride = 2 * m + 2 * N * N;
};

inline int bra_stride(int n) {
  const int m = n * (n + 1) / 2;
  return ((3 * m + 1) + 1) / 2 * 2;
}
inline int ket_stride(int n) { return n * (n + 1) + 2 * n * n; }

// Per-kernel tracing: event pairs around launches, aggregated per label
// after the stream has synchronised.  Off unless enabled on the context.
enum TraceLabel {
  TR_OFFDIAG = 0, TR_ADJOINT, TR_LDL_UPDATE, TR_LDL_DIAG, TR_LDL_PANEL, TR_TRSV_FWD, TR_TRSV_BWD, TR_GEMV,
  TR_FORM_SHIFTED, TR_LDL_TAIL, TR_COUNT   // TR_FORM_SHIFTED: unused (the factorisation forms H - sigma S itself)
};
struct Tracer {
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<int> label;       // per recorded pair
  int used = 0;                 // events used
  bool pending = false;         // a begin() awaits its end()
  double sum_ms[TR_COUNT] = {};
  long long count[TR_COUNT] = {};
  void begin(int lab, hipStream_t s) {
    if (!on || used + 2 > (int)pool.size()) return;
    (void)hipEventRecord(pool[used], s);
    label.push_back(lab);
    used += 2;
    pending = true;
  }
  void end(hipStream_t s) {
    if (!pending) return;
    (void)hipEventRecord(pool[used - 1], s);
    pending = false;
  }
  // call after the stream has synchronised
  void collect() {
    for (size_t i = 0; i < label.size(); ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, pool[2 * i], pool[2 * i + 1]) == hipSuccess) {
        sum_ms[label[i]] += ms;
        count[label[i]] += 1;
      }
    }
    label.clear();
    used = 0;
  }
};

// Device state of one basis (owned by the context in ecg_capi.hip).
struct Ctx {
  SysConst sc;
  int n = 0, T = 0, K = 0, ld = 0;
  int cap = 0;            // allocated basis capacity (functions)
  double* params = nullptr;   // npt x cap
  double* bra = nullptr;      // cap x bra_stride
  double* ket = nullptr;      // cap x T x ket_stride
  double* diagS = nullptr;    // cap
  double* diagH = nullptr;    // cap
  zd* H = nullptr;            // ld x ld
  zd* S = nullptr;            // ld x ld
  zd* rowbuf = nullptr;       // thin_row_entries(cap): raw (H, S) of the pairs of a few rows
  double* adjraw = nullptr;   // thin_adjoint_entries(n, cap): per-pair adjoint values of a few rows
  // optional events bracketing the off-diagonal and adjoint kernels
  hipEvent_t ev_off0 = nullptr, ev_off1 = nullptr, ev_adj0 = nullptr, ev_adj1 = nullptr;
  Tracer* tr = nullptr;
};

// pair assembly (ecg_kernels.hip)
int prep(const Ctx& c, hipStream_t s);
int prep_range(const Ctx& c, int f0, int f1, hipStream_t s);   // records of functions f0..f1-1
// Parameters small enough to travel as a kernel argument (count functions,
// n (n + 1) doubles each, at most PB_DOUBLES in all).
constexpr int PB_DOUBLES = 240;
struct ParamBlock {
  double v[PB_DOUBLES];
};
// Stores the host parameters hparams of functions f0..f0+count-1 and
// computes their records, one launch and no host wait; -1 when they do not
// fit a ParamBlock.
int prep_put(const Ctx& c, int f0, int count, const double* hparams, hipStream_t s);
int assemble(const Ctx& c, int r0, int r1, hipStream_t s);   // rows r0..r1-1 (thin-row kernels for few rows)
size_t thin_row_entries(int cap);
size_t thin_adjoint_entries(int n, int cap);
int assemble_deriv(const Ctx& c, int nfru, zd* D, hipStream_t s);
// E is taken from *Ep (device) when Ep is not null.
int adjoint(const Ctx& c, int nfru, double E, const double* Ep, const zd* cvec, double* part, double* grad,
            hipStream_t s);
size_t adjoint_partial_doubles(int n, int K);
// Overlap penalty of the device S (pairs j > i, j >= nfru): per-function
// partials pen[K - nfru], their fixed-order sum to *total (device), and,
// if grad is not null, the penalty's gradient added to grad (npt per
// function nfru..K-1).  pc = b / (1 - t2).
int overlap_penalty(const Ctx& c, int nfru, double pc, double t2, double* pen, double* total, double* grad,
                    hipStream_t s);

// Expectation values of MatrixElementsForExpcVals' pair operators in the
// eigenvector cvec (K complex, normalised for the normalised S): out
// receives 2 (5 m + 4) doubles, the values then the driver's rounding
// bounds (before their scaling by epsilon), m = n(n+1)/2.  part: scratch of
// expectation_partial_doubles(n, K).
int expectation_values(const Ctx& c, const zd* cvec, double* part, double* out, hipStream_t s);
size_t expectation_partial_doubles(int n, int K);
// Pair correlation functions at ngrid radii (host array grid): out (device,
// n(n+1)/2 x ngrid, pair index fastest); part: correlation_partial_doubles.
int correlation_functions(const Ctx& c, const zd* cvec, int ngrid, const double* grid, double* part, double* out,
                          hipStream_t s);
// Particle densities at ngrid radii from the n + 1 particle masses: out
// (device, (n + 1) x ngrid, particle index fastest).
int particle_densities(const Ctx& c, const zd* cvec, const double* masses, int ngrid, const double* grid,
                       double* part, double* out, hipStream_t s);
// Mass-velocity and Darwin values (out: 2 values, 2 bound sums) from the
// n + 1 particle masses; part: correlation_partial_doubles.
int relativistic_values(const Ctx& c, const zd* cvec, const double* masses, double* part, double* out,
                        hipStream_t s);
size_t correlation_partial_doubles(int n, int K);
// Drachmanized delta of every pair (a, b >= a), orbit-orbit, Drachmanized
// mass-velocity and Darwin (n(n+1)/2 + 3 values, then as many bound sums)
// at the current energy E; part: drachmanized_partial_doubles(n, K).
int drachmanized_values(const Ctx& c, const zd* cvec, const double* masses, double E, double* part, double* out,
                        hipStream_t s);
size_t drachmanized_partial_doubles(int n, int K);

// dense Hermitian linear algebra (ecg_linalg.hip)
struct Factor {
  int K = 0, ld = 0, cap = 0;
  zd* F = nullptr;      // ld x ld: L (strict lower part) and X^H = L^{-H} (strict upper part)
  double* D = nullptr;  // K pivots
  zd* W = nullptr;      // ld x 2 NB panel scratch (double-buffered by block column)
  zd* Dinv = nullptr;   // nblocks x 2 x NB x NB: inverse X of each diagonal block of L, then X^H
  int* status = nullptr;     // [0] negative pivots, [1] first zero pivot (1-based) or 0,
                             // [2] a device-side wait exceeded its time limit
  zd* vec = nullptr;    // 4 x ld work vectors
  zd* tvec = nullptr;   // ld: D^{-1} L^{-1} b between the two solve passes
  double* scal = nullptr;    // 16 scalars
  int* pflag = nullptr;      // sched_flags(ld) ready flags of a factorisation launch (ecg_schedule.hpp)
  int epoch = 0;             // flag value of the current factorisation launch
  zd* tp = nullptr;          // chunk partial sums of the tail update
  size_t tp_cap = 0;         // entries of tp
  Tracer* tr = nullptr;
};
// LDL^H of M = H - sigma S (f.K rows) into f; M's lower triangle is read
// from H and S on the first pass (H null: from f.F, already holding M).
int ldlh_factor(Factor& f, hipStream_t s, const zd* H = nullptr, const zd* S = nullptr, double sigma = 0.0);
// Source of a factorisation's first pass: M = H - sigma S from H and S, or
// (H null) from F itself.
struct ShiftSrc {
  const zd* H = nullptr;
  const zd* S = nullptr;
  double sigma = 0.0;
};
// One problem of a batched factorisation (the storage of its Factor).
struct FacDesc {
  int K = 0, ld = 0;
  zd* F = nullptr;
  double* D = nullptr;
  zd* Dinv = nullptr;
  zd* W = nullptr;
  int* status = nullptr;
  int* pflag = nullptr;
  ShiftSrc src;
};
// Factorisations of independent problems in shared launches (h on the
// host, largest K first; d its copy in device memory).  Flag values continue from epoch,
// which must exceed every flag value the problems' flag arrays hold; on
// return it is the last value used.
int ldlh_factor_batch(const std::vector<FacDesc>& h, const FacDesc* d, int& epoch, hipStream_t s);
// One problem of a batched inverse iteration at a fixed shift (the
// storage of its Ctx, Factor and context vectors).
struct InvDesc {
  int K = 0, ld = 0, warm = 0;   // warm: start from ev[2], else from ones
  const zd* H = nullptr;
  const zd* S = nullptr;
  double sigma = 0.0;
  const zd* F = nullptr;
  const double* D = nullptr;
  zd* ev[3] = {nullptr, nullptr, nullptr};
  zd* w = nullptr;        // S v
  zd* z = nullptr;        // D^{-1} L^{-1} S v
  zd* part = nullptr;     // the solve's and the quadratic forms' partial sums
  double* scal = nullptr;
  const int* status = nullptr;
  double* rep = nullptr;  // pinned host report
};
// nq inverse iterations (x = (H - sigma S)^{-1} S v, GHEPIIS's scaling and
// measure) of every problem, then the Rayleigh quotient, the normalisation
// x^H S x = 1 and the report, in shared launches; the result vector ends in
// ev[nq & 1] and ev[2].  Needs K <= 8192 (returns -1 otherwise).
int invit_batch(const std::vector<InvDesc>& h, const InvDesc* d, int nq, hipStream_t s);
// Complete the factor of H - sigma S from row m (r = K - m <= NB changed
// rows; rows < m factored at the same sigma and still in f).  Clears and
// sets f.status like ldlh_factor.  f.tp must hold tail_partial_entries(ld).
int ldlh_tail(const Ctx& c, Factor& f, double sigma, int m, hipStream_t s);
size_t tail_partial_entries(int ld);
int ldlh_solve(Factor& f, const zd* b, zd* x, hipStream_t s);
int invit_iteration(const Ctx& c, Factor& f, const zd* v, zd* x, hipStream_t s);
int rayleigh(const Ctx& c, Factor& f, const zd* x, double sigma, hipStream_t s);
int normalise(zd* x, int K, const double* t, hipStream_t s);

// packed multi-problem pair launches (ecg_kernels.hip): one descriptor per
// problem, problems of one launch share n; tile ranges ascending.
struct ProbDesc {
  SysConst sc;
  int K = 0, ld = 0, nt = 0, f0 = 0;     // f0: first function in the batch's concatenation
  long long tile0_off = 0, tile0_adj = 0;
  const double* bra = nullptr;
  const double* ket = nullptr;
  const double* diagS = nullptr;
  zd* H = nullptr;
  zd* S = nullptr;
  const zd* cvec = nullptr;
  double* part = nullptr;
  double* grad = nullptr;
  double E = 0.0;
};
int assemble_diag(const Ctx& c, hipStream_t s);   // diagonal pairs of all K functions
int packed_pairs(int n, bool adjoint, const ProbDesc* dpd, int np, long long ntiles, int nf, int maxT,
                 hipStream_t s);

// candidate screening (ecg_screen.hip; pair rows in ecg_kernels.hip)
// Candidate records, diagonal pairs and border rows (rows_only: the rows
// from records already made); cpb candidates per workgroup of the row
// kernel (1, 2, 4 or 8; cpb > 1 stages the shared ket records in LDS,
// cand_rows_lds bytes, 0 when they do not fit 64 KB).
int candidate_rows(const Ctx& c, int Kp, int C, const double* cparams, double* cbra, double* cket,
                   double* cdiag, zd* Hb, zd* Sb, int ldb, int cpb, hipStream_t s, bool rows_only = false);
size_t cand_rows_lds(int n, int T, int cpb);

// Device state of a screening session: functions 1..Kp (Kp = K-1) fixed,
// candidates replace function K.  M' = H' - sigma S' (leading Kp block),
// G = M'^{-1}, P = G S'.  Per-candidate columns are ld x ccap complex.
struct Screen {
  int Kp = 0, ld = 0, ccap = 0, scap = 0, nneg = 0;
  size_t callcap = 0;   // doubles in call
  double sigma = 0.0;
  bool ready = false;
  zd* G = nullptr;       // ld x ld
  zd* P = nullptr;       // ld x ld
  zd* Z = nullptr;       // ld x ld set-up scratch
  double* call = nullptr;    // npt x ccap x scap candidate parameters of a loop
  double* cparams = nullptr; // npt x ccap
  double* cbra = nullptr;
  double* cket = nullptr;
  double* cdiag = nullptr;   // 2 ccap
  zd* Hb = nullptr;      // ld x ccap border h
  zd* Bm = nullptr;      // ld x 2 ccap: [m | s] (m = h - sigma s)
  zd* Wm = nullptr;      // ld x 2 ccap: G [m | s]
  zd* V = nullptr;       // ld x ccap iterates (first Kp entries)
  zd* U = nullptr;       // ld x ccap
  zd* Y1 = nullptr;      // ld x ccap
  zd* Y2 = nullptr;      // ld x ccap
  zd* vK = nullptr;      // ccap last entries of the iterates
  double* cs = nullptr;  // 8 x ccap per-candidate scalars
  zd* v0 = nullptr;      // ld start vector (current eigenvector)
  double* state = nullptr;   // loop state: [0] current E, [1] step
  double* rec = nullptr;     // 2 x scap: best E and index per step
  int tm = 64;               // row tile of the per-step GEMMs (screen_tune)
  int cpb = 1;               // candidates per workgroup of the border rows (screen_tune)
  int tuned_c = 0;           // candidate count tm was timed for
  int tuned_kb = -1;         // and the basis size's 512-wide band
};
// 0 or a status; computes the factor of H - sigma S (whole basis), G and P.
int screen_prepare(const Ctx& c, Factor& f, Screen& sc, double sigma, hipStream_t s);
// Times the per-step GEMMs' row tiles for C candidates and keeps the faster
// in sc.tm (once per C; synchronises, so never inside a graph capture).
int screen_tune(const Ctx& c, Screen& sc, int C, hipStream_t s);
// One batch of C candidates (parameters already in sc.cparams, or, when
// from_loop, taken from sc.call at the step counter): rows, Schur
// complements, `iters` inverse iterations from sc.v0, Rayleigh quotients.
// Per-candidate results in sc.cs: [0] E, [1] convergence measure, [2] delta
// (Schur complement of the candidate's pivot), [3] x^H S x.
int screen_batch(const Ctx& c, const Factor& f, Screen& sc, int C, int iters, bool from_loop, int npt,
                 hipStream_t s);
// Loop bookkeeping kernel: pick the best candidate of the step, accept it
// when it lowers the current energy (parameters into function K of c,
// start vector := its eigenvector), record, advance the step counter.
int screen_select(const Ctx& c, Screen& sc, int C, int npt, hipStream_t s);
// C = alpha op(A) B + beta C (op = identity or A^H for trans = 1), M x N x
// Kd, column-major; tm = rows per workgroup tile (64 or 32); lower = 1 when
// op(A)'s columns and B's rows are zero above the diagonal (A^H B with
// both lower triangular), so each tile skips the leading terms that vanish.
int zgemm(int trans, int M, int N, int Kd, zd alpha, const zd* A, int lda, const zd* B, int ldb, zd beta, zd* Cm,
          int ldc, hipStream_t s, int tm = 64, int lower = 0);

}  // namespace ecg
