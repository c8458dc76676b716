// extern "C" entry points (include/ecg.h): context management, the drop-in
// replacements of MatrixElements / ComputeMatElem(AndDeriv) / GHEPIIS, and
// the device-resident energy and energy+gradient paths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/ecg.h"
#include "ecg_internal.hpp"
#include "ecg_schedule.hpp"

using ecg::zd;

#define HIPCHK(x)                              \
  do {                                         \
    hipError_t e_ = (x);                       \
    if (e_ != hipSuccess) return ECG_ERR_HIP;  \
  } while (0)
#define RCCHK(x)                      \
  do {                                \
    int r_ = (x);                     \
    if (r_ != 0) return r_ < 0 ? r_ : ECG_ERR_HIP; \
  } while (0)

struct ecg_ctx {
  ecg::Ctx c;
  ecg::Factor f;
  hipStream_t stream = nullptr;
  int dev = 0;                // the HIP device current when the context was created
  int npt = 0;
  double* part = nullptr;
  size_t part_cap = 0;
  double* grad = nullptr;
  size_t grad_cap = 0;
  zd* ev[3] = {nullptr, nullptr, nullptr};   // iteration vectors; ev[2] = last eigenvector
  bool have_last = false;     // ev[2] holds a vector of length last_K (warm start of inverse iteration)
  int last_K = 0;
  bool vec_current = false;   // ev[2] is the eigenvector of the current parameters (cleared by every
                              // call that changes them; the post-processing entries need it)
  bool params_set = false;    // the device parameters are the caller's current ones
  hipEvent_t e[8] = {};
  double tms[8] = {};
  double* pinned = nullptr;   // pinned host staging
  double* hrep = nullptr;     // pinned: the inverse iteration's report (k_report)
  size_t pinned_cap = 0;
  ecg::Screen scr;            // candidate screening session
  int scr_tm = 0, scr_cpb = 0;   // fixed screening shapes (0: timed by screen_tune)
  zd* dD = nullptr;           // stored derivatives (2 npt x nfo x K) of ecg_energy_grad_stored
  size_t dD_cap = 0;
  ecg::Tracer tr;             // per-kernel tracing (off by default)
  // overlap penalty settings (Glob_OverlapPenaltyAllowed, Glob_MaxOverlapPenalty,
  // Glob_OverlapPenaltyThreshold2, and Glob_nfru for energy-only calls)
  bool pen_on = false;
  double pen_b = 1.0, pen_t2 = 1.0;
  int pen_nfru = 0;
  double* pen = nullptr;      // per-function penalty partials, then the total (device)
  size_t pen_cap = 0;
  double pen_total = 0.0;     // host copy of the last total
  double* xpart = nullptr;    // expectation-value tile partials, then the 2 NOBS results
  size_t xpart_cap = 0;
  bool s_ready = false;       // the device S is assembled from the current parameters
  // Incremental recomputation.  hs_rows leading rows of the device H, S
  // match the current parameters; fac_rows leading rows of the factor
  // (L, D, X^H in f) belong to the device H - fac_sigma S.  Any write that
  // may change rows >= r of H or S lowers fac_rows to r.  hparams is the
  // host copy of the parameters last uploaded (empty when unknown).
  std::vector<double> hparams;
  int hs_rows = 0;
  int fac_rows = 0;
  double fac_sigma = 0.0;
  int last_tail = -1;         // rows of the last tail update, -1 after a full factorisation
};

// A packed batch of independent problems (one context each).
struct ecg_batch {
  std::vector<ecg_ctx*> ctx;
  hipStream_t lead = nullptr;
  hipEvent_t ev_asm = nullptr;
  std::vector<hipEvent_t> ev_eig;
  struct Class {
    int n = 0, maxT = 0, nf = 0;
    std::vector<int> probs;
    std::vector<ecg::ProbDesc> h;
    ecg::ProbDesc* d = nullptr;
    long long ntiles = 0;
  };
  std::vector<Class> classes;
  int packed = 1;     // 2: pair phases packed and factorisations batched; 1: pair phases packed; 0: per problem
  int nstreams = 4;   // problems whose eigen steps run concurrently
  hipEvent_t ev_fac = nullptr;        // batched factorisations queued on the lead stream
  std::vector<ecg::FacDesc> fh;       // their descriptors (host copy, alive until the upload ran)
  ecg::FacDesc* fd = nullptr;
  std::vector<ecg::InvDesc> ih;       // the batched inverse iterations' descriptors
  ecg::InvDesc* id = nullptr;
};

namespace {

template <class T>
int dev_alloc(T** p, size_t n) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (n == 0) return 0;
  if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return ECG_ERR_NOMEM;
  }
  return 0;
}

// Enqueue the overlap penalty of the device S (pairs j > i, j > nfru) on
// the context's stream, optionally adding its gradient to dgrad, and the
// copy of the total to x->pen_total (valid once the stream synchronises).
int penalty_enqueue(ecg_ctx* x, hipStream_t s, int nfru, double b, double t2, double* dgrad) {
  const int K = x->c.K;
  x->pen_total = 0.0;
  if (nfru >= K) return 0;
  const size_t need = (size_t)(K - nfru) + 1;
  if (need > x->pen_cap) {
    RCCHK(dev_alloc(&x->pen, need));
    x->pen_cap = need;
  }
  double* total = x->pen + (K - nfru);
  if (ecg::overlap_penalty(x->c, nfru, b / (1.0 - t2), t2, x->pen, total, dgrad, s)) return ECG_ERR_HIP;
  HIPCHK(hipMemcpyAsync(&x->pen_total, total, sizeof(double), hipMemcpyDeviceToHost, s));
  return 0;
}

// Rows >= row of the device H, S no longer match the parameters.
void rows_changed(ecg_ctx* x, int row) { x->hs_rows = std::min(x->hs_rows, row); }

// The parameters changed: the last eigenvector no longer belongs to them
// (it stays as a warm start of the next inverse iteration), and a screening
// session built on the old ones is closed.
void params_changed(ecg_ctx* x) {
  x->vec_current = false;
  x->scr.ready = false;
}

// H, S rows >= row are about to be (re)written with values that may differ.
void hs_rewritten(ecg_ctx* x, int row) { x->fac_rows = std::min(x->fac_rows, row); }

// First function whose parameters differ from the last upload (the
// shorter length when one list extends the other).
int first_changed(const ecg_ctx* x, int K, const double* params) {
  const size_t npt = (size_t)x->npt;
  const int n = std::min(K, (int)(x->hparams.size() / npt));
  for (int j = 0; j < n; ++j)
    if (std::memcmp(x->hparams.data() + (size_t)j * npt, params + (size_t)j * npt, npt * sizeof(double)) != 0)
      return j;
  return n;
}

int ensure_capacity(ecg_ctx* x, int K) {
  ecg::Ctx& c = x->c;
  if (K <= c.cap) {
    c.K = K;
    x->f.K = K;
    x->hs_rows = std::min(x->hs_rows, K);
    x->fac_rows = std::min(x->fac_rows, K);
    return 0;
  }
  // Growing: a quarter of headroom (a basis that grows one function at a
  // time reallocates rarely), and the leading rows of H, S, diagS, diagH
  // and of the factor that are still valid move to the new layout, so
  // that the next call continues from them.
  const int keep = (c.cap > 0) ? std::min(c.K, std::max(x->hs_rows, x->fac_rows)) : 0;
  const int want = std::max(std::max(K, 64), c.cap > 0 ? c.cap + c.cap / 4 : 0);
  const int ld = (want + 63) / 64 * 64;
  const int cap = ld;
  const int old_ld = c.ld;
  zd* oH = c.H;
  zd* oS = c.S;
  zd* oF = x->f.F;
  double* oD = x->f.D;
  double* odS = c.diagS;
  double* odH = c.diagH;
  c.H = c.S = x->f.F = nullptr;
  x->f.D = c.diagS = c.diagH = nullptr;
  auto release_old = [&]() {
    for (void* q : {(void*)oH, (void*)oS, (void*)oF, (void*)oD, (void*)odS, (void*)odH})
      if (q) (void)hipFree(q);
  };
  int rc = 0;
  ecg::Factor& f = x->f;
  if (!rc) rc = dev_alloc(&c.params, (size_t)x->npt * cap);
  if (!rc) rc = dev_alloc(&c.bra, (size_t)ecg::bra_stride(c.n) * cap);
  if (!rc) rc = dev_alloc(&c.ket, (size_t)ecg::ket_stride(c.n) * c.T * cap);
  if (!rc) rc = dev_alloc(&c.diagS, (size_t)cap);
  if (!rc) rc = dev_alloc(&c.diagH, (size_t)cap);
  if (!rc) rc = dev_alloc(&c.H, (size_t)ld * ld);
  if (!rc) rc = dev_alloc(&c.S, (size_t)ld * ld);
  if (!rc) rc = dev_alloc(&c.rowbuf, ecg::thin_row_entries(cap));
  if (!rc) rc = dev_alloc(&c.adjraw, ecg::thin_adjoint_entries(c.n, cap));
  if (!rc) rc = dev_alloc(&f.F, (size_t)ld * ld);
  if (!rc) rc = dev_alloc(&f.D, (size_t)ld);
  if (!rc) rc = dev_alloc(&f.W, (size_t)ld * 2 * ecg::NB);
  if (!rc) rc = dev_alloc(&f.Dinv, (size_t)(ld / ecg::NB) * 2 * ecg::NB * ecg::NB);
  if (!rc) rc = dev_alloc(&f.status, (size_t)4);
  if (!rc) rc = dev_alloc(&f.vec, (size_t)4 * ld);
  if (!rc) rc = dev_alloc(&f.tvec, (size_t)ld);
  if (!rc) rc = dev_alloc(&f.pflag, (size_t)ecg::sched_flags(ld));
  if (!rc) rc = dev_alloc(&f.tp, (size_t)0);   // tail-update partials: allocated on first use
  if (!rc) rc = dev_alloc(&f.scal, (size_t)16);
  for (int i = 0; i < 3 && !rc; ++i) rc = dev_alloc(&x->ev[i], (size_t)ld);
  if (rc) {
    release_old();
    c.cap = c.K = 0;
    x->hs_rows = x->fac_rows = 0;
    x->hparams.clear();
    return rc;
  }
  f.tp_cap = 0;
  f.epoch = 0;
  bool ok = hipMemsetAsync(f.pflag, 0, sizeof(int) * ecg::sched_flags(ld), x->stream) == hipSuccess &&
            hipMemsetAsync(f.status, 0, sizeof(int) * 4, x->stream) == hipSuccess &&
            hipMemsetAsync(c.H, 0, sizeof(zd) * ld * ld, x->stream) == hipSuccess &&
            hipMemsetAsync(c.S, 0, sizeof(zd) * ld * ld, x->stream) == hipSuccess;
  if (ok && keep > 0) {
    const size_t w = sizeof(zd) * keep, sp = sizeof(zd) * old_ld, dp = sizeof(zd) * ld;
    ok = hipMemcpy2DAsync(c.H, dp, oH, sp, w, keep, hipMemcpyDeviceToDevice, x->stream) == hipSuccess &&
         hipMemcpy2DAsync(c.S, dp, oS, sp, w, keep, hipMemcpyDeviceToDevice, x->stream) == hipSuccess &&
         hipMemcpy2DAsync(f.F, dp, oF, sp, w, keep, hipMemcpyDeviceToDevice, x->stream) == hipSuccess &&
         hipMemcpyAsync(f.D, oD, sizeof(double) * keep, hipMemcpyDeviceToDevice, x->stream) == hipSuccess &&
         hipMemcpyAsync(c.diagS, odS, sizeof(double) * keep, hipMemcpyDeviceToDevice, x->stream) == hipSuccess &&
         hipMemcpyAsync(c.diagH, odH, sizeof(double) * keep, hipMemcpyDeviceToDevice, x->stream) == hipSuccess;
  }
  ok = ok && hipStreamSynchronize(x->stream) == hipSuccess;
  release_old();
  c.cap = cap;
  c.ld = ld;
  c.K = K;
  f.cap = cap;
  f.ld = ld;
  f.K = K;
  x->have_last = false;
  x->vec_current = false;
  x->params_set = false;   // params/bra/ket were reallocated: the caller uploads them next
  x->hs_rows = ok ? std::min(x->hs_rows, keep) : 0;
  x->fac_rows = ok ? std::min(x->fac_rows, keep) : 0;
  if (!ok) x->hparams.clear();
  return ok ? 0 : ECG_ERR_HIP;
}

int ensure_pinned(ecg_ctx* x, size_t doubles) {
  if (doubles <= x->pinned_cap) return 0;
  if (x->pinned) (void)hipHostFree(x->pinned);
  x->pinned = nullptr;
  x->pinned_cap = 0;
  if (hipHostMalloc((void**)&x->pinned, doubles * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return ECG_ERR_NOMEM;
  x->pinned_cap = doubles;
  return 0;
}

// Screening buffers: ld x ld matrices for the current basis size, C
// candidate columns, nsteps loop steps.
int ensure_screen(ecg_ctx* x, int C, int nsteps) {
  ecg::Screen& sc = x->scr;
  const ecg::Ctx& c = x->c;
  const int ld = c.ld;
  if (sc.ld != ld) {
    RCCHK(dev_alloc(&sc.G, (size_t)ld * ld));
    RCCHK(dev_alloc(&sc.P, (size_t)ld * ld));
    RCCHK(dev_alloc(&sc.Z, (size_t)ld * ld));
    RCCHK(dev_alloc(&sc.v0, (size_t)ld));
    RCCHK(dev_alloc(&sc.state, (size_t)2));
    sc.ld = ld;
    sc.ccap = 0;
    sc.ready = false;
  }
  if (C > sc.ccap) {
    const int cc = std::max(C, 16);
    RCCHK(dev_alloc(&sc.cparams, (size_t)x->npt * cc));
    RCCHK(dev_alloc(&sc.cbra, (size_t)ecg::bra_stride(c.n) * cc));
    RCCHK(dev_alloc(&sc.cket, (size_t)ecg::ket_stride(c.n) * c.T * cc));
    RCCHK(dev_alloc(&sc.cdiag, (size_t)2 * cc));
    RCCHK(dev_alloc(&sc.Hb, (size_t)ld * cc));
    RCCHK(dev_alloc(&sc.Bm, (size_t)ld * 2 * cc));
    RCCHK(dev_alloc(&sc.Wm, (size_t)ld * 2 * cc));
    RCCHK(dev_alloc(&sc.V, (size_t)ld * cc));
    RCCHK(dev_alloc(&sc.U, (size_t)ld * cc));
    RCCHK(dev_alloc(&sc.Y1, (size_t)ld * cc));
    RCCHK(dev_alloc(&sc.Y2, (size_t)ld * cc));
    RCCHK(dev_alloc(&sc.vK, (size_t)cc));
    RCCHK(dev_alloc(&sc.cs, (size_t)8 * cc));
    sc.ccap = cc;
  }
  if (nsteps > 0 && (size_t)x->npt * C * nsteps > sc.callcap) {
    RCCHK(dev_alloc(&sc.call, (size_t)x->npt * C * nsteps));
    sc.callcap = (size_t)x->npt * C * nsteps;
  }
  if (nsteps > sc.scap) {
    RCCHK(dev_alloc(&sc.rec, (size_t)2 * nsteps));
    sc.scap = nsteps;
  }
  return 0;
}

void free_screen(ecg::Screen& sc) {
  void* ptrs[] = {sc.G, sc.P, sc.Z, sc.call, sc.cparams, sc.cbra, sc.cket, sc.cdiag, sc.Hb, sc.Bm, sc.Wm,
                  sc.V, sc.U, sc.Y1, sc.Y2, sc.vK, sc.cs, sc.v0, sc.state, sc.rec};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  sc = ecg::Screen();
}

// Upload parameters and run the per-function set-up for all K functions.
int upload_params(ecg_ctx* x, int K, const double* params) {
  RCCHK(ensure_capacity(x, K));
  HIPCHK(hipMemcpyAsync(x->c.params, params, sizeof(double) * x->npt * K, hipMemcpyHostToDevice, x->stream));
  RCCHK(ecg::prep(x->c, x->stream));
  return 0;
}

void trace_collect(ecg_ctx* x) {
  if (x->tr.on) x->tr.collect();
}

bool finite_all(const double* v, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

// Copy device H, S (normalised, both triangles) and diagH/diagS to the
// caller's arrays with the reference's StoreHS layout for rows r0..r1-1.
int copy_out_hs(ecg_ctx* x, int r0, int r1, char method, double sigma, double* H, double* S, int ld,
                double* diagH, double* diagS) {
  ecg::Ctx& c = x->c;
  const int K = c.K;
  const size_t nd = (size_t)2 * c.ld * K;
  RCCHK(ensure_pinned(x, 2 * nd + 2 * (size_t)K));
  double* hH = x->pinned;
  double* hS = x->pinned + nd;
  double* hdH = x->pinned + 2 * nd;
  double* hdS = hdH + K;
  HIPCHK(hipMemcpyAsync(hH, c.H, sizeof(double) * nd, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(hS, c.S, sizeof(double) * nd, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(hdH, c.diagH, sizeof(double) * K, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(hdS, c.diagS, sizeof(double) * K, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  for (int k = r0; k < r1; ++k) {
    if (!std::isfinite(hdS[k]) || !std::isfinite(hdH[k])) return ECG_ERR_NONFINITE;
    diagS[k] = hdS[k];
    if (method == 'G') {
      diagH[k] = hdH[k];
    } else {
      const size_t o = 2 * ((size_t)k * ld + k);
      S[o] = 1.0;
      S[o + 1] = 0.0;
      H[o] = hdH[k] - sigma;
      H[o + 1] = 0.0;
    }
    for (int l = 0; l < k; ++l) {
      const size_t od = 2 * ((size_t)l * c.ld + k);
      const size_t oh = 2 * ((size_t)l * ld + k);
      const double hr = hH[od], hi = hH[od + 1], sr = hS[od], si = hS[od + 1];
      if (!std::isfinite(hr) || !std::isfinite(hi) || !std::isfinite(sr) || !std::isfinite(si))
        return ECG_ERR_NONFINITE;
      S[oh] = sr;
      S[oh + 1] = si;
      if (method == 'G') {
        H[oh] = hr;
        H[oh + 1] = hi;
      } else {
        H[oh] = hr - sigma * sr;
        H[oh + 1] = hi - sigma * si;
        const size_t ou = 2 * ((size_t)k * ld + l);
        S[ou] = sr;
        S[ou + 1] = -si;
      }
    }
  }
  return 0;
}

// Factor H - sigma S on the device; returns the negative-pivot count and
// the first zero pivot (0 if none).
int factor_at(ecg_ctx* x, double sigma, int* nneg, int* zero) {
  x->fac_rows = 0;
  RCCHK(ecg::ldlh_factor(x->f, x->stream, x->c.H, x->c.S, sigma));
  int st[3];
  HIPCHK(hipMemcpyAsync(st, x->f.status, sizeof(st), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (st[2]) return ECG_ERR_SYNC;   // a device-side wait of the factorisation hit its limit
  *nneg = st[0];
  *zero = st[1];
  if (!st[1]) {
    x->fac_rows = x->c.K;
    x->fac_sigma = sigma;
  }
  return 0;
}

// Queue the factor of H - sigma S: from the first row that differs from
// the last factorisation at the same sigma when at most NB rows do (the
// reference's LDLHF from row Nmin), otherwise in full.
int factor_queue(ecg_ctx* x, double sigma) {
  const int K = x->c.K, m = x->fac_rows;
  ecg::Factor& f = x->f;
  const bool tail = m >= 1 && K - m <= ecg::NB && x->fac_sigma == sigma;
  x->fac_rows = 0;
  if (tail) {
    const size_t need = ecg::tail_partial_entries(f.ld);
    if (f.tp_cap < need) {
      RCCHK(dev_alloc(&f.tp, need));
      f.tp_cap = need;
    }
    RCCHK(ecg::ldlh_tail(x->c, f, sigma, m, x->stream));
  } else {
    RCCHK(ecg::ldlh_factor(f, x->stream, x->c.H, x->c.S, sigma));
  }
  x->fac_sigma = sigma;
  x->last_tail = tail ? K - m : -1;
  return 0;
}

struct InvitResult {
  double lambda = 0.0, nd = 0.0, t1 = 0.0;
  int iters = 0;
  bool converged = false;
};

__global__ void k_fill_ones(int K, zd* v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < K) v[i] = ecg::mkz(1.0, 0.0);
}

// The iteration's scalars and the factorisation status into pinned host
// memory (one small kernel instead of two copies to pageable memory):
// rep[0..4] = scal[0..4], rep[5..7] = status[0..2]; the Rayleigh quotient
// E = scal[3] / scal[2] + sigma (the host's formula, same operations) goes
// to rep[8] and to scal[8] on the device, where a contraction queued
// before the host has read it takes it from.
__global__ void k_report(double* scal, const int* status, double sigma, double* rep) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double E = scal[3] / scal[2] + sigma;
  scal[8] = E;
  for (int i = 0; i < 5; ++i) __hip_atomic_store(&rep[i], scal[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  for (int i = 0; i < 3; ++i)
    __hip_atomic_store(&rep[5 + i], (double)status[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&rep[8], E, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __threadfence_system();
}

// Inverse iteration with the current factorisation (GHEPIIS loop,
// src/linalg.f90:1892-1917).  Start vector in x->ev[0] unless warm.
// Stops when the GHEPIIS measure reaches tol (tol > 0), after maxiter
// iterations, and:
//   * without stop_at_floor and tol <= 0 (GHEPIIS's own floor rule,
//     src/linalg.f90:1902-1906): converged at the first iteration whose
//     measure is larger than the previous one and <= |tol|;
//   * with stop_at_floor (the energy paths): when the measure has stopped
//     halving below 1e-9 (the rounding floor; success is then decided by
//     the caller against |tol|).
// With stop_at_floor (the energy paths) the first nspec iterations, the
// Rayleigh quotient and the normalisation are queued with a single host
// synchronisation; further iterations only run if the measure asks for
// them.  Result vector (normalised x^H S x = 1) in x->ev[2].
// before_sync (optional) is run once, at the first synchronisation, after
// the result vector and k_report's scalars are queued; *hooked reports
// whether the result returned is the vector of that synchronisation.
int inverse_iteration(ecg_ctx* x, double sigma, double tol, int maxiter, bool warm, InvitResult& r,
                      bool stop_at_floor, int* fst = nullptr, int nspec = 3,
                      const std::function<int()>* before_sync = nullptr, bool* hooked = nullptr) {
  ecg::Ctx& c = x->c;
  const int K = c.K;
  zd* v = x->ev[0];
  zd* y = x->ev[1];
  if (warm && x->have_last && x->last_K == K)
    HIPCHK(hipMemcpyAsync(v, x->ev[2], sizeof(zd) * K, hipMemcpyDeviceToDevice, x->stream));
  else
    hipLaunchKernelGGL(k_fill_ones, dim3((K + 255) / 256), dim3(256), 0, x->stream, K, v);
  double prev = std::numeric_limits<double>::infinity();
  r = InvitResult();
  // tail: Rayleigh quotient, normalisation, result copy, host read of the
  // scalars and the status (one synchronisation)
  double q[5];
  int st[4] = {0, 0, 0, 0};
  if (hooked) *hooked = false;
  bool hook_done = false;
  auto read_report = [&]() {
    for (int i = 0; i < 5; ++i) q[i] = x->hrep[i];
    for (int i = 0; i < 3; ++i) st[i] = (int)x->hrep[5 + i];
  };
  auto tail = [&]() -> int {
    RCCHK(ecg::rayleigh(c, x->f, v, sigma, x->stream));
    RCCHK(ecg::normalise(v, K, x->f.scal + 2, x->stream));
    hipLaunchKernelGGL(k_report, dim3(1), dim3(64), 0, x->stream, x->f.scal, x->f.status, sigma, x->hrep);
    HIPCHK(hipMemcpyAsync(x->ev[2], v, sizeof(zd) * K, hipMemcpyDeviceToDevice, x->stream));
    if (before_sync && !hook_done) {
      RCCHK((*before_sync)());
      hook_done = true;
      if (hooked) *hooked = true;
    } else if (hooked) {
      *hooked = false;   // a later result: what the hook queued is stale
    }
    HIPCHK(hipStreamSynchronize(x->stream));
    read_report();
    if (st[2]) return ECG_ERR_SYNC;   // a device-side wait (factor or solve) hit its limit
    if (fst)
      for (int i = 0; i < 3; ++i) fst[i] = st[i];
    return 0;
  };
  int it = 0;
  bool done = false;
  if (stop_at_floor) {
    // energy paths: the first (up to) nspec iterations and the tail are
    // queued without a host round trip; the loop below continues only when
    // the measure asks for more iterations
    const int nq = std::min(maxiter, nspec);
    for (; it < nq; ++it) {
      RCCHK(ecg::invit_iteration(c, x->f, v, y, x->stream));
      std::swap(v, y);
    }
    RCCHK(tail());
    if (fst && st[1]) return 0;   // zero pivot: the caller reports it
    r.iters = nq;
    r.nd = q[0];
    if (!std::isfinite(r.nd)) return ECG_ERR_NONFINITE;
    const double p1 = (nq >= 2) ? q[4] : std::numeric_limits<double>::infinity();   // previous measure
    if (tol > 0.0 && r.nd <= tol) {
      r.converged = true;
      done = true;
    } else if (nq >= 2 && r.nd < 1e-9 && r.nd > 0.5 * p1) {
      done = true;
    } else if (nq == maxiter) {
      done = true;
    }
    prev = r.nd;
  }
  if (!done) {
    for (; it < maxiter; ++it) {
      RCCHK(ecg::invit_iteration(c, x->f, v, y, x->stream));
      hipLaunchKernelGGL(k_report, dim3(1), dim3(64), 0, x->stream, x->f.scal, x->f.status, sigma, x->hrep);
      HIPCHK(hipStreamSynchronize(x->stream));
      read_report();
      if (fst) {   // the factorisation's status, read here instead of right after it
        for (int i = 0; i < 3; ++i) fst[i] = st[i];
        if (st[2]) return ECG_ERR_SYNC;
        if (st[1]) return 0;   // zero pivot: the caller reports it
      }
      r.nd = q[0];
      r.iters = it + 1;
      std::swap(v, y);   // v now holds the newest iterate
      if (!std::isfinite(r.nd)) return ECG_ERR_NONFINITE;
      if (tol > 0.0 ? (r.nd <= tol) : (!stop_at_floor && r.nd > prev && r.nd <= std::fabs(tol))) {
        r.converged = true;
        break;
      }
      if (stop_at_floor && it >= 1 && r.nd < 1e-9 && r.nd > 0.5 * prev) break;
      prev = r.nd;
    }
    RCCHK(tail());
    if (fst && st[1]) return 0;
  }
  r.t1 = q[2];
  r.lambda = q[3] / q[2] + sigma;
  x->have_last = true;
  x->last_K = K;
  return 0;
}

// Root `which` of (H, S) for the assembled device matrices.  With a finite
// sigma, inverse iteration at that shift (method 'I').  Otherwise the shift
// is located from inertia counts (number of negative pivots of
// LDL^H(H - s S) = number of eigenvalues below s), then refined with the
// Rayleigh quotient.
// before_sync / hooked: as for inverse_iteration, in the fixed-shift path
// only (*hooked false otherwise).
int solve_root(ecg_ctx* x, int which, double sigma, double tol, int maxiter, double* E, double* info,
               int* status, const std::function<int()>* before_sync = nullptr, bool* hooked = nullptr,
               bool prefactored = false) {
  if (hooked) *hooked = false;
  const double sigma_in = sigma;
  ecg::Ctx& c = x->c;
  const int K = c.K;
  int nfac = 0, nneg = 0, zero = 0;
  InvitResult r;
  *status = 0;
  if (std::isfinite(sigma)) {
    // factorisation queued without a host round trip; its status is read
    // at the inverse iteration's first synchronisation (prefactored: a
    // batched factorisation of H - sigma S already queued on this stream)
    if (prefactored) {
      x->fac_rows = 0;
      x->fac_sigma = sigma;
      x->last_tail = -1;
    } else {
      RCCHK(factor_queue(x, sigma));
    }
    ++nfac;
    int fst[3] = {0, 0, 0};
    // after a tail update the previous eigenvector is close: two queued
    // iterations usually meet the criterion (more follow if not)
    const bool near = x->last_tail >= 0 && x->have_last && x->last_K == K;
    const int rc = inverse_iteration(x, sigma, tol, maxiter, true, r, true, fst, near ? 2 : 3, before_sync, hooked);
    if (rc == ECG_ERR_SYNC) return rc;
    if (fst[1]) {
      *status = ECG_SINGULAR;
      return 0;
    }
    RCCHK(rc);
    x->fac_rows = K;   // the factor is complete and without zero pivots
    nneg = fst[0];
  } else {
    std::vector<double> dh(K);
    HIPCHK(hipMemcpyAsync(dh.data(), c.diagH, sizeof(double) * K, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    const double s0 = *std::min_element(dh.begin(), dh.end());
    const double scale = std::max(1.0, std::fabs(s0));
    auto count = [&](double s, int* nn) -> int {
      int zz = 0;
      RCCHK(factor_at(x, s, nn, &zz));
      ++nfac;
      if (zz) {   // exactly singular: nudge the shift
        RCCHK(factor_at(x, s * (1.0 + 1e-13) + 1e-300, nn, &zz));
        ++nfac;
      }
      return 0;
    };
    double lo, hi;
    int nlo, nhi;
    int n0;
    RCCHK(count(s0, &n0));
    double step = 1e-2 * scale;
    if (n0 >= which) {
      hi = s0;
      nhi = n0;
      for (;;) {
        double s = s0 - step;
        int nn;
        RCCHK(count(s, &nn));
        if (nn <= which - 1) {
          lo = s;
          nlo = nn;
          break;
        }
        hi = s;
        nhi = nn;
        step *= 4.0;
        if (nfac > 60) {
          *status = ECG_ROOT_NOT_FOUND;
          return 0;
        }
      }
    } else {
      lo = s0;
      nlo = n0;
      for (;;) {
        double s = s0 + step;
        int nn;
        RCCHK(count(s, &nn));
        if (nn >= which) {
          hi = s;
          nhi = nn;
          break;
        }
        lo = s;
        nlo = nn;
        step *= 4.0;
        if (nfac > 60) {
          *status = ECG_ROOT_NOT_FOUND;
          return 0;
        }
      }
    }
    // isolate: nlo == which-1 and nhi == which; for which > 1 also shrink
    // the bracket so that lo is nearer to E_which than to E_{which-1}.
    const double width_goal = (which == 1) ? HUGE_VAL : 1e-6 * scale;
    while (!(nlo == which - 1 && nhi == which) || (hi - lo) > width_goal) {
      const double mid = 0.5 * (lo + hi);
      int nn;
      RCCHK(count(mid, &nn));
      if (nn <= which - 1) {
        lo = mid;
        nlo = nn;
      } else {
        hi = mid;
        nhi = nn;
      }
      if (nfac > 80) {
        *status = ECG_ROOT_NOT_FOUND;
        return 0;
      }
    }
    // Refinement.  For the lowest root the Rayleigh quotient is an upper
    // bound, so the bracket (s, hi) shrinks from both sides: iterate a few
    // steps at s (below the root), lower hi to the quotient, and move s up
    // to just below hi whenever the inertia confirms it is still below the
    // root.  For higher roots the bisection above already made the bracket
    // narrow; iterate at its lower end.
    double s = lo;
    double fac_at = NAN;
    auto refactor = [&](double at) -> int {
      if (fac_at == at) return 0;
      int nn, zz;
      RCCHK(factor_at(x, at, &nn, &zz));
      ++nfac;
      nneg = nn;
      fac_at = at;
      return 0;
    };
    auto probe = [&](double at, int* nn) -> int {
      RCCHK(count(at, nn));
      fac_at = at;
      nneg = *nn;
      return 0;
    };
    if (which == 1) {
      bool warm = false;
      for (int ref = 0; ref < 8; ++ref) {
        RCCHK(refactor(s));
        RCCHK(inverse_iteration(x, s, tol, 3, warm, r, false));
        warm = true;
        if (r.converged) break;
        hi = std::min(hi, r.lambda);
        if (!(hi > s)) break;
        double cand = hi - 1e-3 * (hi - s);
        int nn;
        RCCHK(probe(cand, &nn));
        if (nn == which - 1) {
          s = cand;
          continue;
        }
        hi = cand;
        cand = 0.5 * (s + hi);
        RCCHK(probe(cand, &nn));
        if (nn == which - 1) s = cand;
        else hi = cand;
      }
    }
    RCCHK(refactor(s));
    if (!r.converged) RCCHK(inverse_iteration(x, s, tol, maxiter, which == 1, r, true));
    sigma = s;
  }
  *E = r.lambda;
  // tol <= 0 is GHEPIIS's floor mode: success when the measure at the
  // rounding floor is <= |tol| (src/linalg.f90:1902-1906)
  if (!r.converged && r.nd > std::fabs(tol)) *status = ECG_NOT_CONVERGED;
  if (info) {
    info[0] = r.iters;
    info[1] = r.nd;
    info[2] = sigma;
    info[3] = nfac;
    info[4] = nneg;
    info[5] = r.t1;
    info[7] = std::isfinite(sigma_in) ? x->last_tail : -1;
  }
  return 0;
}

// Milliseconds between two events, or -1 if either was not recorded in
// this call (e.g. the assembly events when no row needed reassembly); the
// failed query's error is cleared so that it does not reach a later check.
float ev_ms(hipEvent_t a, hipEvent_t b) {
  float t = 0.0f;
  if (hipEventElapsedTime(&t, a, b) != hipSuccess) {
    (void)hipGetLastError();
    return -1.0f;
  }
  return t;
}

template <int N>
__global__ void k_one_pair(const ecg::SysConst sc, const double* bra, const double* ket, double* out) {
  constexpr int m = N * (N + 1) / 2;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double* bk = bra;                                   // function 0 = k (bra)
  const double* bl = bra + ecg::BraRec<N>::stride;          // function 1 = l (ket)
  const double* kl = ket + (size_t)1 * sc.T * ecg::KetRec<N>::stride;
  const zd* ck = reinterpret_cast<const zd*>(bk + ecg::BraRec<N>::ck);
  const zd* ct = reinterpret_cast<const zd*>(kl + ecg::KetRec<N>::ct);
  ecg::PairCore<N> pc;
  ecg::pair_energy<N>(sc, ck, ct, reinterpret_cast<const zd*>(kl + ecg::KetRec<N>::nl),
                      bk[ecg::BraRec<N>::adet] * bl[ecg::BraRec<N>::adet], pc);
  zd gb[m], gk[m], gs[m], g2[m], s2[m];
  ecg::pair_adjoint<N>(sc, ck, ct, pc, 0.0, gb, gk);
  for (int e = 0; e < m; ++e) gs[e] = (-1.5) * (pc.S * pc.Ci[e]);
  ecg::perm_congruence<N>(sc.P[0], gk, g2);
  ecg::perm_congruence<N>(sc.P[0], gs, s2);
  const zd H = pc.T + pc.V;
  zd* o = reinterpret_cast<zd*>(out);
  o[0] = H;
  o[1] = pc.S;
  zd* Dk = o + 2;
  zd* Dl = o + 2 + 4 * m;
  const zd aH = 1.5 * H, aS = 1.5 * pc.S;
  ecg::chain_to_params<N>(gb, aH, bk + ecg::BraRec<N>::lp, -1.0, Dk, Dk + m);
  ecg::chain_to_params<N>(gs, aS, bk + ecg::BraRec<N>::lp, -1.0, Dk + 2 * m, Dk + 3 * m);
  ecg::chain_to_params<N>(g2, aH, bl + ecg::BraRec<N>::lp, 1.0, Dl, Dl + m);
  ecg::chain_to_params<N>(s2, aS, bl + ecg::BraRec<N>::lp, 1.0, Dl + 2 * m, Dl + 3 * m);
}

int launch_one_pair(const ecg::Ctx& c, const ecg::SysConst& sc, const double* bra, const double* ket, double* out,
                    hipStream_t s) {
  switch (c.n) {
    case 1: hipLaunchKernelGGL(k_one_pair<1>, dim3(1), dim3(64), 0, s, sc, bra, ket, out); break;
    case 2: hipLaunchKernelGGL(k_one_pair<2>, dim3(1), dim3(64), 0, s, sc, bra, ket, out); break;
    case 3: hipLaunchKernelGGL(k_one_pair<3>, dim3(1), dim3(64), 0, s, sc, bra, ket, out); break;
    case 4: hipLaunchKernelGGL(k_one_pair<4>, dim3(1), dim3(64), 0, s, sc, bra, ket, out); break;
    case 5: hipLaunchKernelGGL(k_one_pair<5>, dim3(1), dim3(64), 0, s, sc, bra, ket, out); break;
    default: return ECG_ERR_ARG;
  }
  return (int)hipGetLastError();
}

// One pair on the device (MatrixElements): the two parameter columns, their
// bra/ket records and the kernel's output are staged in dev (at least
// ME_STAGE_DOUBLES doubles, 256-byte aligned), which the caller provides.
constexpr int ME_OFF_PARAMS = 0, ME_OFF_BRA = 64, ME_OFF_KET = 192, ME_OFF_OUT = 384;
constexpr size_t ME_STAGE_DOUBLES = 640;   // n <= 5: 60 + 92 + 160 + 244 doubles in 256-byte slots
static_assert(2 * 30 <= ME_OFF_BRA - ME_OFF_PARAMS && 2 * 46 <= ME_OFF_KET - ME_OFF_BRA &&
                  2 * 80 <= ME_OFF_OUT - ME_OFF_KET && 2 * (2 + 8 * 15) <= (int)ME_STAGE_DOUBLES - ME_OFF_OUT,
              "staging slots too small for n = 5");

int matrix_elements_run(ecg_ctx* x, const double* paramk, const double* paraml, const int* P, int gradflag,
                        double* Hkl, double* Skl, double* Dk, double* Dl, double* dev) {
  const int n = x->c.n, m = n * (n + 1) / 2, npt = x->npt;
  const bool gk = (gradflag == 1 || gradflag == 3), gl = (gradflag == 2 || gradflag == 3);
  ecg::SysConst sc = x->c.sc;
  sc.T = 1;
  sc.coef[0] = 1.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) sc.P[0][i][j] = (signed char)P[i + n * j];
  // private 2-function basis on the device
  ecg::Ctx c1;
  c1.sc = sc;
  c1.n = n;
  c1.T = 1;
  c1.K = 2;
  c1.params = dev + ME_OFF_PARAMS;
  c1.bra = dev + ME_OFF_BRA;
  c1.ket = dev + ME_OFF_KET;
  double* dout = dev + ME_OFF_OUT;
  const size_t nout = 2 * (2 + 8 * (size_t)m);
  double pp[60], hout[2 * (2 + 8 * 15)];
  std::memcpy(pp, paramk, sizeof(double) * npt);
  std::memcpy(pp + npt, paraml, sizeof(double) * npt);
  HIPCHK(hipMemcpyAsync(c1.params, pp, sizeof(double) * 2 * npt, hipMemcpyHostToDevice, x->stream));
  RCCHK(ecg::prep(c1, x->stream));
  RCCHK(launch_one_pair(c1, sc, c1.bra, c1.ket, dout, x->stream));
  HIPCHK(hipMemcpyAsync(hout, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (!finite_all(hout, 4)) return ECG_ERR_NONFINITE;
  Hkl[0] = hout[0];
  Hkl[1] = hout[1];
  Skl[0] = hout[2];
  Skl[1] = hout[3];
  if (gk) std::memcpy(Dk, hout + 4, sizeof(double) * 8 * m);
  if (gl) std::memcpy(Dl, hout + 4 + 8 * m, sizeof(double) * 8 * m);
  return ECG_OK;
}

// Scratch of ecg_matrix_elements_noalloc: the staging of one pair, in a
// file-scope device array (one copy per GPU, resident until the process
// ends) so that the entry point allocates nothing.  Host code must not
// pass `scratch` itself to a kernel or a copy: its device address comes
// from hipGetSymbolAddress, resolved once per device (with that device
// current) and kept in scratch_ptr.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = 1024;   // 8 KiB
static_assert(ME_STAGE_DOUBLES <= NOALLOC_SCRATCH_DOUBLES, "scratch smaller than one pair's staging");
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];
constexpr int MAX_DEVICES = 64;
double* scratch_ptr[MAX_DEVICES] = {};
std::mutex scratch_mu;

// Device address of `scratch` on device dev (0 when it cannot be resolved).
double* scratch_on(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_ptr[dev]) return scratch_ptr[dev];
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return nullptr;
  if (cur != dev && hipSetDevice(dev) != hipSuccess) return nullptr;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != dev) (void)hipSetDevice(cur);
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return nullptr;
  }
  scratch_ptr[dev] = static_cast<double*>(p);
  return scratch_ptr[dev];
}

}  // namespace

// EnergyGB's contraction on the device (src/workproc.f90:1925-1940), for the
// stored-derivative mode: one workgroup per optimised function k,
//   W = sum_l c_l (D_H(:,k,l) - E D_S(:,k,l)),
//   grad(:,k) = 2 Re(conj(c_k) W) - Re(|c_k|^2 (D_H - E D_S)(:,k,k+nfru)).
// D(m, k, l) at m + 2 npt (k + nfo l); thread -> (m, l-group) so that a
// group of npt threads reads one contiguous column chunk.  Fixed-order
// reduction over the l-groups.
__global__ void __launch_bounds__(256) k_contract_stored(int npt, int nfo, int nfa, int nfru,
                                                         const zd* __restrict__ D, const zd* __restrict__ c,
                                                         double E, double* __restrict__ grad) {
  __shared__ zd part[256];
  const int k = blockIdx.x;
  const int groups = blockDim.x / npt;
  const int t = threadIdx.x;
  const int m = t % npt, g = t / npt;
  zd w = ecg::mkz(0.0, 0.0);
  if (g < groups) {
    for (int l = g; l < nfa; l += groups) {
      const zd* col = D + (size_t)2 * npt * ((size_t)k + (size_t)nfo * l);
      const zd d = ecg::zfms(ecg::mkz(E, 0.0), col[m + npt], col[m]);
      w = ecg::zfma(c[l], d, w);
    }
  }
  part[t] = w;
  __syncthreads();
  if (t < npt) {
    zd W = ecg::mkz(0.0, 0.0);
    for (int q = 0; q < groups; ++q) W = W + part[q * npt + t];
    const zd ck = c[k + nfru];
    const zd* col = D + (size_t)2 * npt * ((size_t)k + (size_t)nfo * (k + nfru));
    const zd d = ecg::zfms(ecg::mkz(E, 0.0), col[t + npt], col[t]);
    const double g2 = 2.0 * (ck.x * W.x + ck.y * W.y);
    const double cc = ck.x * ck.x + ck.y * ck.y;
    grad[(size_t)k * npt + t] = g2 - cc * d.x;
  }
}

extern "C" {

int ecg_ctx_create(int n, int T, const int* yhy_mats, const int* yhy_coef, const double* mass, double q0,
                   const double* q, ecg_ctx** out) {
  if (!out) return ECG_ERR_ARG;
  *out = nullptr;
  if (n < 1 || n > 5 || T < 1 || T > 16 || !yhy_mats || !yhy_coef || !mass || !q) return ECG_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return ECG_ERR_HIP;
  ecg_ctx* x = new ecg_ctx();
  ecg::SysConst& sc = x->c.sc;
  std::memset(&sc, 0, sizeof(sc));
  sc.n = n;
  sc.T = T;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) sc.M[i][j] = mass[i + n * j];
  sc.q0 = q0;
  for (int i = 0; i < n; ++i) sc.q[i] = q[i];
  sc.two3n2 = std::pow(2.0, (3 * n) / 2.0);
  for (int t = 0; t < T; ++t) {
    sc.coef[t] = (double)yhy_coef[t];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const int v = yhy_mats[(size_t)t * n * n + i + n * j];
        if (v < -127 || v > 127) {
          delete x;
          return ECG_ERR_ARG;
        }
        sc.P[t][i][j] = (signed char)v;
      }
  }
  x->c.n = n;
  x->c.T = T;
  x->npt = n * (n + 1);
  if (hipGetDevice(&x->dev) != hipSuccess) {
    delete x;
    return ECG_ERR_HIP;
  }
  if (!scratch_on(x->dev)) {   // resolve the no-allocation entry's scratch once per device
    delete x;
    return ECG_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess) {
    delete x;
    return ECG_ERR_HIP;
  }
  for (int i = 0; i < 8; ++i)
    if (hipEventCreate(&x->e[i]) != hipSuccess) {
      delete x;
      return ECG_ERR_HIP;
    }
  // coherent: the kernel's stores reach host memory without a cache in
  // between, whatever release scope the runtime ends the kernel with
  if (hipHostMalloc((void**)&x->hrep, 16 * sizeof(double), hipHostMallocCoherent) != hipSuccess) {
    x->hrep = nullptr;
    ecg_ctx_destroy(x);
    return ECG_ERR_HIP;
  }
  *out = x;
  return ECG_OK;
}

int ecg_ctx_destroy(ecg_ctx* x) {
  if (!x) return ECG_ERR_ARG;
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  ecg::Ctx& c = x->c;
  void* ptrs[] = {c.params, c.bra, c.ket, c.diagS, c.diagH, c.H, c.S, c.rowbuf, c.adjraw, x->f.F, x->f.D, x->f.W, x->f.Dinv,
                  x->f.status, x->f.vec, x->f.tvec, x->f.pflag, x->f.scal, x->part, x->grad, x->ev[0],
                  x->ev[1], x->ev[2]};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (x->pinned) (void)hipHostFree(x->pinned);
  if (x->hrep) (void)hipHostFree(x->hrep);
  free_screen(x->scr);
  if (x->dD) (void)hipFree(x->dD);
  if (x->pen) (void)hipFree(x->pen);
  if (x->f.tp) (void)hipFree(x->f.tp);
  if (x->xpart) (void)hipFree(x->xpart);
  for (hipEvent_t e : x->tr.pool) (void)hipEventDestroy(e);
  for (int i = 0; i < 8; ++i)
    if (x->e[i]) (void)hipEventDestroy(x->e[i]);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x;
  return ECG_OK;
}

int ecg_set_params(ecg_ctx* x, int K, const double* params) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || K < 1 || !params) return ECG_ERR_ARG;
  const size_t npt = (size_t)x->npt;
  const bool resident = x->params_set && K == x->c.K && x->hparams.size() == npt * K;
  // functions d..e-1: the first to the last one that differs from the
  // resident upload (all of them when nothing is resident); the others
  // equal values already checked
  const int d = first_changed(x, K, params);
  int e = K;
  if (resident)
    while (e > d && std::memcmp(x->hparams.data() + (e - 1) * npt, params + (e - 1) * npt, npt * sizeof(double)) == 0)
      --e;
  if (!finite_all(params + d * npt, npt * (e - d))) return ECG_ERR_ARG;
  x->s_ready = false;
  if (d < K || K != x->c.K || !x->params_set) params_changed(x);
  x->params_set = false;
  if (resident) {
    // the same basis size with its parameters resident: only functions
    // d..e-1 go up, and only their records are recomputed
    if (e > d) {
      if (npt * (e - d) <= (size_t)ecg::PB_DOUBLES) {
        RCCHK(ecg::prep_put(x->c, d, e - d, params + d * npt, x->stream));
      } else {
        HIPCHK(hipMemcpyAsync(x->c.params + d * npt, params + d * npt, sizeof(double) * npt * (e - d),
                              hipMemcpyHostToDevice, x->stream));
        RCCHK(ecg::prep_range(x->c, d, e, x->stream));
        HIPCHK(hipStreamSynchronize(x->stream));
      }
      std::memcpy(x->hparams.data() + d * npt, params + d * npt, sizeof(double) * npt * (e - d));
    }
  } else {
    RCCHK(upload_params(x, K, params));
    HIPCHK(hipStreamSynchronize(x->stream));
    x->hparams.assign(params, params + npt * K);
  }
  x->params_set = true;
  rows_changed(x, d);
  return ECG_OK;
}

int ecg_set_params_device(ecg_ctx* x, int K, const double* dparams) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || K < 1 || !dparams) return ECG_ERR_ARG;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, dparams) != hipSuccess || attr.type != hipMemoryTypeDevice) {
    (void)hipGetLastError();
    return ECG_ERR_ARG;
  }
  x->s_ready = false;
  params_changed(x);
  x->params_set = false;
  RCCHK(ensure_capacity(x, K));
  x->hparams.clear();
  rows_changed(x, 0);
  HIPCHK(hipMemcpyAsync(x->c.params, dparams, sizeof(double) * x->npt * K, hipMemcpyDeviceToDevice, x->stream));
  RCCHK(ecg::prep(x->c, x->stream));
  x->params_set = true;
  return ECG_OK;
}

int ecg_update_params(ecg_ctx* x, int first, int count, const double* params) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !params || first < 1 || count < 1 || first - 1 + count > x->c.K) return ECG_ERR_ARG;
  if (!finite_all(params, (size_t)x->npt * count)) return ECG_ERR_ARG;
  if (!x->params_set) return ECG_ERR_STATE;
  x->s_ready = false;
  params_changed(x);
  rows_changed(x, first - 1);
  if (x->hparams.size() == (size_t)x->npt * x->c.K)
    std::memcpy(x->hparams.data() + (size_t)x->npt * (first - 1), params, sizeof(double) * x->npt * count);
  if ((size_t)x->npt * count <= (size_t)ecg::PB_DOUBLES) {
    // a few functions: the values travel with the launch that stores them
    // and computes their records
    RCCHK(ecg::prep_put(x->c, first - 1, count, params, x->stream));
    return ECG_OK;
  }
  HIPCHK(hipMemcpyAsync(x->c.params + (size_t)x->npt * (first - 1), params, sizeof(double) * x->npt * count,
                        hipMemcpyHostToDevice, x->stream));
  RCCHK(ecg::prep_range(x->c, first - 1, first - 1 + count, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return ECG_OK;
}

int ecg_matrix_elements(ecg_ctx* x, const double* paramk, const double* paraml, const int* P, int gradflag,
                        double* Hkl, double* Skl, double* Dk, double* Dl) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !paramk || !paraml || !P || !Hkl || !Skl) return ECG_ERR_ARG;
  const bool gk = (gradflag == 1 || gradflag == 3), gl = (gradflag == 2 || gradflag == 3);
  if ((gk && !Dk) || (gl && !Dl)) return ECG_ERR_ARG;
  double* dev = nullptr;
  if (hipMalloc((void**)&dev, sizeof(double) * ME_STAGE_DOUBLES) != hipSuccess) return ECG_ERR_NOMEM;
  const int rc = matrix_elements_run(x, paramk, paraml, P, gradflag, Hkl, Skl, Dk, Dl, dev);
  (void)hipStreamSynchronize(x->stream);
  (void)hipFree(dev);
  return rc;
}

int ecg_matrix_elements_noalloc(ecg_ctx* x, const double* paramk, const double* paraml, const int* P,
                                int gradflag, double* Hkl, double* Skl, double* Dk, double* Dl) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !paramk || !paraml || !P || !Hkl || !Skl) return ECG_ERR_ARG;
  const bool gk = (gradflag == 1 || gradflag == 3), gl = (gradflag == 2 || gradflag == 3);
  if ((gk && !Dk) || (gl && !Dl)) return ECG_ERR_ARG;
  // the staging of one pair must fit the scratch (n <= 5 always does)
  const int n = x->c.n, m = n * (n + 1) / 2;
  const size_t need = (size_t)ME_OFF_OUT + 2 * (2 + 8 * (size_t)m);
  if (n < 1 || n > 5 || need > NOALLOC_SCRATCH_DOUBLES) return ECG_ERR_ARG;
  double* dev = scratch_on(x->dev);   // resolved at context creation: no HIP call here
  if (!dev) return ECG_ERR_HIP;
  return matrix_elements_run(x, paramk, paraml, P, gradflag, Hkl, Skl, Dk, Dl, dev);
}

int ecg_compute_mat_elem(ecg_ctx* x, int nmin, int nmax, char method, double approx_energy, double* H,
                         double* S, int ld, double* diagH, double* diagS) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !H || !S || !diagH || !diagS) return ECG_ERR_ARG;
  if (method != 'G' && method != 'I') return ECG_ERR_ARG;
  if (x->c.K < 1 || !x->params_set) return ECG_ERR_STATE;
  if (nmin < 1 || nmax > x->c.K || nmin > nmax || ld < nmax) return ECG_ERR_ARG;
  // rows < nmin keep their device values; their diagS comes from the caller
  if (nmin > 1)
    HIPCHK(hipMemcpyAsync(x->c.diagS, diagS, sizeof(double) * (nmin - 1), hipMemcpyHostToDevice, x->stream));
  if (nmin > 1) {   // rows normalised with the caller's diagS: nothing below is trusted afterwards
    x->scr.ready = false;
    x->hs_rows = 0;
    hs_rewritten(x, 0);
  } else {
    hs_rewritten(x, x->hs_rows);
    x->hs_rows = std::max(x->hs_rows, nmax);
  }
  RCCHK(ecg::assemble(x->c, nmin - 1, nmax, x->stream));
  x->s_ready = (nmax == x->c.K);
  return copy_out_hs(x, nmin - 1, nmax, method, approx_energy, H, S, ld, diagH, diagS);
}

int ecg_compute_mat_elem_and_deriv(ecg_ctx* x, int nfru, char method, double approx_energy, double* H,
                                   double* S, int ld, double* diagH, double* diagS, double* D, int ldD) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !H || !S || !diagH || !diagS || !D) return ECG_ERR_ARG;
  if (method != 'G' && method != 'I') return ECG_ERR_ARG;
  const int K = x->c.K;
  if (K < 1 || !x->params_set) return ECG_ERR_STATE;
  if (nfru < 0 || nfru >= K || ld < K || ldD < K - nfru) return ECG_ERR_ARG;
  const int nfo = K - nfru;
  const size_t nD = (size_t)2 * x->npt * nfo * K;   // complex entries on the device
  zd* dD = nullptr;
  if (hipMalloc((void**)&dD, sizeof(zd) * nD) != hipSuccess) return ECG_ERR_NOMEM;
  int rc = 0;
  if (nfru > 0 &&
      hipMemcpyAsync(x->c.diagS, diagS, sizeof(double) * nfru, hipMemcpyHostToDevice, x->stream) != hipSuccess)
    rc = ECG_ERR_HIP;
  if (nfru > 0) x->scr.ready = false;   // rows normalised with the caller's diagS
  x->hs_rows = 0;
  hs_rewritten(x, 0);
  if (!rc) rc = ecg::assemble_deriv(x->c, nfru, dD, x->stream);
  if (!rc) x->s_ready = true;
  // D(:, k, l) of the caller's Glob_D(2 npt, ldD, *): K column blocks of
  // 2 npt x nfo complex, placed with the pitch of its allocated extent
  const size_t wbytes = sizeof(zd) * 2 * x->npt * nfo, pitch = sizeof(zd) * 2 * x->npt * (size_t)ldD;
  if (!rc && hipMemcpy2DAsync(D, pitch, dD, wbytes, wbytes, K, hipMemcpyDeviceToHost, x->stream) != hipSuccess)
    rc = ECG_ERR_HIP;
  if (!rc && hipStreamSynchronize(x->stream) != hipSuccess) rc = ECG_ERR_HIP;
  (void)hipFree(dD);
  if (rc) return rc < 0 ? rc : ECG_ERR_HIP;
  for (int l = 0; l < K; ++l)
    if (!finite_all(D + (size_t)l * pitch / sizeof(double), wbytes / sizeof(double))) return ECG_ERR_NONFINITE;
  return copy_out_hs(x, nfru, K, method, approx_energy, H, S, ld, diagH, diagS);
}

int ecg_ghepiis(ecg_ctx* x, int k, int n, double* M, int nM, double* invD, const double* B, int nB,
                double apprlambda, const double* v, double tol, int maxiter, int specif_norm, double* lambda,
                double* xout, double* relacc, int* numiter) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !M || !invD || !B || !v || !lambda || !xout || n < 1 || k < 1 || k > n || nM < n || nB < n ||
      maxiter < 1)
    return ECG_ERR_ARG;
  // The device H, S and the eigenvector slot take the caller's problem: the
  // session state of the context's own basis is closed, and when the size
  // differs from the context's basis its parameters are no longer usable
  // (an energy call then returns ECG_ERR_STATE until ecg_set_params).
  params_changed(x);
  if (n != x->c.K) {
    x->params_set = false;
    x->hparams.clear();
  }
  RCCHK(ensure_capacity(x, n));
  ecg::Ctx& c = x->c;
  // Device H := M (Hermitian from its lower triangle), S := B; the factor
  // and the iteration run at shift 0, so that the matrix factored is M
  // itself (no rounding from rebuilding it as (M + lambda B) - lambda B),
  // and apprlambda is added to the Rayleigh quotient at the end.
  const size_t nd = (size_t)2 * c.ld * n;
  RCCHK(ensure_pinned(x, 2 * nd));
  double* hH = x->pinned;
  double* hS = x->pinned + nd;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const size_t ob = 2 * ((size_t)j * nB + i), od = 2 * ((size_t)j * c.ld + i);
      hS[od] = B[ob];
      hS[od + 1] = B[ob + 1];
      const int a = std::max(i, j), b = std::min(i, j);
      const size_t om = 2 * ((size_t)b * nM + a);   // lower element (a, b)
      double mr = M[om], mi = M[om + 1];
      if (i < j) mi = -mi;                          // upper = conj(lower)
      if (i == j) mi = 0.0;
      hH[od] = mr;
      hH[od + 1] = mi;
    }
  x->hs_rows = 0;   // the device H, S now hold the caller's matrices
  hs_rewritten(x, 0);
  HIPCHK(hipMemcpyAsync(c.H, hH, sizeof(double) * nd, hipMemcpyHostToDevice, x->stream));
  x->s_ready = false;
  HIPCHK(hipMemcpyAsync(c.S, hS, sizeof(double) * nd, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(x->ev[2], v, sizeof(zd) * n, hipMemcpyHostToDevice, x->stream));
  x->have_last = true;
  x->last_K = n;
  int nneg = 0, zero = 0;
  RCCHK(factor_at(x, 0.0, &nneg, &zero));
  if (zero) return ECG_SINGULAR;
  InvitResult r;
  RCCHK(inverse_iteration(x, 0.0, tol, maxiter, true, r, false));
  r.lambda += apprlambda;
  // x->ev[2] holds x normalised to x^H B x = 1
  std::vector<double> hx(2 * (size_t)n), hD(n);
  HIPCHK(hipMemcpyAsync(hx.data(), x->ev[2], sizeof(zd) * n, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(hD.data(), x->f.D, sizeof(double) * n, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(hH, x->f.F, sizeof(double) * nd, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (specif_norm == 1 || (specif_norm != 0)) {
    double nrm = 0.0, mx = 0.0;
    for (int i = 0; i < n; ++i) {
      nrm += hx[2 * i] * hx[2 * i] + hx[2 * i + 1] * hx[2 * i + 1];
      mx = std::max(mx, std::max(std::fabs(hx[2 * i]), std::fabs(hx[2 * i + 1])));
    }
    const double s = (specif_norm == 1) ? 1.0 / std::sqrt(nrm) : 1.0 / mx;
    for (int i = 0; i < 2 * n; ++i) hx[i] *= s;
  }
  std::memcpy(xout, hx.data(), sizeof(double) * 2 * n);
  for (int i = 0; i < n; ++i) {
    invD[2 * i] = 1.0 / hD[i];
    invD[2 * i + 1] = 0.0;
  }
  // L^H into the strict upper triangle of M: M(i, j) = conj(L(j, i)), i < j
  for (int j = 1; j < n; ++j)
    for (int i = 0; i < j; ++i) {
      const size_t od = 2 * ((size_t)i * c.ld + j), om = 2 * ((size_t)j * nM + i);
      M[om] = hH[od];
      M[om + 1] = -hH[od + 1];
    }
  *lambda = r.lambda;
  if (relacc) *relacc = r.nd;
  if (numiter) *numiter = r.iters;
  if (!r.converged && r.nd > std::fabs(tol)) return ECG_NOT_CONVERGED;
  return ECG_OK;
}

int ecg_energy(ecg_ctx* x, int which, double sigma, double tol, int maxiter, double* E, double* c, double* info) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !E || which < 1 || maxiter < 1) return ECG_ERR_ARG;
  const int K = x->c.K;
  if (K < 1 || !x->params_set) return ECG_ERR_STATE;
  if (which > K) return ECG_ERR_ARG;
  HIPCHK(hipEventRecord(x->e[0], x->stream));
  hs_rewritten(x, x->hs_rows);   // rows below hs_rows are recomputed nowhere
  RCCHK(ecg::assemble(x->c, x->hs_rows, K, x->stream));
  x->hs_rows = K;
  HIPCHK(hipEventRecord(x->e[1], x->stream));
  int status = 0;
  x->s_ready = true;
  x->vec_current = false;
  RCCHK(solve_root(x, which, sigma, tol, maxiter, E, info, &status));
  x->vec_current = (status == 0 || status == ECG_NOT_CONVERGED);
  x->pen_total = 0.0;
  if (x->pen_on) RCCHK(penalty_enqueue(x, x->stream, x->pen_nfru, x->pen_b, x->pen_t2, nullptr));
  HIPCHK(hipEventRecord(x->e[2], x->stream));
  if (c) HIPCHK(hipMemcpyAsync(c, x->ev[2], sizeof(zd) * K, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  *E += x->pen_total;   // EnergyGA/EnergyIA return Evalue + Glob_TotalOverlapPenalty
  if (info) info[6] = x->pen_total;
  x->tms[0] = ev_ms(x->e[0], x->e[1]);
  x->tms[1] = ev_ms(x->e[1], x->e[2]);
  x->tms[2] = 0.0;
  x->tms[3] = ev_ms(x->e[0], x->e[2]);
  trace_collect(x);
  if (!std::isfinite(*E)) return ECG_ERR_NONFINITE;
  return status;
}

int ecg_energy_grad(ecg_ctx* x, int nfru, int which, double sigma, double tol, int maxiter, double* E,
                    double* grad, double* info) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !E || !grad || which < 1 || maxiter < 1) return ECG_ERR_ARG;
  ecg::Ctx& c = x->c;
  const int K = c.K;
  if (K < 1 || !x->params_set) return ECG_ERR_STATE;
  if (nfru < 0 || nfru >= K || which > K) return ECG_ERR_ARG;
  const size_t np = ecg::adjoint_partial_doubles(c.n, K);
  if (np > x->part_cap) {
    RCCHK(dev_alloc(&x->part, np));
    x->part_cap = np;
  }
  const size_t ng = (size_t)x->npt * (K - nfru);
  if (ng > x->grad_cap) {
    RCCHK(dev_alloc(&x->grad, ng));
    x->grad_cap = ng;
  }
  c.ev_off0 = x->e[5];
  c.ev_off1 = x->e[6];
  HIPCHK(hipEventRecord(x->e[0], x->stream));
  hs_rewritten(x, x->hs_rows);
  RCCHK(ecg::assemble(c, x->hs_rows, K, x->stream));
  x->hs_rows = K;
  HIPCHK(hipEventRecord(x->e[1], x->stream));
  int status = 0;
  x->vec_current = false;
  // the gradient's contraction is queued before the solve's host
  // synchronisation, with E read on the device (k_report); it is redone
  // with the host's E only if the iteration went on past that point
  c.ev_adj0 = x->e[3];
  c.ev_adj1 = x->e[4];
  const std::function<int()> queue_adjoint = [&]() -> int {
    HIPCHK(hipEventRecord(x->e[2], x->stream));
    return ecg::adjoint(c, nfru, 0.0, x->f.scal + 8, x->ev[2], x->part, x->grad, x->stream);
  };
  bool queued = false;
  RCCHK(solve_root(x, which, sigma, tol, maxiter, E, info, &status, &queue_adjoint, &queued));
  x->vec_current = (status == 0 || status == ECG_NOT_CONVERGED);
  if (!queued) {
    HIPCHK(hipEventRecord(x->e[2], x->stream));
    RCCHK(ecg::adjoint(c, nfru, *E, nullptr, x->ev[2], x->part, x->grad, x->stream));
  }
  x->s_ready = true;
  x->pen_total = 0.0;
  if (x->pen_on && K != 1) RCCHK(penalty_enqueue(x, x->stream, nfru, x->pen_b, x->pen_t2, x->grad));
  HIPCHK(hipEventRecord(x->e[7], x->stream));
  HIPCHK(hipMemcpyAsync(grad, x->grad, sizeof(double) * ng, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  *E += x->pen_total;   // EnergyGB/EnergyIB: Evalue + Glob_TotalOverlapPenalty
  if (info) info[6] = x->pen_total;
  c.ev_off0 = c.ev_off1 = c.ev_adj0 = c.ev_adj1 = nullptr;
  x->tms[0] = ev_ms(x->e[0], x->e[1]);
  x->tms[1] = ev_ms(x->e[1], x->e[2]);
  x->tms[2] = ev_ms(x->e[2], x->e[7]);
  x->tms[3] = ev_ms(x->e[0], x->e[7]);
  x->tms[4] = ev_ms(x->e[3], x->e[4]);
  x->tms[5] = ev_ms(x->e[5], x->e[6]);
  trace_collect(x);
  if (!std::isfinite(*E) || !finite_all(grad, ng)) return ECG_ERR_NONFINITE;
  return status;
}

int ecg_energy_grad_stored(ecg_ctx* x, int nfru, int which, double sigma, double tol, int maxiter, double* E,
                           double* grad, double* info) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !E || !grad || which < 1 || maxiter < 1) return ECG_ERR_ARG;
  ecg::Ctx& c = x->c;
  const int K = c.K;
  if (K < 1 || !x->params_set) return ECG_ERR_STATE;
  if (nfru < 0 || nfru >= K || which > K) return ECG_ERR_ARG;
  const int nfo = K - nfru;
  const size_t nD = (size_t)2 * x->npt * nfo * K;
  if (nD > x->dD_cap) {
    RCCHK(dev_alloc(&x->dD, nD));
    x->dD_cap = nD;
  }
  const size_t ng = (size_t)x->npt * nfo;
  if (ng > x->grad_cap) {
    RCCHK(dev_alloc(&x->grad, ng));
    x->grad_cap = ng;
  }
  HIPCHK(hipEventRecord(x->e[0], x->stream));
  // rows >= nfru come from the derivative kernel (its H, S may differ from
  // the pair kernel's in the last bit)
  const int r0 = std::min(x->hs_rows, nfru);
  hs_rewritten(x, r0);
  if (r0 < nfru) RCCHK(ecg::assemble(c, r0, nfru, x->stream));
  RCCHK(ecg::assemble_deriv(c, nfru, x->dD, x->stream));
  x->hs_rows = K;
  HIPCHK(hipEventRecord(x->e[1], x->stream));
  int status = 0;
  x->vec_current = false;
  RCCHK(solve_root(x, which, sigma, tol, maxiter, E, info, &status));
  x->vec_current = (status == 0 || status == ECG_NOT_CONVERGED);
  HIPCHK(hipEventRecord(x->e[2], x->stream));
  hipLaunchKernelGGL(k_contract_stored, dim3(nfo), dim3(256), 0, x->stream, x->npt, nfo, K, nfru, x->dD, x->ev[2],
                     *E, x->grad);
  HIPCHK(hipGetLastError());
  x->s_ready = true;
  x->pen_total = 0.0;
  if (x->pen_on && K != 1) RCCHK(penalty_enqueue(x, x->stream, nfru, x->pen_b, x->pen_t2, x->grad));
  HIPCHK(hipEventRecord(x->e[7], x->stream));
  HIPCHK(hipMemcpyAsync(grad, x->grad, sizeof(double) * ng, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  *E += x->pen_total;
  if (info) info[6] = x->pen_total;
  x->tms[0] = ev_ms(x->e[0], x->e[1]);
  x->tms[1] = ev_ms(x->e[1], x->e[2]);
  x->tms[2] = ev_ms(x->e[2], x->e[7]);
  x->tms[3] = ev_ms(x->e[0], x->e[7]);
  x->tms[4] = x->tms[2];
  x->tms[5] = 0.0;
  trace_collect(x);
  if (!std::isfinite(*E) || !finite_all(grad, ng)) return ECG_ERR_NONFINITE;
  return status;
}

namespace {
// After a batched inverse iteration (ecg::invit_batch, nq iterations and
// the tail, reported in x->hrep): the decisions of inverse_iteration's
// energy path (stop_at_floor) and solve_root's epilogue for one problem.
// A problem whose measure asks for more iterations continues on its own
// stream from the normalised iterate.
int invit_batch_finish(ecg_ctx* x, double sigma, double tol, int maxiter, int nq, double* E, int* status) {
  const int K = x->c.K;
  double q[5];
  int st[3];
  for (int i = 0; i < 5; ++i) q[i] = x->hrep[i];
  for (int i = 0; i < 3; ++i) st[i] = (int)x->hrep[5 + i];
  *status = 0;
  if (st[2]) return ECG_ERR_SYNC;   // a device-side wait of the factorisation hit its limit
  if (st[1]) {
    *status = ECG_SINGULAR;
    return 0;
  }
  InvitResult r;
  r.iters = nq;
  r.nd = q[0];
  if (!std::isfinite(r.nd)) return ECG_ERR_NONFINITE;
  x->have_last = true;
  x->last_K = K;
  const double p1 = (nq >= 2) ? q[4] : std::numeric_limits<double>::infinity();
  bool done = false;
  if (tol > 0.0 && r.nd <= tol) {
    r.converged = true;
    done = true;
  } else if (nq >= 2 && r.nd < 1e-9 && r.nd > 0.5 * p1) {
    done = true;
  } else if (nq >= maxiter) {
    done = true;
  }
  if (done) {
    r.t1 = q[2];
    r.lambda = q[3] / q[2] + sigma;
  } else {
    int fst[3] = {0, 0, 0};
    RCCHK(inverse_iteration(x, sigma, tol, maxiter - nq, true, r, true, fst, 1));
    if (fst[1]) {
      *status = ECG_SINGULAR;
      return 0;
    }
    r.iters += nq;
  }
  x->fac_rows = K;   // the factor is complete and without zero pivots
  *E = r.lambda;
  if (!r.converged && r.nd > std::fabs(tol)) *status = ECG_NOT_CONVERGED;
  return 0;
}
}  // namespace

int ecg_batch_create(int nprob, ecg_ctx** ctxs, ecg_batch** out) {
  if (!out || nprob < 1 || !ctxs) return ECG_ERR_ARG;
  *out = nullptr;
  for (int i = 0; i < nprob; ++i)
    if (!ctxs[i] || ctxs[i]->c.K < 1 || !ctxs[i]->params_set) return ECG_ERR_STATE;
  ecg_batch* b = new ecg_batch();
  b->ctx.assign(ctxs, ctxs + nprob);
  if (hipStreamCreateWithFlags(&b->lead, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_asm, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_fac, hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void**)&b->fd, sizeof(ecg::FacDesc) * nprob) != hipSuccess ||
      hipMalloc((void**)&b->id, sizeof(ecg::InvDesc) * nprob) != hipSuccess) {
    ecg_batch_destroy(b);
    return ECG_ERR_HIP;
  }
  b->ev_eig.resize(nprob, nullptr);
  for (int i = 0; i < nprob; ++i)
    if (hipEventCreateWithFlags(&b->ev_eig[i], hipEventDisableTiming) != hipSuccess) {
      ecg_batch_destroy(b);
      return ECG_ERR_HIP;
    }
  for (int i = 0; i < nprob; ++i) {
    const int n = ctxs[i]->c.n;
    auto it = std::find_if(b->classes.begin(), b->classes.end(), [&](const ecg_batch::Class& c) { return c.n == n; });
    if (it == b->classes.end()) {
      b->classes.emplace_back();
      it = b->classes.end() - 1;
      it->n = n;
    }
    it->probs.push_back(i);
  }
  for (auto& cl : b->classes) {
    cl.h.resize(cl.probs.size());
    if (hipMalloc((void**)&cl.d, sizeof(ecg::ProbDesc) * cl.probs.size()) != hipSuccess) {
      ecg_batch_destroy(b);
      return ECG_ERR_NOMEM;
    }
  }
  *out = b;
  return ECG_OK;
}

int ecg_batch_destroy(ecg_batch* b) {
  if (!b) return ECG_ERR_ARG;
  if (b->lead) (void)hipStreamSynchronize(b->lead);
  for (auto& cl : b->classes)
    if (cl.d) (void)hipFree(cl.d);
  for (hipEvent_t e : b->ev_eig)
    if (e) (void)hipEventDestroy(e);
  if (b->ev_asm) (void)hipEventDestroy(b->ev_asm);
  if (b->ev_fac) (void)hipEventDestroy(b->ev_fac);
  if (b->fd) (void)hipFree(b->fd);
  if (b->id) (void)hipFree(b->id);
  if (b->lead) (void)hipStreamDestroy(b->lead);
  delete b;
  return ECG_OK;
}

int ecg_batch_set_config(ecg_batch* b, int packed, int nstreams) {
  if (!b || packed < 0 || packed > 2 || nstreams < 1 || nstreams > 16) return ECG_ERR_ARG;
  b->packed = packed;
  b->nstreams = nstreams;
  return ECG_OK;
}

namespace {
// descriptors of every class from the contexts' current state
int batch_descriptors(ecg_batch* b, const double* E) {
  HIPCHK(hipStreamSynchronize(b->lead));   // the previous upload has consumed the host tables
  for (auto& cl : b->classes) {
    long long t0 = 0;
    int f0 = 0;
    cl.maxT = 0;
    for (size_t q = 0; q < cl.probs.size(); ++q) {
      const int i = cl.probs[q];
      ecg_ctx* x = b->ctx[i];
      const ecg::Ctx& c = x->c;
      ecg::ProbDesc& d = cl.h[q];
      d.sc = c.sc;
      d.K = c.K;
      d.ld = c.ld;
      d.nt = (c.K + ecg::TILE - 1) / ecg::TILE;
      d.f0 = f0;
      d.tile0_off = t0;
      d.tile0_adj = t0;
      t0 += (long long)d.nt * (d.nt + 1) / 2;
      f0 += c.K;
      d.bra = c.bra;
      d.ket = c.ket;
      d.diagS = c.diagS;
      d.H = c.H;
      d.S = c.S;
      d.cvec = x->ev[2];
      d.part = x->part;
      d.grad = x->grad;
      d.E = E ? E[i] : 0.0;
      cl.maxT = std::max(cl.maxT, c.T);
    }
    cl.ntiles = t0;
    cl.nf = f0;
    HIPCHK(hipMemcpyAsync(cl.d, cl.h.data(), sizeof(ecg::ProbDesc) * cl.h.size(), hipMemcpyHostToDevice, b->lead));
  }
  return 0;
}

int ensure_grad_buffers(ecg_ctx* x) {
  const size_t np = ecg::adjoint_partial_doubles(x->c.n, x->c.K);
  if (np > x->part_cap) {
    RCCHK(dev_alloc(&x->part, np));
    x->part_cap = np;
  }
  const size_t ng = (size_t)x->npt * x->c.K;
  if (ng > x->grad_cap) {
    RCCHK(dev_alloc(&x->grad, ng));
    x->grad_cap = ng;
  }
  return 0;
}
}  // namespace

int ecg_batch_energy_grad(ecg_batch* b, const double* sigma, double tol, int maxiter, double* E, double* grad,
                          double* lat_ms) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!b || !sigma || !E || !grad || maxiter < 1) return ECG_ERR_ARG;
  const int np = (int)b->ctx.size();
  std::vector<size_t> goff(np + 1, 0);
  for (int i = 0; i < np; ++i) {
    RCCHK(ensure_grad_buffers(b->ctx[i]));
    goff[i + 1] = goff[i] + (size_t)b->ctx[i]->npt * b->ctx[i]->c.K;
  }
  const auto t0 = std::chrono::steady_clock::now();
  auto since = [&]() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  };
  // contexts' own streams must have finished earlier work before the lead stream touches their buffers
  for (ecg_ctx* x : b->ctx) HIPCHK(hipStreamSynchronize(x->stream));
  for (ecg_ctx* x : b->ctx) {   // every matrix and factor is rebuilt in full
    x->hs_rows = x->c.K;
    hs_rewritten(x, 0);
  }
  if (b->packed) {
    RCCHK(batch_descriptors(b, nullptr));
    for (ecg_ctx* x : b->ctx) RCCHK(ecg::assemble_diag(x->c, b->lead));
    for (auto& cl : b->classes)
      RCCHK(ecg::packed_pairs(cl.n, false, cl.d, (int)cl.h.size(), cl.ntiles, cl.nf, cl.maxT, b->lead));
    HIPCHK(hipEventRecord(b->ev_asm, b->lead));
  }
  // packed == 2: the factorisations of H - sigma S of every problem with a
  // finite shift in shared launches on the lead stream (the others locate
  // their root by inertia on their own streams)
  std::vector<char> batched(np, 0);
  if (b->packed == 2) {
    b->fh.clear();
    int epoch = 0;
    for (int i = 0; i < np; ++i) {
      if (!std::isfinite(sigma[i])) continue;
      ecg_ctx* x = b->ctx[i];
      ecg::FacDesc d;
      d.K = x->c.K;
      d.ld = x->f.ld;
      d.F = x->f.F;
      d.D = x->f.D;
      d.Dinv = x->f.Dinv;
      d.W = x->f.W;
      d.status = x->f.status;
      d.pflag = x->f.pflag;
      d.src.H = x->c.H;
      d.src.S = x->c.S;
      d.src.sigma = sigma[i];
      b->fh.push_back(d);
      epoch = std::max(epoch, x->f.epoch);
      batched[i] = 1;
    }
    std::stable_sort(b->fh.begin(), b->fh.end(), [](const ecg::FacDesc& a, const ecg::FacDesc& c) {
      return a.K > c.K;
    });
    if (!b->fh.empty()) {
      HIPCHK(hipMemcpyAsync(b->fd, b->fh.data(), sizeof(ecg::FacDesc) * b->fh.size(), hipMemcpyHostToDevice,
                            b->lead));
      RCCHK(ecg::ldlh_factor_batch(b->fh, b->fd, epoch, b->lead));
      for (int i = 0; i < np; ++i)
        if (batched[i]) b->ctx[i]->f.epoch = epoch;
    }
    HIPCHK(hipEventRecord(b->ev_fac, b->lead));
  }
  // packed == 2: the inverse iterations of the batched factorisations in
  // shared launches too (K <= 8192), three iterations and the tail with one
  // synchronisation for all; the rest go to the workers below
  std::vector<char> done(np, 0);
  std::vector<int> st(np, 0);
  if (b->packed == 2) {
    b->ih.clear();
    std::vector<int> which;
    for (int i = 0; i < np; ++i) {
      ecg_ctx* x = b->ctx[i];
      if (!batched[i] || x->c.K > 8192) continue;
      x->fac_rows = 0;
      x->fac_sigma = sigma[i];
      x->last_tail = -1;
      x->vec_current = false;
      ecg::InvDesc q;
      q.K = x->c.K;
      q.ld = x->c.ld;
      q.warm = (x->have_last && x->last_K == x->c.K) ? 1 : 0;
      q.H = x->c.H;
      q.S = x->c.S;
      q.sigma = sigma[i];
      q.F = x->f.F;
      q.D = x->f.D;
      for (int k = 0; k < 3; ++k) q.ev[k] = x->ev[k];
      q.w = x->f.vec;
      q.z = x->f.tvec;
      q.part = x->f.W;
      q.scal = x->f.scal;
      q.status = x->f.status;
      q.rep = x->hrep;
      b->ih.push_back(q);
      which.push_back(i);
    }
    if (!b->ih.empty()) {
      const int nq = std::min(maxiter, 3);
      HIPCHK(hipMemcpyAsync(b->id, b->ih.data(), sizeof(ecg::InvDesc) * b->ih.size(), hipMemcpyHostToDevice,
                            b->lead));
      RCCHK(ecg::invit_batch(b->ih, b->id, nq, b->lead));
      HIPCHK(hipStreamSynchronize(b->lead));
      for (int i : which) {
        ecg_ctx* x = b->ctx[i];
        int status = 0;
        const int rc = invit_batch_finish(x, sigma[i], tol, maxiter, nq, &E[i], &status);
        if (!rc) x->vec_current = (status == 0 || status == ECG_NOT_CONVERGED);
        if (!rc) HIPCHK(hipEventRecord(b->ev_eig[i], x->stream));
        st[i] = rc ? rc : status;
        done[i] = 1;
      }
    }
  }
  // eigen steps (and, unpacked, the whole problem) on nstreams worker
  // threads, problems taken largest first
  std::vector<int> order(np);
  for (int i = 0; i < np; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->ctx[a]->c.K > b->ctx[c]->c.K; });
  order.erase(std::remove_if(order.begin(), order.end(), [&](int i) { return done[i] != 0; }), order.end());
  const int nwork = (int)order.size();
  std::atomic<int> next(0);
  auto worker = [&]() {
    for (;;) {
      const int q = next.fetch_add(1);
      if (q >= nwork) return;
      const int i = order[q];
      ecg_ctx* x = b->ctx[i];
      double info[8];
      int status = 0, rc = 0;
      if (b->packed) {
        if (hipStreamWaitEvent(x->stream, b->packed == 2 ? b->ev_fac : b->ev_asm, 0) != hipSuccess)
          rc = ECG_ERR_HIP;
      } else {
        rc = ecg::assemble(x->c, 0, x->c.K, x->stream);
      }
      x->vec_current = false;
      if (!rc)
        rc = solve_root(x, 1, sigma[i], tol, maxiter, &E[i], info, &status, nullptr, nullptr, batched[i] != 0);
      if (!rc) x->vec_current = (status == 0 || status == ECG_NOT_CONVERGED);
      if (!rc && !b->packed) {
        rc = ecg::adjoint(x->c, 0, E[i], nullptr, x->ev[2], x->part, x->grad, x->stream);
        x->pen_total = 0.0;
        if (!rc && x->pen_on && x->c.K != 1) rc = penalty_enqueue(x, x->stream, 0, x->pen_b, x->pen_t2, x->grad);
        if (!rc && hipMemcpyAsync(grad + goff[i], x->grad, sizeof(double) * (goff[i + 1] - goff[i]),
                                  hipMemcpyDeviceToHost, x->stream) != hipSuccess)
          rc = ECG_ERR_HIP;
        if (!rc && hipStreamSynchronize(x->stream) != hipSuccess) rc = ECG_ERR_HIP;
        if (lat_ms) lat_ms[i] = since();
      }
      if (!rc && b->packed && hipEventRecord(b->ev_eig[i], x->stream) != hipSuccess) rc = ECG_ERR_HIP;
      st[i] = rc ? rc : status;
    }
  };
  {
    std::vector<std::thread> pool;
    const int nt = std::min(b->nstreams, nwork);
    for (int t = 0; t < nt; ++t) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
  }
  for (int i = 0; i < np; ++i)
    if (st[i] < 0) return st[i];
  if (b->packed) {
    RCCHK(batch_descriptors(b, E));
    for (int i = 0; i < np; ++i) HIPCHK(hipStreamWaitEvent(b->lead, b->ev_eig[i], 0));
    for (auto& cl : b->classes)
      RCCHK(ecg::packed_pairs(cl.n, true, cl.d, (int)cl.h.size(), cl.ntiles, cl.nf, cl.maxT, b->lead));
    for (ecg_ctx* x : b->ctx) {
      x->pen_total = 0.0;
      if (x->pen_on && x->c.K != 1) RCCHK(penalty_enqueue(x, b->lead, 0, x->pen_b, x->pen_t2, x->grad));
    }
    for (int i = 0; i < np; ++i)
      HIPCHK(hipMemcpyAsync(grad + goff[i], b->ctx[i]->grad, sizeof(double) * (goff[i + 1] - goff[i]),
                            hipMemcpyDeviceToHost, b->lead));
    HIPCHK(hipStreamSynchronize(b->lead));
    const double t = since();
    if (lat_ms)
      for (int i = 0; i < np; ++i) lat_ms[i] = t;
  }
  int worst = 0;
  for (int i = 0; i < np; ++i) {
    b->ctx[i]->s_ready = true;
    E[i] += b->ctx[i]->pen_total;   // zero unless the context's penalty is on
    if (!std::isfinite(E[i]) || !finite_all(grad + goff[i], goff[i + 1] - goff[i])) return ECG_ERR_NONFINITE;
    worst = std::max(worst, st[i]);
  }
  return worst;
}

int ecg_trace(ecg_ctx* x, int enable) {
  if (!x) return ECG_ERR_ARG;
  HIPCHK(hipStreamSynchronize(x->stream));
  ecg::Tracer& t = x->tr;
  if (enable && t.pool.empty()) {
    t.pool.resize(4096, nullptr);
    for (auto& e : t.pool) HIPCHK(hipEventCreate(&e));
  }
  t.collect();
  for (int i = 0; i < ecg::TR_COUNT; ++i) {
    t.sum_ms[i] = 0.0;
    t.count[i] = 0;
  }
  t.on = enable != 0;
  x->c.tr = t.on ? &t : nullptr;
  x->f.tr = t.on ? &t : nullptr;
  return ECG_OK;
}

int ecg_trace_read(ecg_ctx* x, double* count, double* total_ms, int nmax) {
  if (!x || !count || !total_ms || nmax < 1) return ECG_ERR_ARG;
  HIPCHK(hipStreamSynchronize(x->stream));
  x->tr.collect();
  const int n = std::min(nmax, (int)ecg::TR_COUNT);
  for (int i = 0; i < n; ++i) {
    count[i] = (double)x->tr.count[i];
    total_ms[i] = x->tr.sum_ms[i];
  }
  return n;
}

int ecg_get_params(ecg_ctx* x, double* params) {
  if (!x || !params) return ECG_ERR_ARG;
  if (x->c.K < 1 || !x->params_set) return ECG_ERR_STATE;
  HIPCHK(hipMemcpyAsync(params, x->c.params, sizeof(double) * x->npt * x->c.K, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return ECG_OK;
}

namespace {
// The screening step's shapes for ncand candidates: the caller's fixed
// ones, the others timed.
int screen_shapes(ecg_ctx* x, int ncand) {
  ecg::Screen& sc = x->scr;
  if (!(x->scr_tm && x->scr_cpb)) RCCHK(ecg::screen_tune(x->c, sc, ncand, x->stream));
  if (x->scr_tm) sc.tm = x->scr_tm;
  if (x->scr_cpb) sc.cpb = x->scr_cpb;
  return 0;
}
}  // namespace

int ecg_screen_set_shapes(ecg_ctx* x, int row_tile, int cand_per_block) {
  if (!x || (row_tile != 0 && row_tile != 32 && row_tile != 64)) return ECG_ERR_ARG;
  if (cand_per_block > 1 &&
      ((cand_per_block != 2 && cand_per_block != 4 && cand_per_block != 8) ||
       ecg::cand_rows_lds(x->c.n, x->c.sc.T, cand_per_block) == 0))
    return ECG_ERR_ARG;
  if (cand_per_block < 0) return ECG_ERR_ARG;
  x->scr_tm = row_tile;
  x->scr_cpb = cand_per_block;
  return ECG_OK;
}

int ecg_screen_prepare(ecg_ctx* x, double sigma, double tol, int maxiter, double* info) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !std::isfinite(sigma) || maxiter < 1) return ECG_ERR_ARG;
  ecg::Ctx& c = x->c;
  const int K = c.K;
  if (K < 2 || !x->params_set) return ECG_ERR_STATE;
  RCCHK(ensure_screen(x, 1, 0));
  ecg::Screen& sc = x->scr;
  sc.ready = false;
  x->vec_current = false;
  RCCHK(ecg::prep(c, x->stream));
  hs_rewritten(x, 0);   // the screening factor replaces the energy paths' one
  RCCHK(ecg::assemble(c, 0, K, x->stream));
  x->hs_rows = K;
  RCCHK(ecg::screen_prepare(c, x->f, sc, sigma, x->stream));
  std::vector<double> D(K);
  int st[4];
  HIPCHK(hipMemcpyAsync(D.data(), x->f.D, sizeof(double) * K, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(st, x->f.status, sizeof(st), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  int nneg = 0, zero = 0;
  for (int i = 0; i < K - 1; ++i) {
    if (D[i] < 0.0) ++nneg;
    if (D[i] == 0.0 && !zero) zero = i + 1;
  }
  sc.nneg = nneg;
  if (st[2]) return ECG_ERR_SYNC;
  if (zero || st[1]) return ECG_SINGULAR;
  // energy and eigenvector of the current basis at this shift: start
  // vector of every candidate and the energy a candidate must beat
  InvitResult r;
  RCCHK(inverse_iteration(x, sigma, tol, maxiter, true, r, true));
  HIPCHK(hipMemcpyAsync(sc.v0, x->ev[2], sizeof(zd) * K, hipMemcpyDeviceToDevice, x->stream));
  const double state[2] = {r.lambda, 0.0};
  HIPCHK(hipMemcpyAsync(sc.state, state, sizeof(state), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (info) {
    info[0] = r.lambda;
    info[1] = nneg;
    info[2] = r.iters;
    info[3] = r.nd;
  }
  sc.ready = true;
  x->vec_current = true;
  return (r.converged || r.nd < 1e-9) ? ECG_OK : ECG_NOT_CONVERGED;
}

int ecg_screen(ecg_ctx* x, int ncand, const double* cand, int iters, double* E, double* info) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || ncand < 1 || !cand || iters < 1 || !E) return ECG_ERR_ARG;
  if (!x->scr.ready || x->scr.ld != x->c.ld || x->scr.Kp != x->c.K - 1) return ECG_ERR_STATE;
  if (!finite_all(cand, (size_t)x->npt * ncand)) return ECG_ERR_ARG;
  RCCHK(ensure_screen(x, ncand, 0));
  ecg::Screen& sc = x->scr;
  HIPCHK(hipMemcpyAsync(sc.cparams, cand, sizeof(double) * x->npt * ncand, hipMemcpyHostToDevice, x->stream));
  RCCHK(screen_shapes(x, ncand));
  HIPCHK(hipEventRecord(x->e[0], x->stream));
  RCCHK(ecg::screen_batch(x->c, x->f, sc, ncand, iters, false, x->npt, x->stream));
  HIPCHK(hipEventRecord(x->e[1], x->stream));
  std::vector<double> cs((size_t)8 * ncand);
  HIPCHK(hipMemcpyAsync(cs.data(), sc.cs, sizeof(double) * 8 * ncand, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  x->tms[0] = ev_ms(x->e[0], x->e[1]);
  x->tms[1] = sc.tm;
  x->tms[2] = sc.cpb;
  bool fin = true;
  for (int j = 0; j < ncand; ++j) {
    E[j] = cs[8 * j + 0];
    fin = fin && std::isfinite(E[j]);
    if (info) {
      info[4 * j + 0] = cs[8 * j + 1];
      info[4 * j + 1] = sc.nneg + (cs[8 * j + 2] < 0.0 ? 1 : 0);
      info[4 * j + 2] = cs[8 * j + 2];
      info[4 * j + 3] = cs[8 * j + 3];
    }
  }
  return fin ? ECG_OK : ECG_ERR_NONFINITE;
}

int ecg_screen_loop(ecg_ctx* x, int nsteps, int ncand, const double* cand, int iters, double* Ebest, int* ibest) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || nsteps < 1 || ncand < 1 || !cand || iters < 1 || !Ebest || !ibest) return ECG_ERR_ARG;
  if (!x->scr.ready || x->scr.ld != x->c.ld || x->scr.Kp != x->c.K - 1) return ECG_ERR_STATE;
  if (!finite_all(cand, (size_t)x->npt * ncand * nsteps)) return ECG_ERR_ARG;
  RCCHK(ensure_screen(x, ncand, nsteps));
  x->s_ready = false;
  x->hparams.clear();      // an accepted candidate replaces function K on the device
  rows_changed(x, x->c.K - 1);
  ecg::Screen& sc = x->scr;
  const int npt = x->npt;
  HIPCHK(hipMemcpyAsync(sc.call, cand, sizeof(double) * npt * ncand * nsteps, hipMemcpyHostToDevice, x->stream));
  // the shapes are timed on the first step's candidates
  HIPCHK(hipMemcpyAsync(sc.cparams, sc.call, sizeof(double) * npt * ncand, hipMemcpyDeviceToDevice, x->stream));
  RCCHK(screen_shapes(x, ncand));
  double st1 = 0.0;
  HIPCHK(hipMemcpyAsync(sc.state + 1, &st1, sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  // one step as a graph: fetch, rows, Schur complements, iterations,
  // Rayleigh quotients, selection; replayed nsteps times
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  HIPCHK(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
  int rc = ecg::screen_batch(x->c, x->f, sc, ncand, iters, true, npt, x->stream);
  if (!rc) rc = ecg::screen_select(x->c, sc, ncand, npt, x->stream);
  const hipError_t ec = hipStreamEndCapture(x->stream, &g);
  if (rc || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return rc < 0 ? rc : ECG_ERR_HIP;
  }
  if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    return ECG_ERR_HIP;
  }
  HIPCHK(hipEventRecord(x->e[0], x->stream));
  for (int i = 0; i < nsteps; ++i) {
    if (hipGraphLaunch(ge, x->stream) != hipSuccess) {
      (void)hipStreamSynchronize(x->stream);
      (void)hipGraphExecDestroy(ge);
      (void)hipGraphDestroy(g);
      return ECG_ERR_HIP;
    }
  }
  HIPCHK(hipEventRecord(x->e[1], x->stream));
  std::vector<double> rec(2 * (size_t)nsteps);
  HIPCHK(hipMemcpyAsync(rec.data(), sc.rec, sizeof(double) * 2 * nsteps, hipMemcpyDeviceToHost, x->stream));
  // accepted candidates changed function K: refresh its pair records
  RCCHK(ecg::prep(x->c, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  (void)hipGraphExecDestroy(ge);
  (void)hipGraphDestroy(g);
  x->tms[0] = ev_ms(x->e[0], x->e[1]);
  x->tms[1] = sc.tm;
  x->tms[2] = sc.cpb;
  x->have_last = false;
  x->vec_current = false;
  for (int i = 0; i < nsteps; ++i) {
    Ebest[i] = rec[2 * i];
    ibest[i] = (int)rec[2 * i + 1];
  }
  return ECG_OK;
}

int ecg_set_overlap_penalty(ecg_ctx* x, int allowed, double max_pair_penalty, double threshold2, int nfru) {
  if (!x || nfru < 0 || !std::isfinite(max_pair_penalty) || !std::isfinite(threshold2)) return ECG_ERR_ARG;
  x->pen_on = allowed != 0;
  x->pen_b = max_pair_penalty;
  x->pen_t2 = threshold2;
  x->pen_nfru = nfru;
  return ECG_OK;
}

int ecg_overlap_penalty(ecg_ctx* x, int nfru, double max_pair_penalty, double threshold2, double* total,
                        double* grad) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !total || nfru < 0 || !std::isfinite(max_pair_penalty) || !std::isfinite(threshold2))
    return ECG_ERR_ARG;
  const int K = x->c.K;
  if (K < 1 || !x->s_ready) return ECG_ERR_STATE;
  if (nfru >= K) return ECG_ERR_ARG;
  const size_t ng = (size_t)x->npt * (K - nfru);
  if (grad) {
    if (!finite_all(grad, ng)) return ECG_ERR_ARG;
    if (ng > x->grad_cap) {
      RCCHK(dev_alloc(&x->grad, ng));
      x->grad_cap = ng;
    }
    HIPCHK(hipMemcpyAsync(x->grad, grad, sizeof(double) * ng, hipMemcpyHostToDevice, x->stream));
  }
  RCCHK(penalty_enqueue(x, x->stream, nfru, max_pair_penalty, threshold2, grad ? x->grad : nullptr));
  if (grad) HIPCHK(hipMemcpyAsync(grad, x->grad, sizeof(double) * ng, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  *total = x->pen_total;
  if (!std::isfinite(*total) || (grad && !finite_all(grad, ng))) return ECG_ERR_NONFINITE;
  return ECG_OK;
}

int ecg_expectation_values(ecg_ctx* x, const double* c, double* values, double* errors) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !values) return ECG_ERR_ARG;
  ecg::Ctx& cx = x->c;
  const int K = cx.K;
  if (K < 1 || !x->params_set) return ECG_ERR_STATE;
  if (!c && !(x->vec_current && x->last_K == K)) return ECG_ERR_STATE;
  if (c && !finite_all(c, 2 * (size_t)K)) return ECG_ERR_ARG;
  const int m = cx.n * (cx.n + 1) / 2, nobs = 5 * m + 4;
  const size_t need = ecg::expectation_partial_doubles(cx.n, K) + 2 * (size_t)nobs;
  if (need > x->xpart_cap) {
    x->xpart_cap = 0;   // stays 0 if the allocation fails
    RCCHK(dev_alloc(&x->xpart, need));
    x->xpart_cap = need;
  }
  if (x->hs_rows < K) {   // diagS (the normalisation) from the current parameters
    hs_rewritten(x, x->hs_rows);
    RCCHK(ecg::assemble(cx, x->hs_rows, K, x->stream));
    x->hs_rows = K;
  }
  const zd* cv = x->ev[2];
  if (c) {
    HIPCHK(hipMemcpyAsync(x->ev[1], c, sizeof(zd) * K, hipMemcpyHostToDevice, x->stream));
    cv = x->ev[1];
  }
  double* dout = x->xpart + ecg::expectation_partial_doubles(cx.n, K);
  RCCHK(ecg::expectation_values(cx, cv, x->xpart, dout, x->stream));
  std::vector<double> h(2 * (size_t)nobs);
  HIPCHK(hipMemcpyAsync(h.data(), dout, sizeof(double) * 2 * nobs, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  for (int a = 0; a < nobs; ++a) {
    values[a] = h[a];
    if (errors) errors[a] = h[nobs + a] * 2.220446049250313e-16;   // epsilon(1.0)
  }
  return finite_all(values, nobs) ? ECG_OK : ECG_ERR_NONFINITE;
}

// Shared set-up of the post-processing values (correlation functions,
// densities, relativistic terms): the partial-sum buffer, diagS and the
// vector on the device.
// npart: doubles of per-tile partials the kernel needs (default: those of
// the correlation-function and expectation kernels)
static int post_values_setup(ecg_ctx* x, const double* c, size_t nout, const zd** cv, double** dout,
                             size_t npart = 0) {
  ecg::Ctx& cx = x->c;
  const int K = cx.K;
  if (K < 1 || !x->params_set) return ECG_ERR_STATE;
  if (!c && !(x->vec_current && x->last_K == K)) return ECG_ERR_STATE;
  if (c && !finite_all(c, 2 * (size_t)K)) return ECG_ERR_ARG;
  if (npart == 0) npart = ecg::correlation_partial_doubles(cx.n, K);
  const size_t need = npart + nout;
  if (need > x->xpart_cap) {
    x->xpart_cap = 0;   // stays 0 if the allocation fails
    RCCHK(dev_alloc(&x->xpart, need));
    x->xpart_cap = need;
  }
  if (x->hs_rows < K) {   // diagS (the normalisation) from the current parameters
    hs_rewritten(x, x->hs_rows);
    RCCHK(ecg::assemble(cx, x->hs_rows, K, x->stream));
    x->hs_rows = K;
  }
  *cv = x->ev[2];
  if (c) {
    HIPCHK(hipMemcpyAsync(x->ev[1], c, sizeof(zd) * K, hipMemcpyHostToDevice, x->stream));
    *cv = x->ev[1];
  }
  *dout = x->xpart + npart;
  return ECG_OK;
}

static int grid_values_setup(ecg_ctx* x, const double* c, int ngrid, const double* grid, size_t nout,
                             const zd** cv, double** dout) {
  if (!x || ngrid < 1 || !grid) return ECG_ERR_ARG;
  if (!finite_all(grid, (size_t)ngrid)) return ECG_ERR_ARG;
  return post_values_setup(x, c, nout, cv, dout);
}

static bool masses_ok(const ecg_ctx* x, const double* masses) {
  for (int i = 0; i <= x->c.n; ++i)
    if (!(masses[i] > 0.0) || !std::isfinite(masses[i])) return false;
  return true;
}

int ecg_relativistic_corrections(ecg_ctx* x, const double* c, const double* masses, double* values, double* errors) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !values || !masses || !masses_ok(x, masses)) return ECG_ERR_ARG;
  const zd* cv = nullptr;
  double* dout = nullptr;
  RCCHK(post_values_setup(x, c, 4, &cv, &dout));
  RCCHK(ecg::relativistic_values(x->c, cv, masses, x->xpart, dout, x->stream));
  double h[4];
  HIPCHK(hipMemcpyAsync(h, dout, sizeof(h), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  values[0] = h[0];
  values[1] = h[1];
  if (errors) {
    errors[0] = h[2] * 2.220446049250313e-16;   // epsilon(1.0)
    errors[1] = h[3] * 2.220446049250313e-16;
  }
  return finite_all(values, 2) ? ECG_OK : ECG_ERR_NONFINITE;
}

int ecg_drachmanized_corrections(ecg_ctx* x, const double* c, const double* masses, double E, double* values,
                                 double* errors) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !values || !masses || !masses_ok(x, masses) || !std::isfinite(E)) return ECG_ERR_ARG;
  const int n = x->c.n, nv = n * (n + 1) / 2 + 3;
  const zd* cv = nullptr;
  double* dout = nullptr;
  RCCHK(post_values_setup(x, c, 2 * (size_t)nv, &cv, &dout, ecg::drachmanized_partial_doubles(n, x->c.K)));
  RCCHK(ecg::drachmanized_values(x->c, cv, masses, E, x->xpart, dout, x->stream));
  double h[2 * 18];
  HIPCHK(hipMemcpyAsync(h, dout, sizeof(double) * 2 * nv, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  for (int a = 0; a < nv; ++a) {
    values[a] = h[a];
    if (errors) errors[a] = h[nv + a] * 2.220446049250313e-16;   // epsilon(1.0)
  }
  return finite_all(values, nv) ? ECG_OK : ECG_ERR_NONFINITE;
}

int ecg_correlation_functions(ecg_ctx* x, const double* c, int ngrid, const double* grid, double* values) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !values) return ECG_ERR_ARG;
  const size_t nout = (size_t)(x->c.n * (x->c.n + 1) / 2) * (ngrid > 0 ? ngrid : 0);
  const zd* cv = nullptr;
  double* dout = nullptr;
  RCCHK(grid_values_setup(x, c, ngrid, grid, nout, &cv, &dout));
  RCCHK(ecg::correlation_functions(x->c, cv, ngrid, grid, x->xpart, dout, x->stream));
  HIPCHK(hipMemcpyAsync(values, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return finite_all(values, nout) ? ECG_OK : ECG_ERR_NONFINITE;
}

int ecg_particle_densities(ecg_ctx* x, const double* c, const double* masses, int ngrid, const double* grid,
                           double* values) {
  (void)hipGetLastError();   // start from a clean runtime error state
  if (!x || !values || !masses || !masses_ok(x, masses)) return ECG_ERR_ARG;
  const int npart = x->c.n + 1;
  const size_t nout = (size_t)npart * (ngrid > 0 ? ngrid : 0);
  const zd* cv = nullptr;
  double* dout = nullptr;
  RCCHK(grid_values_setup(x, c, ngrid, grid, nout, &cv, &dout));
  RCCHK(ecg::particle_densities(x->c, cv, masses, ngrid, grid, x->xpart, dout, x->stream));
  HIPCHK(hipMemcpyAsync(values, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return finite_all(values, nout) ? ECG_OK : ECG_ERR_NONFINITE;
}

int ecg_eigenvector(ecg_ctx* x, double* c) {
  if (!x || !c) return ECG_ERR_ARG;
  const int K = x->c.K;
  if (K < 1 || !x->vec_current || x->last_K != K) return ECG_ERR_STATE;
  HIPCHK(hipMemcpyAsync(c, x->ev[2], sizeof(zd) * K, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return ECG_OK;
}

int ecg_last_timings(ecg_ctx* x, double* ms, int nmax) {
  if (!x || !ms || nmax < 1) return ECG_ERR_ARG;
  const int n = std::min(nmax, 6);
  for (int i = 0; i < n; ++i) ms[i] = x->tms[i];
  return n;
}

}  // extern "C"
