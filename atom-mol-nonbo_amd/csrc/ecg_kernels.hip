// Pair-assembly kernels: per-function set-up, diagonal and lower-triangle
// pair tiles (energy), pair tiles with stored derivatives (drop-in for
// StoreHSD) and the fused adjoint pass that contracts the gradient with the
// eigenvector without storing per-pair derivatives.
//
// Storage (device, column-major, ld = K rounded up to a multiple of 64):
//   H, S   complex, normalised off-diagonal elements H_kl / sqrt(S_kk S_ll)
//          in both triangles (the upper one conjugated), unit diagonal of S,
//          diagH_k / diagS_k on the diagonal of H.
//   diagS  real S_kk (un-normalised), diagH real H_kk / S_kk.
// The reference stores the same numbers in the lower triangles only
// (src/matform.f90:11-51); the host copy-out restores that layout.
#include <algorithm>
#include <type_traits>

#include "ecg_internal.hpp"

namespace ecg {

// ----------------------------------------------------------- set-up
// Records of function k from its parameters p (vech L, vech B): bra record
// (Ck* packed, |det L|, L packed) and, for every Y^dagger Y term, the ket
// record (P^T (A + iB) P packed, M tCl).
template <int N>
__device__ __attribute__((always_inline)) inline void prep_one(const SysConst& sc, int k, const double* p,
                                                               double* __restrict__ bra, double* __restrict__ ket) {
  constexpr int m = N * (N + 1) / 2;
  double L[N][N], B[N][N], A[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) L[i][j] = 0.0;
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int i = j; i < N; ++i) {
      L[i][j] = p[vidx(N, i, j)];
      B[i][j] = B[j][i] = p[m + vidx(N, i, j)];
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double s = 0.0;
#pragma unroll
      for (int c = 0; c <= j; ++c) s = fma(L[i][c], L[j][c], s);
      A[i][j] = A[j][i] = s;
    }
  double det = 1.0;
#pragma unroll
  for (int i = 0; i < N; ++i) det *= L[i][i];
  double* br = bra + (size_t)k * BraRec<N>::stride;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      br[BraRec<N>::ck + 2 * sidx(i, j)] = A[i][j];
      br[BraRec<N>::ck + 2 * sidx(i, j) + 1] = -B[i][j];
      br[BraRec<N>::lp + sidx(i, j)] = L[i][j];
    }
  br[BraRec<N>::adet] = fabs(det);
  for (int t = 0; t < sc.T; ++t) {
    double* kr = ket + ((size_t)k * sc.T + t) * KetRec<N>::stride;
    double tA[N][N], tB[N][N], W1[N][N], W2[N][N];
    // P^T X P with integer P
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int c = 0; c < N; ++c) {
          a += double(sc.P[t][c][i]) * A[c][j];
          b += double(sc.P[t][c][i]) * B[c][j];
        }
        W1[i][j] = a;
        W2[i][j] = b;
      }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int c = 0; c < N; ++c) {
          a += W1[i][c] * double(sc.P[t][c][j]);
          b += W2[i][c] * double(sc.P[t][c][j]);
        }
        tA[i][j] = a;
        tB[i][j] = b;
      }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        kr[KetRec<N>::ct + 2 * sidx(i, j)] = tA[i][j];
        kr[KetRec<N>::ct + 2 * sidx(i, j) + 1] = tB[i][j];
      }
    // M tCl, row-major
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int b = 0; b < N; ++b) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int c = 0; c < N; ++c) {
          re = fma(sc.M[a][c], tA[c][b], re);
          im = fma(sc.M[a][c], tB[c][b], im);
        }
        kr[KetRec<N>::nl + 2 * (a * N + b)] = re;
        kr[KetRec<N>::nl + 2 * (a * N + b) + 1] = im;
      }
  }
}

// One thread per function k in [f0, f1).
template <int N>
__global__ void k_prep(const SysConst sc, int f0, int f1, const double* __restrict__ params,
                       double* __restrict__ bra, double* __restrict__ ket) {
  constexpr int m = N * (N + 1) / 2;
  const int k = f0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= f1) return;
  prep_one<N>(sc, k, params + (size_t)k * 2 * m, bra, ket);
}

// The parameters of functions f0..f0+count-1 passed by value: stored into
// params and their records computed in one launch, with no copy from host
// memory to wait for.
template <int N>
__global__ void k_prep_put(const SysConst sc, int f0, int count, const ParamBlock pb, double* __restrict__ params,
                           double* __restrict__ bra, double* __restrict__ ket) {
  constexpr int m = N * (N + 1) / 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  double v[2 * m];
#pragma unroll
  for (int e = 0; e < 2 * m; ++e) v[e] = pb.v[i * 2 * m + e];
  double* dst = params + (size_t)(f0 + i) * 2 * m;
#pragma unroll
  for (int e = 0; e < 2 * m; ++e) dst[e] = v[e];
  prep_one<N>(sc, f0 + i, v, bra, ket);
}

// Sum over Y^dagger Y terms of one pair (energy only).
template <int N>
__device__ inline void pair_sum_energy(const SysConst& sc, const double* br, const double* kbase, double adet_l,
                                       zd& H, zd& S) {
  H = mkz(0.0, 0.0);
  S = mkz(0.0, 0.0);
  const zd* ck = reinterpret_cast<const zd*>(br + BraRec<N>::ck);
  const double adet = br[BraRec<N>::adet] * adet_l;
  for (int t = 0; t < sc.T; ++t) {
    const double* kr = kbase + (size_t)t * KetRec<N>::stride;
    PairCore<N> pc;
    pair_energy<N>(sc, ck, reinterpret_cast<const zd*>(kr + KetRec<N>::ct),
                   reinterpret_cast<const zd*>(kr + KetRec<N>::nl), adet, pc);
    H = H + sc.coef[t] * (pc.T + pc.V);
    S = S + sc.coef[t] * pc.S;
  }
}

// ----------------------------------------------------- diagonal pairs
// S_kk, H_kk of rows r0..r1-1; diagS, diagH and the unit/normalised
// diagonal of the device H, S.
template <int N>
__global__ void __launch_bounds__(128) k_diag(const SysConst sc, int r0, int r1, int ld, const double* __restrict__ bra,
                       const double* __restrict__ ket, double* __restrict__ diagS,
                       double* __restrict__ diagH, zd* __restrict__ Hm, zd* __restrict__ Sm) {
  const int k = r0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= r1) return;
  const double* br = bra + (size_t)k * BraRec<N>::stride;
  zd H, S;
  pair_sum_energy<N>(sc, br, ket + (size_t)k * sc.T * KetRec<N>::stride, br[BraRec<N>::adet], H, S);
  const double s = S.x;
  const double h = H.x / s;
  diagS[k] = s;
  diagH[k] = h;
  Hm[(size_t)k * ld + k] = mkz(h, 0.0);
  Sm[(size_t)k * ld + k] = mkz(1.0, 0.0);
}

// Tile coordinates of the b-th lower-triangle tile among tile rows
// >= tr0 (tile row tr holds tiles tc = 0..tr).
__device__ inline void tri_tile(int b, int tr0, int& tr, int& tc) {
  const long long base = (long long)tr0 * (tr0 + 1) / 2;
  const long long x = base + b;
  int r = (int)((sqrt(8.0 * (double)x + 1.0) - 1.0) * 0.5);
  while ((long long)(r + 1) * (r + 2) / 2 <= x) ++r;
  while ((long long)r * (r + 1) / 2 > x) --r;
  tr = r;
  tc = (int)(x - (long long)r * (r + 1) / 2);
}

// ---------------------------------------------- lower-triangle pair tiles
// Block = TILE x TILE pairs; thread x -> row k (bra), thread y -> column l
// (ket).  Bra and ket records of the tile are staged in LDS.
template <int N>
__device__ __attribute__((always_inline)) inline void offdiag_tile(const SysConst& sc, int r0, int r1, int ld, int tr,
                                                                   int tc, const double* __restrict__ bra,
                                                                   const double* __restrict__ ket,
                                                                   const double* __restrict__ diagS,
                                                                   zd* __restrict__ Hm, zd* __restrict__ Sm,
                                                                   double* lds) {
  const int T = sc.T;
  const int kb = tr * TILE, lb = tc * TILE;
  double* sbra = lds;
  double* sket = lds + TILE * BraRec<N>::stride;
  const int tid = threadIdx.y * TILE + threadIdx.x;
  for (int i = tid; i < TILE * BraRec<N>::stride; i += TILE * TILE) {
    const int f = kb + i / BraRec<N>::stride;
    sbra[i] = (f < r1) ? bra[(size_t)kb * BraRec<N>::stride + i] : 0.0;
  }
  const int kstride = T * KetRec<N>::stride;
  for (int i = tid; i < TILE * kstride; i += TILE * TILE) {
    const int f = lb + i / kstride;
    sket[i] = (f < r1) ? ket[(size_t)lb * kstride + i] : 0.0;
  }
  __syncthreads();
  const int k = kb + threadIdx.x, l = lb + threadIdx.y;
  if (k < r0 || k >= r1 || l >= k) return;
  const double* br = sbra + threadIdx.x * BraRec<N>::stride;
  const double* kr = sket + threadIdx.y * kstride;
  zd H, S;
  pair_sum_energy<N>(sc, br, kr, bra[(size_t)l * BraRec<N>::stride + BraRec<N>::adet], H, S);
  const double f = 1.0 / sqrt(diagS[k] * diagS[l]);
  const zd hn = H * f, sn = S * f;
  Hm[(size_t)l * ld + k] = hn;
  Sm[(size_t)l * ld + k] = sn;
  Hm[(size_t)k * ld + l] = zconj(hn);
  Sm[(size_t)k * ld + l] = zconj(sn);
}

template <int N>
__global__ void __launch_bounds__(TILE* TILE) k_offdiag(const SysConst sc, int r0, int r1, int ld, int tr0,
                                                      const double* __restrict__ bra,
                                                      const double* __restrict__ ket,
                                                      const double* __restrict__ diagS,
                                                      zd* __restrict__ Hm, zd* __restrict__ Sm) {
  extern __shared__ double lds[];
  int tr, tc;
  tri_tile(blockIdx.x, tr0, tr, tc);
  offdiag_tile<N>(sc, r0, r1, ld, tr, tc, bra, ket, diagS, Hm, Sm, lds);
}

// ------------------------------------------- stored derivatives (StoreHSD)
// Per-pair dH, dS with respect to both functions, summed over terms.
template <int N>
__device__ inline void pair_sum_deriv(const SysConst& sc, const double* br, const double* kbase,
                                      const double* bra_l, bool want_l, zd& H, zd& S, zd* Dk, zd* Dl) {
  constexpr int m = N * (N + 1) / 2;
  const zd* ck = reinterpret_cast<const zd*>(br + BraRec<N>::ck);
  const double adet = br[BraRec<N>::adet] * bra_l[BraRec<N>::adet];
  zd GHb[m], GHk[m], GSs[m], GSk[m];
#pragma unroll
  for (int e = 0; e < m; ++e) GHb[e] = GHk[e] = GSs[e] = GSk[e] = mkz(0.0, 0.0);
  H = S = mkz(0.0, 0.0);
  for (int t = 0; t < sc.T; ++t) {
    const double* kr = kbase + (size_t)t * KetRec<N>::stride;
    const zd* ct = reinterpret_cast<const zd*>(kr + KetRec<N>::ct);
    PairCore<N> pc;
    pair_energy<N>(sc, ck, ct, reinterpret_cast<const zd*>(kr + KetRec<N>::nl), adet, pc);
    zd gb[m], gk[m];
    pair_adjoint<N>(sc, ck, ct, pc, 0.0, gb, gk);
    const double c = sc.coef[t];
    H = H + c * (pc.T + pc.V);
    S = S + c * pc.S;
    const zd cS = (-1.5 * c) * pc.S;
#pragma unroll
    for (int e = 0; e < m; ++e) {
      GHb[e] = GHb[e] + c * gb[e];
      GSs[e] = zfma(cS, pc.Ci[e], GSs[e]);
    }
    if (want_l) {
      zd g2[m], s2[m], gs[m];
#pragma unroll
      for (int e = 0; e < m; ++e) gs[e] = cS * pc.Ci[e];
      perm_congruence<N>(sc.P[t], gk, g2);
      perm_congruence<N>(sc.P[t], gs, s2);
#pragma unroll
      for (int e = 0; e < m; ++e) {
        GHk[e] = GHk[e] + c * g2[e];
        GSk[e] = GSk[e] + s2[e];
      }
    }
  }
  const double* Lk = br + BraRec<N>::lp;
  const zd aH = 1.5 * H, aS = 1.5 * S;
  chain_to_params<N>(GHb, aH, Lk, -1.0, Dk, Dk + m);
  chain_to_params<N>(GSs, aS, Lk, -1.0, Dk + 2 * m, Dk + 3 * m);
  if (want_l) {
    const double* Ll = bra_l + BraRec<N>::lp;
    chain_to_params<N>(GHk, aH, Ll, 1.0, Dl, Dl + m);
    chain_to_params<N>(GSk, aS, Ll, 1.0, Dl + 2 * m, Dl + 3 * m);
  }
}

// Rows k in [nfru, K) with all l <= k; D layout (2 npt) x nfo x ld complex,
// as Glob_D (src/globvars.f90:211-227, src/matform.f90:55-115).
template <int N>
__global__ void __launch_bounds__(128) k_deriv(const SysConst sc, int K, int nfru, int ld, const double* __restrict__ bra,
                        const double* __restrict__ ket, const double* __restrict__ diagS,
                        zd* __restrict__ Hm, zd* __restrict__ Sm, zd* __restrict__ D) {
  constexpr int m = N * (N + 1) / 2, np4 = 4 * m;
  const int T = sc.T;
  const int nfo = K - nfru;
  // rows k = nfru..K-1, columns l = 0..k (diagonal included), row by row
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long base = (long long)nfru * (nfru + 1) / 2;
  const long long x = base + idx;
  const long long total = (long long)K * (K + 1) / 2;
  if (x >= total) return;
  int k = (int)((sqrt(8.0 * (double)x + 1.0) - 1.0) * 0.5);
  while ((long long)(k + 1) * (k + 2) / 2 <= x) ++k;
  while ((long long)k * (k + 1) / 2 > x) --k;
  const int l = (int)(x - (long long)k * (k + 1) / 2);
  const double* br = bra + (size_t)k * BraRec<N>::stride;
  const double* brl = bra + (size_t)l * BraRec<N>::stride;
  const bool want_l = (l >= nfru) && (l != k);
  zd H, S, Dk[np4], Dl[np4];
  pair_sum_deriv<N>(sc, br, ket + (size_t)l * T * KetRec<N>::stride, brl, want_l, H, S, Dk, Dl);
  zd* Dcol = D + (size_t)np4 * ((size_t)(k - nfru) + (size_t)nfo * l);
  if (k == l) {
    const double s = diagS[k];
#pragma unroll
    for (int e = 0; e < np4; ++e) Dcol[e] = mkz(2.0 * Dk[e].x / s, 0.0);
    return;
  }
  const double f = 1.0 / sqrt(diagS[k] * diagS[l]);
#pragma unroll
  for (int e = 0; e < np4; ++e) Dcol[e] = Dk[e] * f;
  if (want_l) {
    zd* Dc2 = D + (size_t)np4 * ((size_t)(l - nfru) + (size_t)nfo * k);
#pragma unroll
    for (int e = 0; e < np4; ++e) Dc2[e] = zconj(Dl[e] * f);
  }
  const zd hn = H * f, sn = S * f;
  Hm[(size_t)l * ld + k] = hn;
  Sm[(size_t)l * ld + k] = sn;
  Hm[(size_t)k * ld + l] = zconj(hn);
  Sm[(size_t)k * ld + l] = zconj(sn);
}

// ---------------------------------------------------- fused adjoint pass
// For a pair (k bra, l ket) with weight w = conj(c_k) c_l / sqrt(S_kk S_ll)
// the gradient picks up 2 Re(w dh) with h = H - E S.  Per function the
// contributions reduce to 13 reals: RA (Re of the adjoint, both roles),
// IB (Im of the adjoint, bra minus ket) and Ra (Re of the det-term scalar).
template <int N> struct AdjAcc {
  static constexpr int m = N * (N + 1) / 2;
  static constexpr int nvals = 2 * m + 1;
};

template <int N>
__device__ inline void pair_adjoint_sum(const SysConst& sc, const double* br, const double* kbase,
                                        double adet_l, double E, bool want_l, zd w, double* vb, double* vk) {
  constexpr int m = N * (N + 1) / 2;
  const zd* ck = reinterpret_cast<const zd*>(br + BraRec<N>::ck);
  const double adet = br[BraRec<N>::adet] * adet_l;
  zd Gb[m], Gk[m];
#pragma unroll
  for (int e = 0; e < m; ++e) Gb[e] = Gk[e] = mkz(0.0, 0.0);
  zd hs = mkz(0.0, 0.0);
  for (int t = 0; t < sc.T; ++t) {
    const double* kr = kbase + (size_t)t * KetRec<N>::stride;
    const zd* ct = reinterpret_cast<const zd*>(kr + KetRec<N>::ct);
    PairCore<N> pc;
    pair_energy<N>(sc, ck, ct, reinterpret_cast<const zd*>(kr + KetRec<N>::nl), adet, pc);
    zd gb[m], gk[m];
    pair_adjoint<N>(sc, ck, ct, pc, E, gb, gk);
    const double c = sc.coef[t];
    hs = hs + c * (pc.T + pc.V - E * pc.S);
#pragma unroll
    for (int e = 0; e < m; ++e) Gb[e] = Gb[e] + c * gb[e];
    if (want_l) {
      zd g2[m];
      perm_congruence<N>(sc.P[t], gk, g2);
#pragma unroll
      for (int e = 0; e < m; ++e) Gk[e] = Gk[e] + c * g2[e];
    }
  }
  const zd a = w * (1.5 * hs);
#pragma unroll
  for (int e = 0; e < m; ++e) {
    const zd g = w * Gb[e];
    vb[e] = g.x;
    vb[m + e] = g.y;
  }
  vb[2 * m] = a.x;
  if (want_l) {
#pragma unroll
    for (int e = 0; e < m; ++e) {
      const zd g = w * Gk[e];
      vk[e] = g.x;
      vk[m + e] = -g.y;
    }
    vk[2 * m] = a.x;
  } else {
#pragma unroll
    for (int e = 0; e < 2 * m + 1; ++e) vk[e] = 0.0;
  }
}

// Tile pass over all pairs with k >= nfru (rows) and l <= k.  Partials:
// part[(f * (nt + 1) + slot) * nv + v]; function f in tile row r gets its
// bra partial from tile (r, c) in slot c and its ket partial from tile
// (r', r) in slot r' + 1.
template <int N>
__device__ __attribute__((always_inline)) inline void adjoint_tile(const SysConst& sc, int K, int nfru, int nt,
                                                                   int tr, int tc, double E,
                                                                   const double* __restrict__ bra,
                                                                   const double* __restrict__ ket,
                                                                   const double* __restrict__ diagS,
                                                                   const zd* __restrict__ cvec,
                                                                   double* __restrict__ part, double* lds) {
  constexpr int nv = AdjAcc<N>::nvals;
  const int T = sc.T;
  const int kb = tr * TILE, lb = tc * TILE;
  double* sbra = lds;
  double* sket = lds + TILE * BraRec<N>::stride;
  double* sred = sket + TILE * T * KetRec<N>::stride;   // TILE*TILE*nv
  const int tid = threadIdx.y * TILE + threadIdx.x;
  for (int i = tid; i < TILE * BraRec<N>::stride; i += TILE * TILE) {
    const int f = kb + i / BraRec<N>::stride;
    sbra[i] = (f < K) ? bra[(size_t)kb * BraRec<N>::stride + i] : 0.0;
  }
  const int kstride = T * KetRec<N>::stride;
  for (int i = tid; i < TILE * kstride; i += TILE * TILE) {
    const int f = lb + i / kstride;
    sket[i] = (f < K) ? ket[(size_t)lb * kstride + i] : 0.0;
  }
  __syncthreads();
  const int k = kb + threadIdx.x, l = lb + threadIdx.y;
  double vb[nv], vk[nv];
  const bool active = (k >= nfru) && (k < K) && (l <= k);
  if (active) {
    const double* br = sbra + threadIdx.x * BraRec<N>::stride;
    const double* kr = sket + threadIdx.y * kstride;
    const zd ckc = zconj(cvec[k]);
    const zd w = (ckc * cvec[l]) * (1.0 / sqrt(diagS[k] * diagS[l]));
    const bool want_l = (l != k) && (l >= nfru);
    pair_adjoint_sum<N>(sc, br, kr, bra[(size_t)l * BraRec<N>::stride + BraRec<N>::adet], E, want_l, w,
                        vb, vk);
  } else {
#pragma unroll
    for (int e = 0; e < nv; ++e) vb[e] = vk[e] = 0.0;
  }
  // reduce bra partials over threadIdx.y (same k), ket partials over x
  for (int e = 0; e < nv; ++e) sred[(threadIdx.y * TILE + threadIdx.x) * nv + e] = vb[e];
  __syncthreads();
  if (threadIdx.y == 0 && kb + (int)threadIdx.x < K) {
    const int f = kb + threadIdx.x;
    double* out = part + ((size_t)f * (nt + 1) + tc) * nv;
    for (int e = 0; e < nv; ++e) {
      double s = 0.0;
      for (int y = 0; y < TILE; ++y) s += sred[(y * TILE + threadIdx.x) * nv + e];
      out[e] = s;
    }
  }
  __syncthreads();
  for (int e = 0; e < nv; ++e) sred[(threadIdx.x * TILE + threadIdx.y) * nv + e] = vk[e];
  __syncthreads();
  if (threadIdx.y == 0 && lb + (int)threadIdx.x < K) {
    const int f = lb + threadIdx.x;
    double* out = part + ((size_t)f * (nt + 1) + tr + 1) * nv;
    for (int e = 0; e < nv; ++e) {
      double s = 0.0;
      for (int y = 0; y < TILE; ++y) s += sred[(y * TILE + threadIdx.x) * nv + e];
      out[e] = s;
    }
  }
}

template <int N>
__global__ void __launch_bounds__(TILE* TILE, 2) k_adjoint(const SysConst sc, int K, int nfru, int tr0, int nt,
                                                      double E, const double* __restrict__ Ep,
                                                      const double* __restrict__ bra,
                                                      const double* __restrict__ ket,
                                                      const double* __restrict__ diagS,
                                                      const zd* __restrict__ cvec,
                                                      double* __restrict__ part) {
  extern __shared__ double lds[];
  int tr, tc;
  tri_tile(blockIdx.x, tr0, tr, tc);
  adjoint_tile<N>(sc, K, nfru, nt, tr, tc, Ep ? *Ep : E, bra, ket, diagS, cvec, part, lds);
}

// Per function f >= nfru: sum its partial slots in fixed order and apply
// the chain rule to vech L, vech B:
//   dE/dL(i,j) = 4 (RA L)(i,j) + 2 Ra delta_ij / L_jj
//   dE/dB(i,i) = 2 IB_ii,   dE/dB(i,j) = 4 IB_ij
// One wave per function: lanes take the slots s = lane, lane + 64, ...
// (contiguous in memory for one function), then a fixed-order butterfly
// reduction; lane 0 applies the chain rule.
template <int N>
__device__ __attribute__((always_inline)) inline void adjoint_finish_one(int f, int nfru, int nt, int tr0,
                                                                         const double* __restrict__ bra,
                                                                         const double* __restrict__ part,
                                                                         double* __restrict__ grad) {
  constexpr int m = N * (N + 1) / 2, nv = AdjAcc<N>::nvals;
  const int r = f / TILE;
  const int lane = threadIdx.x & 63;
  double acc[nv];
#pragma unroll
  for (int e = 0; e < nv; ++e) acc[e] = 0.0;
  // valid slots: bra slots c = 0..r (tile rows >= tr0 only), ket slots r'+1 for r' = max(r,tr0)..nt-1
  for (int s = lane; s <= nt; s += 64) {
    bool valid;
    if (s <= r) valid = (r >= tr0);
    else valid = (s - 1 >= tr0);
    if (!valid) continue;
    const double* p = part + ((size_t)f * (nt + 1) + s) * nv;
#pragma unroll
    for (int e = 0; e < nv; ++e) acc[e] += p[e];
  }
#pragma unroll
  for (int e = 0; e < nv; ++e)
    for (int o = 32; o > 0; o >>= 1) acc[e] += __shfl_xor(acc[e], o);
  if (lane != 0) return;
  const double* L = bra + (size_t)f * BraRec<N>::stride + BraRec<N>::lp;
  double* g = grad + (size_t)(f - nfru) * 2 * m;
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int i = j; i < N; ++i) {
      double s = 0.0;
#pragma unroll
      for (int c = j; c < N; ++c) s = fma(acc[sidx(i, c)], L[sidx(c, j)], s);
      double v = 4.0 * s;
      if (i == j) v += 2.0 * acc[2 * m] / L[sidx(j, j)];
      g[vidx(N, i, j)] = v;
      g[m + vidx(N, i, j)] = ((i == j) ? 2.0 : 4.0) * acc[m + sidx(i, j)];
    }
}

template <int N>
__global__ void __launch_bounds__(256) k_adjoint_finish(int K, int nfru, int nt, int tr0,
                                                        const double* __restrict__ bra,
                                                        const double* __restrict__ part,
                                                        double* __restrict__ grad) {
  const int f = nfru + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (f >= K) return;
  adjoint_finish_one<N>(f, nfru, nt, tr0, bra, part, grad);
}

// ---------------------------------------------------- expectation values
// MatrixElementsForExpcVals (src/matelem.f90:1236-1349) for the operators
// 1/r^2, 1/r, r, r^2, delta(r) of every pseudo-particle pair (a, b >= a;
// a = b: r_a, the distance from particle 1), with rs = tr(C^-1 J_ab)^-1/2
// from the energy's Coulomb part:
//   1/r^2 = 2 S rs^2, 1/r = (2/sqrt(pi)) S rs, r = (2/sqrt(pi)) S / rs,
//   r^2 = (3/2) S / rs^2, delta(r) = S rs^3 / pi^(3/2);
// then H, S, T, V.  Contracted on the fly as the reference's driver does
// (src/workproc.f90:8158-8230): per pair k >= l and term t,
// beta = Re(conj(c_k) c_l O), mu = the driver's bound on its rounding,
// sum += f coef_t beta, err += |f coef_t (mu + 2 |beta|)|, f = 1/diagS_k
// (k = l) or 2/sqrt(diagS_k diagS_l).  Per tile: fixed-order reduction
// (wave butterflies, then the four waves in order) into part[tile][2 NOBS];
// k_expect_finish adds the tiles in order.
template <int N> struct Expc {
  static constexpr int m = N * (N + 1) / 2;
  static constexpr int nobs = 5 * m + 4;
};

template <int N>
__global__ void __launch_bounds__(TILE* TILE) k_expect(const SysConst sc, int K, const double* __restrict__ bra,
                                                     const double* __restrict__ ket,
                                                     const double* __restrict__ diagS, const zd* __restrict__ cvec,
                                                     double* __restrict__ part) {
  constexpr int m = Expc<N>::m, NO = Expc<N>::nobs;
  extern __shared__ double lds[];
  int tr, tc;
  tri_tile(blockIdx.x, 0, tr, tc);
  const int T = sc.T;
  const int kb = tr * TILE, lb = tc * TILE;
  double* sbra = lds;
  double* sket = lds + TILE * BraRec<N>::stride;
  const int tid = threadIdx.y * TILE + threadIdx.x;
  for (int i = tid; i < TILE * BraRec<N>::stride; i += TILE * TILE) {
    const int f = kb + i / BraRec<N>::stride;
    sbra[i] = (f < K) ? bra[(size_t)kb * BraRec<N>::stride + i] : 0.0;
  }
  const int kstride = T * KetRec<N>::stride;
  for (int i = tid; i < TILE * kstride; i += TILE * TILE) {
    const int f = lb + i / kstride;
    sket[i] = (f < K) ? ket[(size_t)lb * kstride + i] : 0.0;
  }
  __syncthreads();
  double sum[NO], err[NO];
#pragma unroll
  for (int a = 0; a < NO; ++a) sum[a] = err[a] = 0.0;
  const int k = kb + threadIdx.x, l = lb + threadIdx.y;
  if (k < K && l <= k) {
    const double* br = sbra + threadIdx.x * BraRec<N>::stride;
    const double* kr0 = sket + threadIdx.y * kstride;
    const zd* ck = reinterpret_cast<const zd*>(br + BraRec<N>::ck);
    const double adet = br[BraRec<N>::adet] * bra[(size_t)l * BraRec<N>::stride + BraRec<N>::adet];
    const double f = (k == l) ? 1.0 / diagS[k] : 2.0 / sqrt(diagS[k] * diagS[l]);
    const zd ci = cvec[k], cj = cvec[l];
    const double wr = ci.x * cj.x + ci.y * cj.y, wi = ci.x * cj.y - ci.y * cj.x;   // conj(c_k) c_l
    constexpr double two_over_sqrtpi = 1.12837916709551257389615890312154517;
    constexpr double inv_pi32 = 0.179587122125166564899490489694101733;    // pi^(-3/2)
    for (int t = 0; t < T; ++t) {
      const double* kr = kr0 + (size_t)t * KetRec<N>::stride;
      PairCore<N> pc;
      pair_energy<N>(sc, ck, reinterpret_cast<const zd*>(kr + KetRec<N>::ct),
                     reinterpret_cast<const zd*>(kr + KetRec<N>::nl), adet, pc);
      zd O[NO];
      int q = 0;
#pragma unroll
      for (int a = 0; a < N; ++a)
#pragma unroll
        for (int b = a; b < N; ++b, ++q) {
          const zd rs = pc.rs[(a == b) ? sidx(a, a) : sidx(b, a)];
          const zd rs2 = rs * rs;
          const zd Srs = pc.S * rs;
          O[q] = 2.0 * (pc.S * rs2);
          O[m + q] = two_over_sqrtpi * Srs;
          O[2 * m + q] = two_over_sqrtpi * (pc.S * zrcp(rs));
          O[3 * m + q] = 1.5 * (pc.S * zrcp(rs2));
          O[4 * m + q] = inv_pi32 * (Srs * rs2);
        }
      O[5 * m] = pc.T + pc.V;
      O[5 * m + 1] = pc.S;
      O[5 * m + 2] = pc.T;
      O[5 * m + 3] = pc.V;
      const double fc = f * sc.coef[t];
#pragma unroll
      for (int a = 0; a < NO; ++a) {
        const double ro = O[a].x, io = O[a].y;
        const double beta = wr * ro - wi * io;
        const double mu = fabs(ci.x * (cj.x * ro - cj.y * io)) + fabs(cj.x * (ci.x * ro + ci.y * io)) +
                          fabs(ci.y * (cj.y * ro + cj.x * io)) + fabs(cj.y * (ci.y * ro - ci.x * io));
        sum[a] += fc * beta;
        err[a] += fabs(fc * (mu + 2.0 * fabs(beta)));
      }
    }
  }
  // fixed-order reduction of the 256 pairs' sums
  double* red = lds;   // 4 waves x 2 NO (the records are no longer read)
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int a = 0; a < NO; ++a) {
    double v = sum[a], e = err[a];
    for (int o = 32; o > 0; o >>= 1) {
      v += __shfl_xor(v, o);
      e += __shfl_xor(e, o);
    }
    if (lane == 0) {
      red[wave * 2 * NO + a] = v;
      red[wave * 2 * NO + NO + a] = e;
    }
  }
  __syncthreads();
  for (int a = tid; a < 2 * NO; a += TILE * TILE)
    part[(size_t)blockIdx.x * 2 * NO + a] = ((red[a] + red[2 * NO + a]) + red[4 * NO + a]) + red[6 * NO + a];
}

__global__ void __launch_bounds__(256) k_expect_finish(int nvals, long long ntiles, const double* __restrict__ part,
                                                       double* __restrict__ out) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nvals) return;
  double s = 0.0;
  for (long long t = 0; t < ntiles; ++t) s += part[(size_t)t * nvals + a];
  out[a] = s;
}

template <int N>
static int launch_expect(const Ctx& c, const zd* cvec, double* part, double* out, hipStream_t s) {
  const int K = c.K;
  const int nt = (K + TILE - 1) / TILE;
  const long long ntiles = (long long)nt * (nt + 1) / 2;
  size_t lds = sizeof(double) * TILE * (BraRec<N>::stride + c.T * KetRec<N>::stride);
  lds = std::max(lds, sizeof(double) * 8 * Expc<N>::nobs);
  hipLaunchKernelGGL(k_expect<N>, dim3((unsigned)ntiles), dim3(TILE, TILE), lds, s, c.sc, K, c.bra, c.ket, c.diagS,
                     cvec, part);
  const int nv = 2 * Expc<N>::nobs;
  hipLaunchKernelGGL(k_expect_finish, dim3((nv + 255) / 256), dim3(256), 0, s, nv, ntiles, part, out);
  return (int)hipGetLastError();
}

int expectation_values(const Ctx& c, const zd* cvec, double* part, double* out, hipStream_t s) {
  switch (c.n) {
    case 1: return launch_expect<1>(c, cvec, part, out, s);
    case 2: return launch_expect<2>(c, cvec, part, out, s);
    case 3: return launch_expect<3>(c, cvec, part, out, s);
    case 4: return launch_expect<4>(c, cvec, part, out, s);
    case 5: return launch_expect<5>(c, cvec, part, out, s);
  }
  return -1;
}

size_t expectation_partial_doubles(int n, int K) {
  const int nt = (K + TILE - 1) / TILE;
  const int m = n * (n + 1) / 2;
  return (size_t)nt * (nt + 1) / 2 * 2 * (5 * m + 4);
}

// ------------------------------------------------ correlation functions
// Pair correlation functions on a radial grid, the AreCorrFuncNeeded part
// of MatrixElementsForExpcVals (src/matelem.f90:1934-1946):
//   CF_kl(q, r) = S_kl exp(-r^2 / TrCJ_q) / (sqrt(TrCJ_q) TrCJ_q) / pi^(3/2)
//              = S_kl rs_q^3 exp(-r^2 rs_q^2) / pi^(3/2),
// q over the pseudo-particle pairs (a, b >= a) as the other pair values,
// contracted with the vector like the driver's CFDMEkl_s
// (src/workproc.f90:8234-8243; no rounding estimate there).  CF_q(0) is
// <delta(r_q)>.  One launch per CF_CHUNK grid points (r^2 in r2), values
// (point, pair) with the pair index fastest; fixed-order reduction.
// Particle densities (ArePartDensNeeded, src/matelem.f90:1948-1963) are
// the same form with TrCJ replaced by b_i^T C^-1 b_i for the vectors b_i of
// Glob_bvc (src/workproc.f90:612-621), i = 1..n+1 (DENS = true).
constexpr int CF_CHUNK = 4;

struct DensVec {
  double b[6][5];   // b[i][p]: component p of particle i's vector
};

template <int N, bool DENS>
__global__ void __launch_bounds__(TILE* TILE) k_corr(const SysConst sc, int K, const double* __restrict__ bra,
                                                   const double* __restrict__ ket,
                                                   const double* __restrict__ diagS, const zd* __restrict__ cvec,
                                                   double r2_0, double r2_1, double r2_2, double r2_3, int ng,
                                                   const DensVec dv, double* __restrict__ part) {
  constexpr int m = DENS ? N + 1 : N * (N + 1) / 2, NV = CF_CHUNK * m;
  extern __shared__ double lds[];
  int tr, tc;
  tri_tile(blockIdx.x, 0, tr, tc);
  const int T = sc.T;
  const int kb = tr * TILE, lb = tc * TILE;
  double* sbra = lds;
  double* sket = lds + TILE * BraRec<N>::stride;
  const int tid = threadIdx.y * TILE + threadIdx.x;
  for (int i = tid; i < TILE * BraRec<N>::stride; i += TILE * TILE) {
    const int f = kb + i / BraRec<N>::stride;
    sbra[i] = (f < K) ? bra[(size_t)kb * BraRec<N>::stride + i] : 0.0;
  }
  const int kstride = T * KetRec<N>::stride;
  for (int i = tid; i < TILE * kstride; i += TILE * TILE) {
    const int f = lb + i / kstride;
    sket[i] = (f < K) ? ket[(size_t)lb * kstride + i] : 0.0;
  }
  __syncthreads();
  const double r2[CF_CHUNK] = {r2_0, r2_1, r2_2, r2_3};
  double sum[NV];
#pragma unroll
  for (int a = 0; a < NV; ++a) sum[a] = 0.0;
  const int k = kb + threadIdx.x, l = lb + threadIdx.y;
  if (k < K && l <= k) {
    const double* br = sbra + threadIdx.x * BraRec<N>::stride;
    const double* kr0 = sket + threadIdx.y * kstride;
    const zd* ck = reinterpret_cast<const zd*>(br + BraRec<N>::ck);
    const double adet = br[BraRec<N>::adet] * bra[(size_t)l * BraRec<N>::stride + BraRec<N>::adet];
    const double f = (k == l) ? 1.0 / diagS[k] : 2.0 / sqrt(diagS[k] * diagS[l]);
    const zd ci = cvec[k], cj = cvec[l];
    const double wr = ci.x * cj.x + ci.y * cj.y, wi = ci.x * cj.y - ci.y * cj.x;   // conj(c_k) c_l
    constexpr double inv_pi32 = 0.179587122125166564899490489694101733;    // pi^(-3/2)
    for (int t = 0; t < T; ++t) {
      const double* kr = kr0 + (size_t)t * KetRec<N>::stride;
      PairCore<N> pc;
      pair_energy<N>(sc, ck, reinterpret_cast<const zd*>(kr + KetRec<N>::ct),
                     reinterpret_cast<const zd*>(kr + KetRec<N>::nl), adet, pc);
      const double fc = f * sc.coef[t];
      auto accumulate = [&](int q, zd rs) {
        const zd rs2 = rs * rs;
        const zd A = inv_pi32 * ((pc.S * rs) * rs2);
#pragma unroll
        for (int g = 0; g < CF_CHUNK; ++g) {
          if (g >= ng) break;
          // exp(-r^2 rs^2) for complex rs^2
          const double er = exp(-r2[g] * rs2.x);
          double sn, cs;
          sincos(-r2[g] * rs2.y, &sn, &cs);
          const zd O = A * mkz(er * cs, er * sn);
          sum[g * m + q] += fc * (wr * O.x - wi * O.y);
        }
      };
      if constexpr (DENS) {
#pragma unroll
        for (int i = 0; i <= N; ++i) {
          // b_i^T C^-1 b_i in the reference's order (diagonal, then twice the lower part)
          zd tq = mkz(0.0, 0.0);
#pragma unroll
          for (int p = 0; p < N; ++p) {
            tq = tq + (dv.b[i][p] * dv.b[i][p]) * pc.Ci[sidx(p, p)];
#pragma unroll
            for (int q = p + 1; q < N; ++q) tq = tq + (2.0 * dv.b[i][q] * dv.b[i][p]) * pc.Ci[sidx(q, p)];
          }
          accumulate(i, zrcp(zsqrt(tq)));
        }
      } else {
        int q = 0;
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int b = a; b < N; ++b, ++q) accumulate(q, pc.rs[(a == b) ? sidx(a, a) : sidx(b, a)]);
      }
    }
  }
  double* red = lds;   // 4 waves x NV
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    double v = sum[a];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) red[wave * NV + a] = v;
  }
  __syncthreads();
  for (int a = tid; a < NV; a += TILE * TILE)
    part[(size_t)blockIdx.x * NV + a] = ((red[a] + red[NV + a]) + red[2 * NV + a]) + red[3 * NV + a];
}

__global__ void __launch_bounds__(256) k_corr_finish(int nv, int ncopy, long long ntiles, const double* __restrict__ part,
                                                     double* __restrict__ out) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ncopy) return;
  double s = 0.0;
  for (long long t = 0; t < ntiles; ++t) s += part[(size_t)t * nv + a];
  out[a] = s;
}

template <int N, bool DENS>
static int launch_corr(const Ctx& c, const zd* cvec, int ngrid, const double* grid, const DensVec& dv, double* part,
                       double* out, hipStream_t s) {
  constexpr int m = DENS ? N + 1 : N * (N + 1) / 2, NV = CF_CHUNK * m;
  const int K = c.K;
  const int nt = (K + TILE - 1) / TILE;
  const long long ntiles = (long long)nt * (nt + 1) / 2;
  size_t lds = sizeof(double) * TILE * (BraRec<N>::stride + c.T * KetRec<N>::stride);
  lds = std::max(lds, sizeof(double) * 4 * NV);
  for (int g0 = 0; g0 < ngrid; g0 += CF_CHUNK) {
    double r2[CF_CHUNK];
    const int ng = std::min(CF_CHUNK, ngrid - g0);
    for (int g = 0; g < CF_CHUNK; ++g) r2[g] = (g < ng) ? grid[g0 + g] * grid[g0 + g] : 0.0;
    hipLaunchKernelGGL((k_corr<N, DENS>), dim3((unsigned)ntiles), dim3(TILE, TILE), lds, s, c.sc, K, c.bra, c.ket,
                       c.diagS, cvec, r2[0], r2[1], r2[2], r2[3], ng, dv, part);
    // the chunk's values (ng of its CF_CHUNK points) go to their place in out
    hipLaunchKernelGGL(k_corr_finish, dim3((ng * m + 255) / 256), dim3(256), 0, s, NV, ng * m, ntiles, part,
                       out + (size_t)g0 * m);
    if (ng < CF_CHUNK) break;
  }
  return (int)hipGetLastError();
}

int correlation_functions(const Ctx& c, const zd* cvec, int ngrid, const double* grid, double* part, double* out,
                          hipStream_t s) {
  const DensVec dv{};
  switch (c.n) {
    case 1: return launch_corr<1, false>(c, cvec, ngrid, grid, dv, part, out, s);
    case 2: return launch_corr<2, false>(c, cvec, ngrid, grid, dv, part, out, s);
    case 3: return launch_corr<3, false>(c, cvec, ngrid, grid, dv, part, out, s);
    case 4: return launch_corr<4, false>(c, cvec, ngrid, grid, dv, part, out, s);
    case 5: return launch_corr<5, false>(c, cvec, ngrid, grid, dv, part, out, s);
  }
  return -1;
}

int particle_densities(const Ctx& c, const zd* cvec, const double* masses, int ngrid, const double* grid,
                       double* part, double* out, hipStream_t s) {
  // Glob_bvc (src/workproc.f90:612-621): b_i = -m_(2..n+1) / Mtot for every
  // particle i, plus 1 in component i-1 for i >= 2
  const int n = c.n, npart = n + 1;
  DensVec dv{};
  double mtot = 0.0;
  for (int i = 0; i < npart; ++i) mtot += masses[i];
  for (int i = 0; i < npart; ++i)
    for (int p = 0; p < n; ++p) dv.b[i][p] = -masses[p + 1] / mtot + ((i >= 1 && p == i - 1) ? 1.0 : 0.0);
  switch (n) {
    case 1: return launch_corr<1, true>(c, cvec, ngrid, grid, dv, part, out, s);
    case 2: return launch_corr<2, true>(c, cvec, ngrid, grid, dv, part, out, s);
    case 3: return launch_corr<3, true>(c, cvec, ngrid, grid, dv, part, out, s);
    case 4: return launch_corr<4, true>(c, cvec, ngrid, grid, dv, part, out, s);
    case 5: return launch_corr<5, true>(c, cvec, ngrid, grid, dv, part, out, s);
  }
  return -1;
}

// ------------------------------------------ mass-velocity and Darwin
// The relativistic pair values of MatrixElementsForExpcVals without the
// Drachman forms (src/matelem.f90:1479-1667, 1866-1884), contracted with
// the driver's bound as the other expectation values.  For particle 1
// the direction vector is the column sums, for particle i + 1 column i:
// u = Ck^* j, v = tCl j, a = j^T Ck^* j, b = j^T tCl j, and with
// Wc1 = u u^T, Wc2 = v v^T the reference's traces become
//   tr4 = u^T C^-1 u, tr5 = v^T C^-1 v, cnumb = (u^T C^-1 v)^2,
//   caux = a tr5 + b tr4,
//   MV_i = 36 S (tr4 tr5 + 2/3 cnumb - caux + a b) / m_i^3,  MV = -sum MV_i / 8;
// Darwin = sum_q dq_q delta_q with the host's coefficients dq (charges,
// masses and -pi/2).
struct RelConst {
  double inv_m3[6];   // 1 / m^3 of particles 1..n+1
  double dq[15];      // Darwin coefficient of each pair's delta
};

template <int N>
__global__ void __launch_bounds__(TILE* TILE) k_relat(const SysConst sc, int K, const double* __restrict__ bra,
                                                    const double* __restrict__ ket,
                                                    const double* __restrict__ diagS, const zd* __restrict__ cvec,
                                                    const RelConst rc, double* __restrict__ part) {
  constexpr int NO = 2;
  extern __shared__ double lds[];
  int tr, tc;
  tri_tile(blockIdx.x, 0, tr, tc);
  const int T = sc.T;
  const int kb = tr * TILE, lb = tc * TILE;
  double* sbra = lds;
  double* sket = lds + TILE * BraRec<N>::stride;
  const int tid = threadIdx.y * TILE + threadIdx.x;
  for (int i = tid; i < TILE * BraRec<N>::stride; i += TILE * TILE) {
    const int f = kb + i / BraRec<N>::stride;
    sbra[i] = (f < K) ? bra[(size_t)kb * BraRec<N>::stride + i] : 0.0;
  }
  const int kstride = T * KetRec<N>::stride;
  for (int i = tid; i < TILE * kstride; i += TILE * TILE) {
    const int f = lb + i / kstride;
    sket[i] = (f < K) ? ket[(size_t)lb * kstride + i] : 0.0;
  }
  __syncthreads();
  double sum[NO], err[NO];
#pragma unroll
  for (int a = 0; a < NO; ++a) sum[a] = err[a] = 0.0;
  const int k = kb + threadIdx.x, l = lb + threadIdx.y;
  if (k < K && l <= k) {
    const double* br = sbra + threadIdx.x * BraRec<N>::stride;
    const double* kr0 = sket + threadIdx.y * kstride;
    const zd* ck = reinterpret_cast<const zd*>(br + BraRec<N>::ck);
    const double adet = br[BraRec<N>::adet] * bra[(size_t)l * BraRec<N>::stride + BraRec<N>::adet];
    const double f = (k == l) ? 1.0 / diagS[k] : 2.0 / sqrt(diagS[k] * diagS[l]);
    const zd ci = cvec[k], cj = cvec[l];
    const double wr = ci.x * cj.x + ci.y * cj.y, wi = ci.x * cj.y - ci.y * cj.x;   // conj(c_k) c_l
    constexpr double inv_pi32 = 0.179587122125166564899490489694101733;    // pi^(-3/2)
    for (int t = 0; t < T; ++t) {
      const double* kr = kr0 + (size_t)t * KetRec<N>::stride;
      const zd* ct = reinterpret_cast<const zd*>(kr + KetRec<N>::ct);
      PairCore<N> pc;
      pair_energy<N>(sc, ck, ct, reinterpret_cast<const zd*>(kr + KetRec<N>::nl), adet, pc);
      zd mv = mkz(0.0, 0.0);
#pragma unroll
      for (int part_i = 0; part_i <= N; ++part_i) {
        zd u[N], v[N], a, b;
        if (part_i == 0) {
#pragma unroll
          for (int p = 0; p < N; ++p) {
            u[p] = v[p] = mkz(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < N; ++q) {
              u[p] = u[p] + ck[p >= q ? sidx(p, q) : sidx(q, p)];
              v[p] = v[p] + ct[p >= q ? sidx(p, q) : sidx(q, p)];
            }
          }
          a = b = mkz(0.0, 0.0);
#pragma unroll
          for (int p = 0; p < N; ++p) {
            a = a + u[p];
            b = b + v[p];
          }
        } else {
          const int c0 = part_i - 1;
#pragma unroll
          for (int p = 0; p < N; ++p) {
            u[p] = ck[p >= c0 ? sidx(p, c0) : sidx(c0, p)];
            v[p] = ct[p >= c0 ? sidx(p, c0) : sidx(c0, p)];
          }
          a = u[c0];
          b = v[c0];
        }
        // C^-1 u and C^-1 v, then the three quadratic forms
        zd tr4 = mkz(0.0, 0.0), tr5 = mkz(0.0, 0.0), x = mkz(0.0, 0.0);
#pragma unroll
        for (int p = 0; p < N; ++p) {
          zd cu = mkz(0.0, 0.0), cv = mkz(0.0, 0.0);
#pragma unroll
          for (int q = 0; q < N; ++q) {
            const zd cpq = pc.Ci[p >= q ? sidx(p, q) : sidx(q, p)];
            cu = zfma(cpq, u[q], cu);
            cv = zfma(cpq, v[q], cv);
          }
          tr4 = zfma(u[p], cu, tr4);
          tr5 = zfma(v[p], cv, tr5);
          x = zfma(u[p], cv, x);
        }
        const zd br_ = tr4 * tr5 + (2.0 / 3.0) * (x * x) - (a * tr5 + b * tr4) + a * b;
        mv = zfma(mkz(36.0 * rc.inv_m3[part_i], 0.0), pc.S * br_, mv);
      }
      zd O[NO];
      O[0] = -0.125 * mv;
      zd dw = mkz(0.0, 0.0);
      int q = 0;
#pragma unroll
      for (int a = 0; a < N; ++a)
#pragma unroll
        for (int b = a; b < N; ++b, ++q) {
          const zd rs = pc.rs[(a == b) ? sidx(a, a) : sidx(b, a)];
          dw = zfma(mkz(rc.dq[q], 0.0), inv_pi32 * ((pc.S * rs) * (rs * rs)), dw);
        }
      O[1] = dw;
      const double fc = f * sc.coef[t];
#pragma unroll
      for (int a = 0; a < NO; ++a) {
        const double ro = O[a].x, io = O[a].y;
        const double beta = wr * ro - wi * io;
        const double mu = fabs(ci.x * (cj.x * ro - cj.y * io)) + fabs(cj.x * (ci.x * ro + ci.y * io)) +
                          fabs(ci.y * (cj.y * ro + cj.x * io)) + fabs(cj.y * (ci.y * ro - ci.x * io));
        sum[a] += fc * beta;
        err[a] += fabs(fc * (mu + 2.0 * fabs(beta)));
      }
    }
  }
  double* red = lds;
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int a = 0; a < NO; ++a) {
    double v = sum[a], e = err[a];
    for (int o = 32; o > 0; o >>= 1) {
      v += __shfl_xor(v, o);
      e += __shfl_xor(e, o);
    }
    if (lane == 0) {
      red[wave * 2 * NO + a] = v;
      red[wave * 2 * NO + NO + a] = e;
    }
  }
  __syncthreads();
  for (int a = tid; a < 2 * NO; a += TILE * TILE)
    part[(size_t)blockIdx.x * 2 * NO + a] = ((red[a] + red[2 * NO + a]) + red[4 * NO + a]) + red[6 * NO + a];
}

template <int N>
static int launch_relat(const Ctx& c, const zd* cvec, const RelConst& rc, double* part, double* out, hipStream_t s) {
  const int K = c.K;
  const int nt = (K + TILE - 1) / TILE;
  const long long ntiles = (long long)nt * (nt + 1) / 2;
  size_t lds = sizeof(double) * TILE * (BraRec<N>::stride + c.T * KetRec<N>::stride);
  lds = std::max(lds, sizeof(double) * 16);
  hipLaunchKernelGGL(k_relat<N>, dim3((unsigned)ntiles), dim3(TILE, TILE), lds, s, c.sc, K, c.bra, c.ket, c.diagS,
                     cvec, rc, part);
  hipLaunchKernelGGL(k_expect_finish, dim3(1), dim3(256), 0, s, 4, ntiles, part, out);
  return (int)hipGetLastError();
}

int relativistic_values(const Ctx& c, const zd* cvec, const double* masses, double* part, double* out,
                        hipStream_t s) {
  // Darwin coefficients (src/matelem.f90:1866-1884): -pi/2 (1/m_0^2 + 1/m_a^2)
  // q0 q_a for the pair (a, a), -pi/2 (1/m_a^2 + 1/m_b^2) q_a q_b for (a, b)
  const int n = c.n;
  RelConst rc{};
  for (int i = 0; i <= n; ++i) rc.inv_m3[i] = 1.0 / (masses[i] * masses[i] * masses[i]);
  const double hp = -1.57079632679489661923132169163975144;
  int q = 0;
  for (int a = 0; a < n; ++a)
    for (int b = a; b < n; ++b, ++q) {
      const double ia = 1.0 / (masses[a + 1] * masses[a + 1]);
      rc.dq[q] = (a == b) ? hp * (1.0 / (masses[0] * masses[0]) + ia) * c.sc.q0 * c.sc.q[a]
                          : hp * (ia + 1.0 / (masses[b + 1] * masses[b + 1])) * c.sc.q[a] * c.sc.q[b];
    }
  switch (n) {
    case 1: return launch_relat<1>(c, cvec, rc, part, out, s);
    case 2: return launch_relat<2>(c, cvec, rc, part, out, s);
    case 3: return launch_relat<3>(c, cvec, rc, part, out, s);
    case 4: return launch_relat<4>(c, cvec, rc, part, out, s);
    case 5: return launch_relat<5>(c, cvec, rc, part, out, s);
  }
  return -1;
}

// ------------------------------- Drachmanized forms and orbit-orbit term
// The remaining pair values of MatrixElementsForExpcVals: the Drachmanized
// delta(r_pq) of every pair (a, b >= a) (src/matelem.f90:1353-1392), the
// orbit-orbit term (:1394-1477), the Drachmanized mass-velocity (:1669-1865)
// and Darwin (:1890-1933) terms, with the helper integrals
// <r'Xr / r_ij> (ME_rXr_over_rij, :753-818), <(r'Xr)(r'Yr) / r_ij>
// (ME_rXr_rYr_over_rij, :821-945), ME_dXd_dYd, ME_dXd, ME_dXd_over_rij
// (:1998-2044), on full n x n matrices in registers (indices known at
// compile time after unrolling).
//
// <1/(r_ij r_pq)> (:1305-1343) is evaluated per entry instead of as a table,
// with the value the reference's table holds: 2 S / TrCJ_ij for the same
// pair; else, with t = tr[C^-1 J_ij C^-1 J_pq] as the reference's loops
// (:1275-1302) leave it, (4 S / pi) asin(sqrt t / sqrt(TrCJ_ij TrCJ_pq))
// / sqrt(t) (asin z = -i log(i z + sqrt(1 - z^2))), or (4 S / pi) /
// sqrt(TrCJ_ij TrCJ_pq) when t is zero.  Those loops leave t unwritten --
// zero in the reference's static storage -- for a diagonal pair (x, x)
// against an off-diagonal pair (y, z) with x < y; the other entries are
// squares of the entry's own combination of C^-1 in the order of the loop
// iteration that wrote it last (tc_entry).
struct DrachConst {
  double mass[6];   // particle masses, particle 1 first (Glob_Mass)
  double E;         // current energy (Glob_CurrEnergy)
};

__device__ inline zd zdiv(zd a, zd b) { return a * zrcp(b); }
__device__ inline zd zlog(zd z) { return mkz(log(hypot(z.x, z.y)), atan2(z.y, z.x)); }

template <int N>
__device__ __attribute__((always_inline)) inline zd dr_rxr_over_rij(const zd (&X)[N][N], int i, int j,
                                                                    const zd (&Ci)[N][N], zd rm, zd trcj) {
  zd trcx = mkz(0.0, 0.0);
#pragma unroll
  for (int m = 0; m < N; ++m)
#pragma unroll
    for (int p = 0; p < N; ++p) trcx = zfma(Ci[m][p], X[p][m], trcx);
  zd t = mkz(0.0, 0.0);
#pragma unroll
  for (int m = 0; m < N; ++m) {
    zd cxi = mkz(0.0, 0.0), cxj = mkz(0.0, 0.0);
#pragma unroll
    for (int p = 0; p < N; ++p) {
      cxi = zfma(Ci[i][p], X[p][m], cxi);
      if (j != i) cxj = zfma(Ci[j][p], X[p][m], cxj);
    }
    if (i == j) t = zfma(cxi, Ci[m][i], t);
    else t = zfma(cxi - cxj, Ci[m][i] - Ci[m][j], t);
  }
  return (0.5 * rm) * (3.0 * trcx - zdiv(t, trcj));
}

template <int N>
__device__ __attribute__((always_inline)) inline zd dr_rxr_ryr_over_rij(const zd (&X)[N][N], const zd (&Y)[N][N],
                                                                        int i, int j, const zd (&Ci)[N][N], zd rm,
                                                                        zd trcj) {
  zd CX[N][N], CY[N][N];
#pragma unroll
  for (int p = 0; p < N; ++p)
#pragma unroll
    for (int q = 0; q < N; ++q) {
      zd ax = mkz(0.0, 0.0), ay = mkz(0.0, 0.0);
#pragma unroll
      for (int m = 0; m < N; ++m) {
        const zd xs = (m == q) ? X[m][m] : 0.5 * (X[m][q] + X[q][m]);
        const zd ys = (m == q) ? Y[m][m] : 0.5 * (Y[m][q] + Y[q][m]);
        ax = zfma(Ci[p][m], xs, ax);
        ay = zfma(Ci[p][m], ys, ay);
      }
      CX[p][q] = ax;
      CY[p][q] = ay;
    }
  zd trcx = mkz(0.0, 0.0), trcy = mkz(0.0, 0.0), trcxcy = mkz(0.0, 0.0);
#pragma unroll
  for (int m = 0; m < N; ++m) {
    trcy = trcy + CY[m][m];
    trcx = trcx + CX[m][m];
  }
#pragma unroll
  for (int m = 0; m < N; ++m)
#pragma unroll
    for (int p = 0; p < N; ++p) trcxcy = zfma(CX[m][p], CY[p][m], trcxcy);
  zd tx = mkz(0.0, 0.0), ty = mkz(0.0, 0.0), txy = mkz(0.0, 0.0);
#pragma unroll
  for (int m = 0; m < N; ++m) {
    zd xyi = mkz(0.0, 0.0), xyj = mkz(0.0, 0.0);
#pragma unroll
    for (int p = 0; p < N; ++p) {
      xyi = zfma(CX[i][p], CY[p][m], xyi);
      if (j != i) xyj = zfma(CX[j][p], CY[p][m], xyj);
    }
    if (i == j) {
      tx = zfma(CX[i][m], Ci[m][i], tx);
      ty = zfma(CY[i][m], Ci[m][i], ty);
      txy = zfma(xyi, Ci[m][i], txy);
    } else {
      const zd d = Ci[m][i] - Ci[m][j];
      tx = zfma(CX[i][m] - CX[j][m], d, tx);
      ty = zfma(CY[i][m] - CY[j][m], d, ty);
      txy = zfma(xyi - xyj, d, txy);
    }
  }
  const zd r = zrcp(trcj);
  return (0.25 * rm) * (9.0 * (trcy * trcx) + 6.0 * trcxcy - (3.0 * (trcy * tx + trcx * ty) + 4.0 * txy) * r +
                        3.0 * (tx * ty) * (r * r));
}

template <int N>
__device__ __attribute__((always_inline)) inline void dr_matmul(const zd (&A)[N][N], const zd (&B)[N][N],
                                                                zd (&C)[N][N]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      zd s = mkz(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < N; ++k) s = zfma(A[i][k], B[k][j], s);
      C[i][j] = s;
    }
}
template <int N> __device__ __attribute__((always_inline)) inline zd dr_trace(const zd (&A)[N][N]) {
  zd s = mkz(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < N; ++i) s = s + A[i][i];
  return s;
}
template <int N>
__device__ __attribute__((always_inline)) inline zd dr_four_traces(const zd (&A3)[N][N], const zd (&A4)[N][N],
                                                                   zd t1, zd t2, zd t3, zd t4, zd S) {
  zd t5 = mkz(0.0, 0.0);
#pragma unroll
  for (int p = 0; p < N; ++p)
#pragma unroll
    for (int j = 0; j < N; ++j) t5 = zfma(A3[p][j], A4[j][p], t5);
  return S * (36.0 * (t3 * t4) + 24.0 * t5 - 36.0 * (t1 * t4) - 36.0 * (t2 * t3) + 36.0 * (t1 * t2)) * 0.125;
}

// ME_dXd_over_rij (X real here: a mass-weight matrix)
template <int N>
__device__ __attribute__((always_inline)) inline zd dr_dxd_over_rij(const double (&X)[N][N], int i, int j,
                                                                    const zd (&Ci)[N][N], const zd (&Cl)[N][N],
                                                                    zd rm, zd trcj) {
  zd XC[N][N], M[N][N];
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = 0; b < N; ++b) {
      zd s = mkz(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < N; ++k) s = s + X[a][k] * Cl[k][b];
      XC[a][b] = s;
    }
  dr_matmul<N>(Cl, XC, M);
  zd tr1 = mkz(0.0, 0.0);
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int k = 0; k < N; ++k) tr1 = tr1 + Cl[a][k] * X[k][a];
  return 4.0 * dr_rxr_over_rij<N>(M, i, j, Ci, rm, trcj) - 6.0 * (tr1 * rm);
}

// ME_dXd_dYd for real X, Y
template <int N>
__device__ __attribute__((always_inline)) inline zd dr_dxd_dyd(const double (&X)[N][N], const double (&Y)[N][N],
                                                               const zd (&Ci)[N][N], const zd (&Ck)[N][N],
                                                               const zd (&Cl)[N][N], zd S) {
  zd T1[N][N], T2[N][N], M[N][N], MM[N][N];
  zd tr1 = mkz(0.0, 0.0), tr2 = mkz(0.0, 0.0);
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int k = 0; k < N; ++k) {
      tr1 = tr1 + X[a][k] * Ck[k][a];
      tr2 = tr2 + Cl[a][k] * Y[k][a];
    }
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = 0; b < N; ++b) {
      zd s = mkz(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < N; ++k) s = s + X[a][k] * Ck[k][b];
      T1[a][b] = s;
    }
  dr_matmul<N>(Ck, T1, T2);
  dr_matmul<N>(Ci, T2, M);
  const zd tr3 = dr_trace<N>(M);
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = 0; b < N; ++b) {
      zd s = mkz(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < N; ++k) s = s + Y[a][k] * Cl[k][b];
      T1[a][b] = s;
    }
  dr_matmul<N>(Cl, T1, T2);
  dr_matmul<N>(Ci, T2, MM);
  const zd tr4 = dr_trace<N>(MM);
  return 8.0 * dr_four_traces<N>(M, MM, tr1, tr2, tr3, tr4, S);
}

// tr[C^-1 J_ab C^-1 J_cd] as the reference's table holds it (see above);
// pairs (a <= b), (c <= d).
template <int N>
__device__ __attribute__((always_inline)) inline zd tc_entry(int a, int b, int c, int d, const zd (&Ci)[N][N]) {
  if (a == b && c == d) {
    const zd v = (a == c) ? Ci[a][a] : Ci[a > c ? a : c][a > c ? c : a];
    return v * v;
  }
  if (a == b || c == d) {   // diagonal (x, x) against (y, z), y < z
    const int x = (a == b) ? a : c, y = (a == b) ? c : a, z = (a == b) ? d : b;
    if (x < y) return mkz(0.0, 0.0);   // never written
    const zd t = Ci[x][y] - Ci[x][z];
    return t * t;
  }
  // two off-diagonal pairs: the iteration (i, j, p, q) with p >= i, the
  // last one in loop order when a == c
  int i, j, p, q;
  if (a < c) { i = a; j = b; p = c; q = d; }
  else if (c < a) { i = c; j = d; p = a; q = b; }
  else { i = a; j = b > d ? b : d; p = a; q = b > d ? d : b; }
  const zd u = Ci[q][i] - Ci[p][i] - Ci[q][j] + Ci[p][j];
  return u * u;
}

template <int N>
__device__ __attribute__((always_inline)) inline zd rr_entry(int a, int b, int c, int d, const zd (&Ci)[N][N],
                                                             const zd (&trcj)[N][N], zd S) {
  constexpr double inv_pi = 0.318309886183790671537767526745028724;
  if (a == c && b == d) return zdiv(2.0 * S, trcj[a][b]);
  const zd ca = zsqrt(tc_entry<N>(a, b, c, d, Ci));
  const zd c4 = (4.0 * inv_pi) * S;
  const zd prod = trcj[a][b] * trcj[c][d];
  if (ca.x != 0.0 || ca.y != 0.0) {
    const zd ca1 = zdiv(ca, zsqrt(prod));
    const zd w = mkz(-ca1.y, ca1.x) + zsqrt(mkz(1.0, 0.0) - ca1 * ca1);   // i ca1 + sqrt(1 - ca1^2)
    const zd lw = zlog(w);
    return zdiv(c4 * mkz(lw.y, -lw.x), ca);   // -i log(w) / ca
  }
  return zdiv(c4, zsqrt(prod));
}

// The n(n+1)/2 + 3 values of one pair and term.
template <int N>
__device__ __attribute__((always_inline)) inline void drach_values(const SysConst& sc, const zd* ck, const zd* ct,
                                                                   const PairCore<N>& pc, const DrachConst& dc,
                                                                   zd* out) {
  constexpr double pi = 3.14159265358979323846264338327950288, sqrtpi = 1.77245385090551602729816748334115;
  zd Ci[N][N], Ck[N][N], Cl[N][N];
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = 0; b < N; ++b) {
      const int e = a >= b ? sidx(a, b) : sidx(b, a);
      Ci[a][b] = pc.Ci[e];
      Ck[a][b] = ck[e];
      Cl[a][b] = ct[e];
    }
  const zd S = pc.S;
  zd trcj[N][N], rm[N][N];
  const zd cn = (2.0 / sqrtpi) * S;
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int b = 0; b < N; ++b) {
      trcj[a][b] = (a == b) ? Ci[a][a] : Ci[a][a] + Ci[b][b] - Ci[a > b ? a : b][a > b ? b : a] -
                                             Ci[a > b ? a : b][a > b ? b : a];
      rm[a][b] = zdiv(cn, zsqrt(trcj[a][b]));
    }
  const double* mass = dc.mass;
  // Drachmanized delta(r_pq).  The loops over particle indices stay rolled
  // (the matrix work inside them is unrolled): unrolled, they multiply the
  // code of these helpers by the number of index pairs.
  {
    zd W2[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        zd t = mkz(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < N; ++k) {
          zd mc = mkz(0.0, 0.0);
#pragma unroll
          for (int r = 0; r < N; ++r) mc = mc + sc.M[k][r] * Cl[r][j];
          t = zfma(Ck[i][k], mc, t);
        }
        W2[i][j] = t;
      }
    int o = 0;
#pragma unroll 1
    for (int p = 0; p < N; ++p)
#pragma unroll 1
      for (int q = p; q < N; ++q, ++o) {
        zd cts = mkz(0.0, 0.0);
#pragma unroll 1
        for (int i = 0; i < N; ++i) {
          const bool first_ii = p < i || (p == i && q <= i);
          cts = cts + (sc.q0 * sc.q[i]) * (first_ii ? rr_entry<N>(p, q, i, i, Ci, trcj, S)
                                                    : rr_entry<N>(i, i, p, q, Ci, trcj, S));
#pragma unroll 1
          for (int j = i + 1; j < N; ++j) cts = cts + (sc.q[i] * sc.q[j]) * rr_entry<N>(p, q, i, j, Ci, trcj, S);
        }
        const double ca = (p == q) ? 2.0 * pi * sc.M[p][p] : 2.0 * pi * (sc.M[p][p] + sc.M[q][q] - 2.0 * sc.M[p][q]);
        const zd cnm = dr_rxr_over_rij<N>(W2, p, q, Ci, rm[p][q], trcj[p][q]);
        out[o] = (1.0 / ca) * (dc.E * rm[p][q] - cts - 4.0 * cnm);
      }
  }
  constexpr int m = N * (N + 1) / 2;
  // orbit-orbit
  {
    zd OO = mkz(0.0, 0.0);
#pragma unroll 1
    for (int i = 0; i < N; ++i)
#pragma unroll 1
      for (int j = 0; j < N; ++j) {
        const zd tr1 = Cl[j][i], tr3 = 3.0 * Cl[j][j];
        zd W1[N][N], W2[N][N], W3[N][N];
#pragma unroll
        for (int p = 0; p < N; ++p)
#pragma unroll
          for (int q = 0; q < N; ++q) {
            W1[q][p] = Cl[q][i] * Cl[p][j] + Cl[q][j] * Cl[p][i];
            W2[p][q] = (Ck[p][j] + Cl[p][j]) * Cl[q][j];
            W3[p][q] = mkz(0.0, 0.0);
          }
#pragma unroll
        for (int p = 0; p < N; ++p) W1[p][j] = W1[p][j] + (Ck[p][j] + Cl[p][j]) * Cl[j][i];
#pragma unroll
        for (int p = 0; p < N; ++p) W1[j][p] = W1[j][p] + Cl[i][j] * Cl[p][j] + tr3 * Cl[p][i];
#pragma unroll
        for (int p = 0; p < N; ++p) W3[j][p] = Cl[p][i];
        const zd cnm = dr_rxr_over_rij<N>(W1, j, j, Ci, rm[j][j], trcj[j][j]);
        const zd ctt = dr_rxr_ryr_over_rij<N>(W2, W3, j, j, Ci, rm[j][j], trcj[j][j]);
        const zd ca = -6.0 * ((tr1 + tr1) * rm[j][j]) + 4.0 * cnm - 8.0 * ctt;
        OO = OO - (sc.q[j] / mass[j + 1]) * ca;
      }
    OO = (sc.q0 / mass[0]) * OO;
#pragma unroll 1
    for (int i = 0; i < N; ++i)
#pragma unroll 1
      for (int j = i + 1; j < N; ++j) {
        const zd tr1 = Cl[j][i], tr3 = 3.0 * Cl[j][j];
        zd W1[N][N], W2[N][N], W3[N][N];
#pragma unroll
        for (int p = 0; p < N; ++p)
#pragma unroll
          for (int q = 0; q < N; ++q) {
            W1[q][p] = Cl[q][i] * Cl[p][j] + Cl[q][j] * Cl[p][i];
            W2[p][q] = (Ck[p][j] + Cl[p][j]) * Cl[q][j];
            W3[p][q] = mkz(0.0, 0.0);
          }
#pragma unroll
        for (int p = 0; p < N; ++p) {
          const zd t = (Ck[p][j] + Cl[p][j]) * Cl[j][i];
          W1[p][j] = W1[p][j] + t;
          W1[p][i] = W1[p][i] - t;
        }
#pragma unroll
        for (int p = 0; p < N; ++p) {
          const zd t = Cl[i][j] * Cl[p][j] + tr3 * Cl[p][i];
          W1[j][p] = W1[j][p] + t;
          W1[i][p] = W1[i][p] - t;
        }
#pragma unroll
        for (int p = 0; p < N; ++p) {
          W3[j][p] = Cl[p][i];
          W3[i][p] = -Cl[p][i];
        }
        const zd cnm = dr_rxr_over_rij<N>(W1, i, j, Ci, rm[i][j], trcj[i][j]);
        const zd ctt = dr_rxr_ryr_over_rij<N>(W2, W3, i, j, Ci, rm[i][j], trcj[i][j]);
        const zd ca = -6.0 * ((tr1 + tr1) * rm[i][j]) + 4.0 * cnm - 8.0 * ctt;
        OO = OO + (sc.q[i] * sc.q[j] / (mass[i + 1] * mass[j + 1])) * ca;
      }
    out[m] = 0.5 * OO;
  }
  // Drachmanized mass-velocity
  {
    double Wm[N][N];
    const double c0 = 1.0 / (mass[0] * mass[0]);
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
      for (int q = 0; q < N; ++q) Wm[p][q] = (p == q) ? c0 + 1.0 / (mass[p + 1] * mass[p + 1]) : c0;
    zd W2[N][N], W3[N][N];
    zd tr1 = mkz(0.0, 0.0);
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
      for (int q = 0; q < N; ++q) {
        zd t = mkz(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < N; ++j) t = t + Cl[j][q] * Wm[j][p];
        W2[q][p] = t;
      }
#pragma unroll
    for (int p = 0; p < N; ++p) tr1 = tr1 + W2[p][p];
    dr_matmul<N>(W2, Cl, W3);
    zd tr2 = mkz(0.0, 0.0);
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
      for (int j = 0; j < N; ++j) tr2 = zfma(Ci[j][p], W3[j][p], tr2);
    zd dmv = (6.0 * 0.25 * dc.E) * (S * (tr2 - tr1));
    zd cnm = mkz(0.0, 0.0);
#pragma unroll 1
    for (int i = 0; i < N; ++i)
      cnm = cnm + (sc.q0 * sc.q[i]) *
                      (4.0 * dr_rxr_over_rij<N>(W3, i, i, Ci, rm[i][i], trcj[i][i]) - 6.0 * (tr1 * rm[i][i]));
#pragma unroll 1
    for (int i = 0; i < N; ++i)
#pragma unroll 1
      for (int j = i + 1; j < N; ++j)
        cnm = cnm + (sc.q[i] * sc.q[j]) *
                        (4.0 * dr_rxr_over_rij<N>(W3, i, j, Ci, rm[i][j], trcj[i][j]) - 6.0 * (tr1 * rm[i][j]));
    dmv = dmv - 0.25 * cnm;
    // third term: J Ck* and the first particle's pair weights
    {
      zd W4[N][N], W5[N][N], A3[N][N];
      tr1 = mkz(0.0, 0.0);
      zd colsum[N];
#pragma unroll
      for (int p = 0; p < N; ++p) {
        zd c = mkz(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < N; ++q) c = c + Ck[q][p];
        colsum[p] = c;
        tr1 = tr1 + c;
      }
      zd tr2b = mkz(0.0, 0.0);
#pragma unroll
      for (int p = 0; p < N; ++p) {
        const double cnp = (mass[0] + mass[p + 1]) / (mass[0] * mass[0] * mass[p + 1] * mass[p + 1]);
#pragma unroll
        for (int q = 0; q < N; ++q) W2[q][p] = cnp * Cl[q][p];
        tr2b = tr2b + W2[p][p];
      }
      dr_matmul<N>(Ci, Ck, W5);
      zd tr3 = mkz(0.0, 0.0);
#pragma unroll
      for (int q = 0; q < N; ++q)
#pragma unroll
        for (int p = 0; p < N; ++p) {
          zd t = mkz(0.0, 0.0);
#pragma unroll
          for (int j = 0; j < N; ++j) t = t + W5[q][j];
          A3[q][p] = t * colsum[p];   // (C^-1 Ck* J Ck*)(q, p) = rowsum_q(C^-1 Ck*) colsum_p(Ck*)
        }
#pragma unroll
      for (int p = 0; p < N; ++p) tr3 = tr3 + A3[p][p];
      dr_matmul<N>(Ci, W2, W5);
      dr_matmul<N>(W5, Cl, W4);
      const zd tr4 = dr_trace<N>(W4);
      dmv = dmv + dr_four_traces<N>(A3, W4, tr1, tr2b, tr3, tr4, S);
    }
    // fourth term: pairs of pseudo-particles
#pragma unroll 1
    for (int k = 0; k < N - 1; ++k) {
      zd W4[N][N], W5[N][N], A3[N][N];
      const zd t1 = Ck[k][k];
      zd t2 = mkz(0.0, 0.0);
#pragma unroll
      for (int p = 0; p < N; ++p) {
        const double cnp = (p > k) ? (mass[k + 1] + mass[p + 1]) /
                                         (mass[k + 1] * mass[k + 1] * mass[p + 1] * mass[p + 1])
                                   : 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) W2[q][p] = cnp * Cl[q][p];
        if (p > k) t2 = t2 + W2[p][p];
      }
#pragma unroll
      for (int p = 0; p < N; ++p)
#pragma unroll
        for (int q = 0; q < N; ++q) W5[q][p] = Ck[q][k] * Ck[p][k];
      dr_matmul<N>(Ci, W5, A3);
      const zd t3 = dr_trace<N>(A3);
      dr_matmul<N>(Ci, W2, W5);
      dr_matmul<N>(W5, Cl, W4);
      const zd t4 = dr_trace<N>(W4);
      dmv = dmv + dr_four_traces<N>(A3, W4, t1, t2, t3, t4, S);
    }
    out[m + 1] = dmv;
  }
  // Drachmanized Darwin
  {
    double W1[N][N], W2[N][N];
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int b = 0; b < N; ++b)
        W1[a][b] = 1.0 / (mass[0] * mass[0]) + ((a == b) ? 1.0 / (mass[a + 1] * mass[a + 1]) : 0.0);
    // ME_dXd
    zd XC[N][N], T2[N][N], M[N][N];
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int b = 0; b < N; ++b) {
        zd s = mkz(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < N; ++k) s = s + W1[a][k] * Cl[k][b];
        XC[a][b] = s;
      }
    dr_matmul<N>(Cl, XC, T2);
    dr_matmul<N>(Ci, T2, M);
    zd tr1 = mkz(0.0, 0.0);
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int k = 0; k < N; ++k) tr1 = tr1 + Cl[a][k] * W1[k][a];
    const zd dxd = 6.0 * ((dr_trace<N>(M) - tr1) * S);
    zd dd = (0.25 * dc.E) * dxd;
#pragma unroll 1
    for (int i = 0; i < N; ++i)
      dd = dd - (0.25 * sc.q0 * sc.q[i]) * dr_dxd_over_rij<N>(W1, i, i, Ci, Cl, rm[i][i], trcj[i][i]);
#pragma unroll 1
    for (int i = 0; i < N; ++i)
#pragma unroll 1
      for (int j = i + 1; j < N; ++j)
        dd = dd - (0.25 * sc.q[i] * sc.q[j]) * dr_dxd_over_rij<N>(W1, i, j, Ci, Cl, rm[i][j], trcj[i][j]);
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int b = 0; b < N; ++b) {
        W1[a][b] = 1.0;
        W2[a][b] = (a == b) ? (mass[0] + mass[a + 1]) / (mass[0] * mass[0] * mass[a + 1] * mass[a + 1]) : 0.0;
      }
    dd = dd + 0.125 * dr_dxd_dyd<N>(W1, W2, Ci, Ck, Cl, S);
#pragma unroll 1
    for (int i = 0; i < N; ++i)
#pragma unroll 1
      for (int j = i + 1; j < N; ++j) {
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int b = 0; b < N; ++b) {
            W1[a][b] = (a == i && b == i) ? 1.0 : 0.0;
            W2[a][b] = (a == j && b == j) ? 1.0 : 0.0;
          }
        const double cnm = (mass[i + 1] + mass[j + 1]) / (mass[i + 1] * mass[i + 1] * mass[j + 1] * mass[j + 1]);
        dd = dd + (0.125 * cnm) * dr_dxd_dyd<N>(W1, W2, Ci, Ck, Cl, S);
      }
    out[m + 2] = dd;
  }
}

template <int N>
__global__ void __launch_bounds__(TILE* TILE) k_drach(const SysConst sc, int K, const double* __restrict__ bra,
                                                    const double* __restrict__ ket,
                                                    const double* __restrict__ diagS, const zd* __restrict__ cvec,
                                                    const DrachConst dc, double* __restrict__ part) {
  constexpr int NO = N * (N + 1) / 2 + 3;
  extern __shared__ double lds[];
  int tr, tc;
  tri_tile(blockIdx.x, 0, tr, tc);
  const int T = sc.T;
  const int kb = tr * TILE, lb = tc * TILE;
  double* sbra = lds;
  double* sket = lds + TILE * BraRec<N>::stride;
  const int tid = threadIdx.y * TILE + threadIdx.x;
  for (int i = tid; i < TILE * BraRec<N>::stride; i += TILE * TILE) {
    const int f = kb + i / BraRec<N>::stride;
    sbra[i] = ld_or0(bra, (size_t)kb * BraRec<N>::stride + i, f < K);
  }
  const int kstride = T * KetRec<N>::stride;
  for (int i = tid; i < TILE * kstride; i += TILE * TILE) {
    const int f = lb + i / kstride;
    sket[i] = ld_or0(ket, (size_t)lb * kstride + i, f < K);
  }
  __syncthreads();
  double sum[NO], err[NO];
#pragma unroll
  for (int a = 0; a < NO; ++a) sum[a] = err[a] = 0.0;
  const int k = kb + threadIdx.x, l = lb + threadIdx.y;
  if (k < K && l <= k) {
    const double* br = sbra + threadIdx.x * BraRec<N>::stride;
    const double* kr0 = sket + threadIdx.y * kstride;
    const zd* ck = reinterpret_cast<const zd*>(br + BraRec<N>::ck);
    const double adet = br[BraRec<N>::adet] * bra[(size_t)l * BraRec<N>::stride + BraRec<N>::adet];
    const double f = (k == l) ? 1.0 / diagS[k] : 2.0 / sqrt(diagS[k] * diagS[l]);
    const zd ci = cvec[k], cj = cvec[l];
    const double wr = ci.x * cj.x + ci.y * cj.y, wi = ci.x * cj.y - ci.y * cj.x;   // conj(c_k) c_l
    for (int t = 0; t < T; ++t) {
      const double* kr = kr0 + (size_t)t * KetRec<N>::stride;
      const zd* ct = reinterpret_cast<const zd*>(kr + KetRec<N>::ct);
      PairCore<N> pc;
      pair_energy<N>(sc, ck, ct, reinterpret_cast<const zd*>(kr + KetRec<N>::nl), adet, pc);
      zd O[NO];
      drach_values<N>(sc, ck, ct, pc, dc, O);
      const double fc = f * sc.coef[t];
#pragma unroll
      for (int a = 0; a < NO; ++a) {
        const double ro = O[a].x, io = O[a].y;
        const double beta = wr * ro - wi * io;
        const double mu = fabs(ci.x * (cj.x * ro - cj.y * io)) + fabs(cj.x * (ci.x * ro + ci.y * io)) +
                          fabs(ci.y * (cj.y * ro + cj.x * io)) + fabs(cj.y * (ci.y * ro - ci.x * io));
        sum[a] += fc * beta;
        err[a] += fabs(fc * (mu + 2.0 * fabs(beta)));
      }
    }
  }
  double* red = lds;
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int a = 0; a < NO; ++a) {
    double v = sum[a], e = err[a];
    for (int o = 32; o > 0; o >>= 1) {
      v += __shfl_xor(v, o);
      e += __shfl_xor(e, o);
    }
    if (lane == 0) {
      red[wave * 2 * NO + a] = v;
      red[wave * 2 * NO + NO + a] = e;
    }
  }
  __syncthreads();
  for (int a = tid; a < 2 * NO; a += TILE * TILE)
    part[(size_t)blockIdx.x * 2 * NO + a] = ((red[a] + red[2 * NO + a]) + red[4 * NO + a]) + red[6 * NO + a];
}

template <int N>
static int launch_drach(const Ctx& c, const zd* cvec, const DrachConst& dc, double* part, double* out,
                        hipStream_t s) {
  constexpr int NO = N * (N + 1) / 2 + 3;
  const int K = c.K;
  const int nt = (K + TILE - 1) / TILE;
  const long long ntiles = (long long)nt * (nt + 1) / 2;
  size_t lds = sizeof(double) * TILE * (BraRec<N>::stride + c.T * KetRec<N>::stride);
  lds = std::max(lds, sizeof(double) * 8 * NO);
  hipLaunchKernelGGL(k_drach<N>, dim3((unsigned)ntiles), dim3(TILE, TILE), lds, s, c.sc, K, c.bra, c.ket, c.diagS,
                     cvec, dc, part);
  hipLaunchKernelGGL(k_expect_finish, dim3(1), dim3(256), 0, s, 2 * NO, ntiles, part, out);
  return (int)hipGetLastError();
}

int drachmanized_values(const Ctx& c, const zd* cvec, const double* masses, double E, double* part, double* out,
                        hipStream_t s) {
  DrachConst dc{};
  for (int i = 0; i <= c.n; ++i) dc.mass[i] = masses[i];
  dc.E = E;
  switch (c.n) {
    case 1: return launch_drach<1>(c, cvec, dc, part, out, s);
    case 2: return launch_drach<2>(c, cvec, dc, part, out, s);
    case 3: return launch_drach<3>(c, cvec, dc, part, out, s);
    case 4: return launch_drach<4>(c, cvec, dc, part, out, s);
    case 5: return launch_drach<5>(c, cvec, dc, part, out, s);
  }
  return -1;
}

size_t drachmanized_partial_doubles(int n, int K) {
  const int nt = (K + TILE - 1) / TILE;
  return (size_t)nt * (nt + 1) / 2 * 2 * (n * (n + 1) / 2 + 3);
}

size_t correlation_partial_doubles(int n, int K) {
  const int nt = (K + TILE - 1) / TILE;
  const int m = std::max(n * (n + 1) / 2, n + 1);   // pair functions or particle densities
  return (size_t)nt * (nt + 1) / 2 * CF_CHUNK * m;
}

// ------------------------------------------------------ overlap penalty
// ComputeOverlapPenalty / ComputeOverlapPenaltyAndAddGradient
// (src/workproc.f90:1491-1693).  P sums pc (|s_ji|^2 - t2), pc = b/(1-t2),
// over the pairs j > i with j > nfru whose normalised overlap has
// |s_ji|^2 > t2.  Gradient terms for function a and partner b of a hit
// pair (j, i) = (max(a,b), min(a,b)), f = 1/sqrt(S_aa S_bb),
// D_aa = (dS_aa/dp_a) / S_aa, written as the reference writes them:
//   a = j:  2 pc f Re(conj(s) dS_ji/dp_a) - pc |s|^2   sqrt(S_bb/S_aa) D_aa
//   a = i:  2 pc f Re(conj(s) dS_ji/dp_a) - pc Re(s^2) sqrt(S_bb/S_aa) D_aa
// (the exact derivative of P has |s|^2 and no square-root factor in both;
// the two agree when S_aa = S_bb and s is real).
// One wave per function a >= nfru reads column a of the device S, which
// holds both triangles (s_ji is the element (b, a) for b > a, its
// conjugate for b < a).  Only hit pairs, rare by the purpose of the
// penalty, recompute their derivatives.  Pair values of P go to pen[a -
// nfru] (pairs with j = a); wave sums use a fixed butterfly order.
template <int N>
__global__ void __launch_bounds__(256) k_overlap_penalty(const SysConst sc, int K, int nfru, int ld,
                                                        const double* __restrict__ bra,
                                                        const double* __restrict__ ket,
                                                        const double* __restrict__ diagS,
                                                        const zd* __restrict__ Sm, double pc, double t2,
                                                        double* __restrict__ pen, double* __restrict__ grad) {
  constexpr int m = N * (N + 1) / 2, npt = 2 * m, np4 = 4 * m;
  const int a = nfru + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (a >= K) return;
  const int lane = threadIdx.x & 63;
  const int kstride = sc.T * KetRec<N>::stride;
  const double Sa = diagS[a];
  const zd* col = Sm + (size_t)a * ld;
  double p = 0.0, beta = 0.0, g[npt];
#pragma unroll
  for (int e = 0; e < npt; ++e) g[e] = 0.0;
  for (int b = lane; b < K; b += 64) {
    if (b == a) continue;
    zd s = col[b];
    if (b < a) s = zconj(s);
    const double s2 = s.x * s.x + s.y * s.y;
    if (!(s2 > t2)) continue;
    if (b < a) p += pc * (s2 - t2);
    if (!grad) continue;
    const double Sb = diagS[b];
    const bool ket_side = b > a;
    const int j = ket_side ? b : a, i = ket_side ? a : b;
    zd H, S, Dk[np4], Dl[np4];
    pair_sum_deriv<N>(sc, bra + (size_t)j * BraRec<N>::stride, ket + (size_t)i * kstride,
                      bra + (size_t)i * BraRec<N>::stride, ket_side, H, S, Dk, Dl);
    const double c2 = 2.0 * pc / sqrt(Sa * Sb);
#pragma unroll
    for (int e = 0; e < npt; ++e) {
      const zd d = ket_side ? Dl[2 * m + e] : Dk[2 * m + e];
      g[e] += c2 * (s.x * d.x + s.y * d.y);
    }
    const double w = ket_side ? (s.x * s.x - s.y * s.y) : s2;
    beta += pc * w * sqrt(Sb / Sa);
  }
  for (int o = 32; o > 0; o >>= 1) p += __shfl_xor(p, o);
  if (lane == 0) pen[a - nfru] = p;
  if (!grad) return;
  for (int o = 32; o > 0; o >>= 1) beta += __shfl_xor(beta, o);
  bool any = beta != 0.0;
#pragma unroll
  for (int e = 0; e < npt; ++e) {
    for (int o = 32; o > 0; o >>= 1) g[e] += __shfl_xor(g[e], o);
    any = any || g[e] != 0.0;
  }
  if (lane != 0 || !any) return;
  zd H, S, Dk[np4], Dl[np4];
  pair_sum_deriv<N>(sc, bra + (size_t)a * BraRec<N>::stride, ket + (size_t)a * kstride,
                    bra + (size_t)a * BraRec<N>::stride, false, H, S, Dk, Dl);
  double* ga = grad + (size_t)(a - nfru) * npt;
#pragma unroll
  for (int e = 0; e < npt; ++e) ga[e] += g[e] - beta * (2.0 * Dk[2 * m + e].x / Sa);
}

// Sum of v[0..n) in a fixed order by one workgroup of 256 threads.
__global__ void __launch_bounds__(256) k_sum_fixed(const double* __restrict__ v, int n, double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += v[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// ------------------------------------------------- packed problem batches
// One launch covers the pair tiles of every problem of a batch that has
// the same n: workgroup b belongs to the problem whose tile range
// [tile0, tile0 + tiles) holds b (ranges are ascending in the table).
__device__ inline int find_prob(const ProbDesc* __restrict__ pd, int np, long long b, bool adj) {
  int p = 0;
  while (p + 1 < np && (adj ? pd[p + 1].tile0_adj : pd[p + 1].tile0_off) <= b) ++p;
  return p;
}

template <int N>
__global__ void __launch_bounds__(TILE* TILE) k_offdiag_multi(const ProbDesc* __restrict__ pd, int np) {
  extern __shared__ double lds[];
  const int p = find_prob(pd, np, blockIdx.x, false);
  const ProbDesc& P = pd[p];
  int tr, tc;
  tri_tile((int)(blockIdx.x - P.tile0_off), 0, tr, tc);
  offdiag_tile<N>(P.sc, 0, P.K, P.ld, tr, tc, P.bra, P.ket, P.diagS, P.H, P.S, lds);
}

template <int N>
__global__ void __launch_bounds__(TILE* TILE, 2) k_adjoint_multi(const ProbDesc* __restrict__ pd, int np) {
  extern __shared__ double lds[];
  const int p = find_prob(pd, np, blockIdx.x, true);
  const ProbDesc& P = pd[p];
  int tr, tc;
  tri_tile((int)(blockIdx.x - P.tile0_adj), 0, tr, tc);
  adjoint_tile<N>(P.sc, P.K, 0, P.nt, tr, tc, P.E, P.bra, P.ket, P.diagS, P.cvec, P.part, lds);
}

// thread -> function f0 + t of the batch's concatenated functions
template <int N>
__global__ void __launch_bounds__(256) k_adjoint_finish_multi(const ProbDesc* __restrict__ pd, int np, int nf) {
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= nf) return;
  int p = 0;
  while (p + 1 < np && pd[p + 1].f0 <= g) ++p;
  const ProbDesc& P = pd[p];
  adjoint_finish_one<N>(g - P.f0, 0, P.nt, 0, P.bra, P.part, P.grad);
}

// ------------------------------------------------------------ launchers
// ------------------------------------------------------ screening rows
// Candidate functions for the last basis position (the reference's trial
// loops, src/workproc.f90:1375-1487, 3397-3413): the candidate is the bra of
// row K (it is the reference's row index k >= l), the fixed functions are
// the kets.
// Diagonal pair of candidate j: cdiag[2j] = Re S_jj, cdiag[2j+1] = Re H_jj / S_jj.
template <int N>
__global__ void __launch_bounds__(128) k_cand_diag(const SysConst sc, int C, const double* __restrict__ cbra,
                                                   const double* __restrict__ cket, double* __restrict__ cdiag) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= C) return;
  const double* br = cbra + (size_t)j * BraRec<N>::stride;
  zd H, S;
  pair_sum_energy<N>(sc, br, cket + (size_t)j * sc.T * KetRec<N>::stride, br[BraRec<N>::adet], H, S);
  cdiag[2 * j] = S.x;
  cdiag[2 * j + 1] = H.x / S.x;
}

// Border columns of candidate j against the fixed functions l < Kp, as the
// upper-right column of the bordered matrices (conjugate of row K):
//   Hb(l, j) = conj(H_Kl) f, Sb(l, j) = conj(S_Kl) f, f = 1/sqrt(S_jj diagS_l).
template <int N>
__global__ void __launch_bounds__(128) k_cand_rows(const SysConst sc, int Kp, int ld,
                                                   const double* __restrict__ bra, const double* __restrict__ ket,
                                                   const double* __restrict__ diagS,
                                                   const double* __restrict__ cbra,
                                                   const double* __restrict__ cdiag, zd* __restrict__ Hb,
                                                   zd* __restrict__ Sb) {
  __shared__ double sbra[BraRec<N>::stride];
  const int j = blockIdx.y;
  for (int i = threadIdx.x; i < BraRec<N>::stride; i += blockDim.x)
    sbra[i] = cbra[(size_t)j * BraRec<N>::stride + i];
  __syncthreads();
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= Kp) return;
  zd H, S;
  pair_sum_energy<N>(sc, sbra, ket + (size_t)l * sc.T * KetRec<N>::stride,
                     bra[(size_t)l * BraRec<N>::stride + BraRec<N>::adet], H, S);
  const double f = 1.0 / sqrt(cdiag[2 * j] * diagS[l]);
  Hb[(size_t)j * ld + l] = zconj(H) * f;
  Sb[(size_t)j * ld + l] = zconj(S) * f;
}

// The same rows with CB candidates per workgroup: 128 / CB functions'
// ket records (all T terms) staged in LDS once and shared by the CB
// candidates, whose bra records are staged too, so each function's records
// are read from L2 once per CB candidates instead of once per candidate.
// Thread t: function l0 + t % (128 / CB), candidate j0 + t / (128 / CB).
template <int N, int CB>
__global__ void __launch_bounds__(128) k_cand_rows_cb(const SysConst sc, int Kp, int C, int ld,
                                                      const double* __restrict__ bra, const double* __restrict__ ket,
                                                      const double* __restrict__ diagS,
                                                      const double* __restrict__ cbra,
                                                      const double* __restrict__ cdiag, zd* __restrict__ Hb,
                                                      zd* __restrict__ Sb) {
  constexpr int FB = 128 / CB;
  extern __shared__ double sm[];
  const int kst = sc.T * KetRec<N>::stride;
  double* sket = sm;                        // FB x T ket records
  double* sbra = sm + (size_t)FB * kst;     // CB bra records
  const int l0 = blockIdx.x * FB, j0 = blockIdx.y * CB;
  for (int i = threadIdx.x; i < FB * kst; i += blockDim.x)
    sket[i] = (l0 + i / kst < Kp) ? ket[(size_t)l0 * kst + i] : 0.0;
  for (int i = threadIdx.x; i < CB * BraRec<N>::stride; i += blockDim.x)
    sbra[i] = (j0 + i / BraRec<N>::stride < C) ? cbra[(size_t)j0 * BraRec<N>::stride + i] : 0.0;
  __syncthreads();
  const int fl = threadIdx.x % FB, cj = threadIdx.x / FB;
  const int l = l0 + fl, j = j0 + cj;
  if (l >= Kp || j >= C) return;
  zd H, S;
  pair_sum_energy<N>(sc, sbra + cj * BraRec<N>::stride, sket + (size_t)fl * kst,
                     bra[(size_t)l * BraRec<N>::stride + BraRec<N>::adet], H, S);
  const double f = 1.0 / sqrt(cdiag[2 * j] * diagS[l]);
  Hb[(size_t)j * ld + l] = zconj(H) * f;
  Sb[(size_t)j * ld + l] = zconj(S) * f;
}

template <int N>
static int launch_cand(const Ctx& c, int Kp, int C, const double* cparams, double* cbra, double* cket,
                       double* cdiag, zd* Hb, zd* Sb, int ldb, int cpb, bool rows_only, hipStream_t s) {
  if (cpb != 1 && cand_rows_lds(c.n, c.sc.T, cpb) == 0) return -1;
  if (!rows_only) {
    hipLaunchKernelGGL(k_prep<N>, dim3((C + 127) / 128), dim3(128), 0, s, c.sc, 0, C, cparams, cbra, cket);
    hipLaunchKernelGGL(k_cand_diag<N>, dim3((C + 127) / 128), dim3(128), 0, s, c.sc, C, cbra, cket, cdiag);
  }
  const size_t lds = cand_rows_lds(c.n, c.sc.T, cpb);
  switch (cpb) {
    case 1:
      hipLaunchKernelGGL(k_cand_rows<N>, dim3((Kp + 127) / 128, C), dim3(128), 0, s, c.sc, Kp, ldb, c.bra, c.ket,
                         c.diagS, cbra, cdiag, Hb, Sb);
      break;
    case 2:
      hipLaunchKernelGGL((k_cand_rows_cb<N, 2>), dim3((Kp + 63) / 64, (C + 1) / 2), dim3(128), lds, s, c.sc, Kp, C,
                         ldb, c.bra, c.ket, c.diagS, cbra, cdiag, Hb, Sb);
      break;
    case 4:
      hipLaunchKernelGGL((k_cand_rows_cb<N, 4>), dim3((Kp + 31) / 32, (C + 3) / 4), dim3(128), lds, s, c.sc, Kp, C,
                         ldb, c.bra, c.ket, c.diagS, cbra, cdiag, Hb, Sb);
      break;
    case 8:
      hipLaunchKernelGGL((k_cand_rows_cb<N, 8>), dim3((Kp + 15) / 16, (C + 7) / 8), dim3(128), lds, s, c.sc, Kp, C,
                         ldb, c.bra, c.ket, c.diagS, cbra, cdiag, Hb, Sb);
      break;
    default:
      return -1;
  }
  return (int)hipGetLastError();
}

template <int N>
static int launch_prep(const Ctx& c, int f0, int f1, hipStream_t s) {
  hipLaunchKernelGGL(k_prep<N>, dim3((f1 - f0 + 127) / 128), dim3(128), 0, s, c.sc, f0, f1, c.params, c.bra, c.ket);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------ thin rows
// Rows r0..r1-1 when there are few of them (THIN_ROWS or fewer): the tile
// kernels would leave most threads of each tile idle while one thread walks
// all T terms of its pair.  Here a group of G >= T lanes takes one pair, a
// lane per term; lane 0 adds the terms in term order (the same sums as
// pair_sum_energy) into raw (H, S) of the pair.  Pairs (k, l), l <= k,
// numbered row by row from (r0, 0).
constexpr int THIN_ROWS = 32;

__device__ inline void row_pair(long long p, int r0, int& k, int& l) {
  const long long x = (long long)r0 * (r0 + 1) / 2 + p;
  int r = (int)((sqrt(8.0 * (double)x + 1.0) - 1.0) * 0.5);
  while ((long long)(r + 1) * (r + 2) / 2 <= x) ++r;
  while ((long long)r * (r + 1) / 2 > x) --r;
  k = r;
  l = (int)(x - (long long)r * (r + 1) / 2);
}

template <int N, int G>
__global__ void __launch_bounds__(256) k_rows_raw(const SysConst sc, int r0, long long npairs,
                                                  const double* __restrict__ bra, const double* __restrict__ ket,
                                                  zd* __restrict__ raw) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long p = gid / G;
  const int t = (int)(gid % G);
  zd h = mkz(0.0, 0.0), sv = mkz(0.0, 0.0);
  if (p < npairs && t < sc.T) {
    int k, l;
    row_pair(p, r0, k, l);
    const double* br = bra + (size_t)k * BraRec<N>::stride;
    const double* kr = ket + ((size_t)l * sc.T + t) * KetRec<N>::stride;
    const double adet = br[BraRec<N>::adet] * bra[(size_t)l * BraRec<N>::stride + BraRec<N>::adet];
    PairCore<N> pc;
    pair_energy<N>(sc, reinterpret_cast<const zd*>(br + BraRec<N>::ck), reinterpret_cast<const zd*>(kr + KetRec<N>::ct),
                   reinterpret_cast<const zd*>(kr + KetRec<N>::nl), adet, pc);
    h = sc.coef[t] * (pc.T + pc.V);
    sv = sc.coef[t] * pc.S;
  }
  // lane 0 of the group adds the terms in order: H = ((0 + h_0) + h_1) + ...
  const int base = (threadIdx.x & 63) - t;
  zd H = mkz(0.0, 0.0), S = mkz(0.0, 0.0);
  for (int q = 0; q < G; ++q) {
    const zd hq = mkz(__shfl(h.x, base + q), __shfl(h.y, base + q));
    const zd sq = mkz(__shfl(sv.x, base + q), __shfl(sv.y, base + q));
    if (q < sc.T) {
      H = H + hq;
      S = S + sq;
    }
  }
  if (p < npairs && t == 0) {
    raw[2 * p] = H;
    raw[2 * p + 1] = S;
  }
}

// Normalisation and stores of the thin rows: diagonal pairs as k_diag,
// the others as offdiag_tile (both triangles).  The rows' own diagS come
// from raw (this launch), earlier rows' from diagS.
__global__ void __launch_bounds__(256) k_rows_fin(int r0, long long npairs, int ld, const zd* __restrict__ raw,
                                                  double* __restrict__ diagS, double* __restrict__ diagH,
                                                  zd* __restrict__ Hm, zd* __restrict__ Sm) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  int k, l;
  row_pair(p, r0, k, l);
  const zd H = raw[2 * p], S = raw[2 * p + 1];
  const long long base = (long long)r0 * (r0 + 1) / 2;
  auto dS = [&](int i) {
    return i >= r0 ? raw[2 * ((long long)i * (i + 1) / 2 + i - base) + 1].x : diagS[i];
  };
  if (k == l) {
    const double sd = S.x, hd = H.x / sd;
    diagS[k] = sd;
    diagH[k] = hd;
    Hm[(size_t)k * ld + k] = mkz(hd, 0.0);
    Sm[(size_t)k * ld + k] = mkz(1.0, 0.0);
    return;
  }
  const double f = 1.0 / sqrt(dS(k) * dS(l));
  const zd hn = H * f, sn = S * f;
  Hm[(size_t)l * ld + k] = hn;
  Sm[(size_t)l * ld + k] = sn;
  Hm[(size_t)k * ld + l] = zconj(hn);
  Sm[(size_t)k * ld + l] = zconj(sn);
}

template <int N>
static int launch_rows(const Ctx& c, int r0, int r1, hipStream_t s) {
  const long long npairs = (long long)r1 * (r1 + 1) / 2 - (long long)r0 * (r0 + 1) / 2;
  const int T = c.sc.T;
  const int G = T <= 1 ? 1 : T <= 2 ? 2 : T <= 4 ? 4 : T <= 8 ? 8 : 16;
  const unsigned nb = (unsigned)((npairs * G + 255) / 256);
  if (c.ev_off0) (void)hipEventRecord(c.ev_off0, s);
  if (c.tr) c.tr->begin(TR_OFFDIAG, s);
  switch (G) {
    case 1: hipLaunchKernelGGL((k_rows_raw<N, 1>), dim3(nb), dim3(256), 0, s, c.sc, r0, npairs, c.bra, c.ket, c.rowbuf); break;
    case 2: hipLaunchKernelGGL((k_rows_raw<N, 2>), dim3(nb), dim3(256), 0, s, c.sc, r0, npairs, c.bra, c.ket, c.rowbuf); break;
    case 4: hipLaunchKernelGGL((k_rows_raw<N, 4>), dim3(nb), dim3(256), 0, s, c.sc, r0, npairs, c.bra, c.ket, c.rowbuf); break;
    case 8: hipLaunchKernelGGL((k_rows_raw<N, 8>), dim3(nb), dim3(256), 0, s, c.sc, r0, npairs, c.bra, c.ket, c.rowbuf); break;
    default: hipLaunchKernelGGL((k_rows_raw<N, 16>), dim3(nb), dim3(256), 0, s, c.sc, r0, npairs, c.bra, c.ket, c.rowbuf); break;
  }
  hipLaunchKernelGGL(k_rows_fin, dim3((unsigned)((npairs + 255) / 256)), dim3(256), 0, s, r0, npairs, c.ld, c.rowbuf,
                     c.diagS, c.diagH, c.H, c.S);
  if (c.tr) c.tr->end(s);
  if (c.ev_off1) (void)hipEventRecord(c.ev_off1, s);
  return (int)hipGetLastError();
}

size_t thin_row_entries(int cap) { return (size_t)2 * THIN_ROWS * cap; }

template <int N>
static int launch_assemble(const Ctx& c, int r0, int r1, hipStream_t s) {
  if (r1 <= r0) return 0;
  if (r1 - r0 <= THIN_ROWS && c.rowbuf) return launch_rows<N>(c, r0, r1, s);
  hipLaunchKernelGGL(k_diag<N>, dim3((r1 - r0 + 127) / 128), dim3(128), 0, s, c.sc, r0, r1, c.ld, c.bra, c.ket,
                     c.diagS, c.diagH, c.H, c.S);
  const int tr0 = r0 / TILE, tr1 = (r1 + TILE - 1) / TILE;
  const long long ntiles = (long long)tr1 * (tr1 + 1) / 2 - (long long)tr0 * (tr0 + 1) / 2;
  const size_t lds = sizeof(double) * TILE * (BraRec<N>::stride + c.T * KetRec<N>::stride);
  if (c.ev_off0) (void)hipEventRecord(c.ev_off0, s);
  if (c.tr) c.tr->begin(TR_OFFDIAG, s);
  hipLaunchKernelGGL(k_offdiag<N>, dim3((unsigned)ntiles), dim3(TILE, TILE), lds, s, c.sc, r0, r1, c.ld, tr0,
                     c.bra, c.ket, c.diagS, c.H, c.S);
  if (c.tr) c.tr->end(s);
  if (c.ev_off1) (void)hipEventRecord(c.ev_off1, s);
  return (int)hipGetLastError();
}

template <int N>
static int launch_deriv(const Ctx& c, int nfru, zd* D, hipStream_t s) {
  const int K = c.K;
  hipLaunchKernelGGL(k_diag<N>, dim3((K - nfru + 127) / 128), dim3(128), 0, s, c.sc, nfru, K, c.ld, c.bra, c.ket,
                     c.diagS, c.diagH, c.H, c.S);
  const long long n = (long long)K * (K + 1) / 2 - (long long)nfru * (nfru + 1) / 2;
  hipLaunchKernelGGL(k_deriv<N>, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, c.sc, K, nfru, c.ld, c.bra,
                     c.ket, c.diagS, c.H, c.S, D);
  return (int)hipGetLastError();
}

// Thin rows of the adjoint (functions nfru..K-1, at most THIN_ROWS of
// them): a group of G >= T lanes per pair (k, l), k >= nfru, l <= k, a
// lane per term; lane 0 adds the terms in order (the sums of
// pair_adjoint_sum) and stores the pair's bra and ket values (nv each) in
// raw.  k_adj_rows_slots then fills the partial slots of each function as
// adjoint_tile would (same 16-pair sums in the same order), so
// k_adjoint_finish is unchanged.
template <int N, int G>
__global__ void __launch_bounds__(256) k_adj_rows(const SysConst sc, int K, int nfru, long long npairs, double Ein,
                                                  const double* __restrict__ Ep,
                                                  const double* __restrict__ bra, const double* __restrict__ ket,
                                                  const double* __restrict__ diagS, const zd* __restrict__ cvec,
                                                  double* __restrict__ raw) {
  constexpr int m = N * (N + 1) / 2, nv = AdjAcc<N>::nvals;
  const double E = Ep ? *Ep : Ein;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long p = gid / G;
  const int t = (int)(gid % G);
  int k = 0, l = 0;
  if (p < npairs) row_pair(p, nfru, k, l);
  const bool want_l = (l != k) && (l >= nfru);
  zd hs = mkz(0.0, 0.0), gb[m], g2[m];
#pragma unroll
  for (int e = 0; e < m; ++e) gb[e] = g2[e] = mkz(0.0, 0.0);
  if (p < npairs && t < sc.T) {
    const double* br = bra + (size_t)k * BraRec<N>::stride;
    const double* kr = ket + ((size_t)l * sc.T + t) * KetRec<N>::stride;
    const zd* ck = reinterpret_cast<const zd*>(br + BraRec<N>::ck);
    const zd* ct = reinterpret_cast<const zd*>(kr + KetRec<N>::ct);
    const double adet = br[BraRec<N>::adet] * bra[(size_t)l * BraRec<N>::stride + BraRec<N>::adet];
    PairCore<N> pc;
    pair_energy<N>(sc, ck, ct, reinterpret_cast<const zd*>(kr + KetRec<N>::nl), adet, pc);
    zd gk[m], gbt[m];
    pair_adjoint<N>(sc, ck, ct, pc, E, gbt, gk);
    const double c = sc.coef[t];
    hs = c * (pc.T + pc.V - E * pc.S);
#pragma unroll
    for (int e = 0; e < m; ++e) gb[e] = c * gbt[e];
    if (want_l) {
      zd gp[m];
      perm_congruence<N>(sc.P[t], gk, gp);
#pragma unroll
      for (int e = 0; e < m; ++e) g2[e] = c * gp[e];
    }
  }
  // lane 0 of the group: the terms in order
  const int base = (threadIdx.x & 63) - t;
  zd HS = mkz(0.0, 0.0), Gb[m], Gk[m];
#pragma unroll
  for (int e = 0; e < m; ++e) Gb[e] = Gk[e] = mkz(0.0, 0.0);
  for (int q = 0; q < G; ++q) {
    const bool use = q < sc.T;
    const zd hq = mkz(__shfl(hs.x, base + q), __shfl(hs.y, base + q));
    if (use) HS = HS + hq;
#pragma unroll
    for (int e = 0; e < m; ++e) {
      const zd bq = mkz(__shfl(gb[e].x, base + q), __shfl(gb[e].y, base + q));
      const zd kq = mkz(__shfl(g2[e].x, base + q), __shfl(g2[e].y, base + q));
      if (use) {
        Gb[e] = Gb[e] + bq;
        if (want_l) Gk[e] = Gk[e] + kq;
      }
    }
  }
  if (p >= npairs || t != 0) return;
  const zd w = (zconj(cvec[k]) * cvec[l]) * (1.0 / sqrt(diagS[k] * diagS[l]));
  const zd a = w * (1.5 * HS);
  double* vb = raw + (size_t)p * 2 * nv;
  double* vk = vb + nv;
#pragma unroll
  for (int e = 0; e < m; ++e) {
    const zd g = w * Gb[e];
    vb[e] = g.x;
    vb[m + e] = g.y;
  }
  vb[2 * m] = a.x;
  if (want_l) {
#pragma unroll
    for (int e = 0; e < m; ++e) {
      const zd g = w * Gk[e];
      vk[e] = g.x;
      vk[m + e] = -g.y;
    }
    vk[2 * m] = a.x;
  } else {
#pragma unroll
    for (int e = 0; e < nv; ++e) vk[e] = 0.0;
  }
}

// One thread per (function f >= nfru, slot): slot c <= f/16 sums the bra
// values of pairs (f, 16c..16c+15) with l <= f; slot r' + 1 (r' >= f/16)
// sums the ket values of pairs (16r'..16r'+15, f) with k > f.
template <int N>
__global__ void __launch_bounds__(256) k_adj_rows_slots(int K, int nfru, int nt, const double* __restrict__ raw,
                                                        double* __restrict__ part) {
  constexpr int nv = AdjAcc<N>::nvals;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = nfru + (int)(gid / (nt + 1));
  const int slot = (int)(gid % (nt + 1));
  if (f >= K) return;
  const int r = f / TILE;
  const long long base = (long long)nfru * (nfru + 1) / 2;
  double sum[nv];
#pragma unroll
  for (int e = 0; e < nv; ++e) sum[e] = 0.0;
  if (slot <= r) {
    for (int y = 0; y < TILE; ++y) {
      const int l = slot * TILE + y;
      if (l > f) break;
      const double* v = raw + ((long long)f * (f + 1) / 2 + l - base) * 2 * nv;
#pragma unroll
      for (int e = 0; e < nv; ++e) sum[e] += v[e];
    }
  } else {
    const int rp = slot - 1;
    for (int y = 0; y < TILE; ++y) {
      const int k = rp * TILE + y;
      if (k >= K) break;
      if (k <= f) continue;
      const double* v = raw + ((long long)k * (k + 1) / 2 + f - base) * 2 * nv + nv;
#pragma unroll
      for (int e = 0; e < nv; ++e) sum[e] += v[e];
    }
  }
  double* out = part + ((size_t)f * (nt + 1) + slot) * nv;
#pragma unroll
  for (int e = 0; e < nv; ++e) out[e] = sum[e];
}

template <int N>
static void launch_adj_rows(const Ctx& c, int nfru, double E, const double* Ep, const zd* cvec, double* part,
                            hipStream_t s) {
  const int K = c.K, nt = (K + TILE - 1) / TILE;
  const long long npairs = (long long)K * (K + 1) / 2 - (long long)nfru * (nfru + 1) / 2;
  const int T = c.sc.T;
  const int G = T <= 1 ? 1 : T <= 2 ? 2 : T <= 4 ? 4 : T <= 8 ? 8 : 16;
  const unsigned nb = (unsigned)((npairs * G + 255) / 256);
  switch (G) {
    case 1: hipLaunchKernelGGL((k_adj_rows<N, 1>), dim3(nb), dim3(256), 0, s, c.sc, K, nfru, npairs, E, Ep, c.bra, c.ket, c.diagS, cvec, c.adjraw); break;
    case 2: hipLaunchKernelGGL((k_adj_rows<N, 2>), dim3(nb), dim3(256), 0, s, c.sc, K, nfru, npairs, E, Ep, c.bra, c.ket, c.diagS, cvec, c.adjraw); break;
    case 4: hipLaunchKernelGGL((k_adj_rows<N, 4>), dim3(nb), dim3(256), 0, s, c.sc, K, nfru, npairs, E, Ep, c.bra, c.ket, c.diagS, cvec, c.adjraw); break;
    case 8: hipLaunchKernelGGL((k_adj_rows<N, 8>), dim3(nb), dim3(256), 0, s, c.sc, K, nfru, npairs, E, Ep, c.bra, c.ket, c.diagS, cvec, c.adjraw); break;
    default: hipLaunchKernelGGL((k_adj_rows<N, 16>), dim3(nb), dim3(256), 0, s, c.sc, K, nfru, npairs, E, Ep, c.bra, c.ket, c.diagS, cvec, c.adjraw); break;
  }
  const long long nslots = (long long)(K - nfru) * (nt + 1);
  hipLaunchKernelGGL(k_adj_rows_slots<N>, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s, K, nfru, nt,
                     c.adjraw, part);
}

size_t thin_adjoint_entries(int n, int cap) {
  const int m = n * (n + 1) / 2;
  return (size_t)THIN_ROWS * cap * 2 * (2 * m + 1);
}

template <int N>
static int launch_adjoint(const Ctx& c, int nfru, double E, const double* Ep, const zd* cvec, double* part,
                          double* grad, hipStream_t s) {
  const int K = c.K;
  const int nt = (K + TILE - 1) / TILE;
  const int tr0 = nfru / TILE;
  const long long ntiles = (long long)nt * (nt + 1) / 2 - (long long)tr0 * (tr0 + 1) / 2;
  const size_t lds = sizeof(double) * TILE *
                     (BraRec<N>::stride + c.T * KetRec<N>::stride + TILE * AdjAcc<N>::nvals);
  if (c.ev_adj0) (void)hipEventRecord(c.ev_adj0, s);
  if (c.tr) c.tr->begin(TR_ADJOINT, s);
  if (K - nfru <= THIN_ROWS && c.adjraw)
    launch_adj_rows<N>(c, nfru, E, Ep, cvec, part, s);
  else
    hipLaunchKernelGGL(k_adjoint<N>, dim3((unsigned)ntiles), dim3(TILE, TILE), lds, s, c.sc, K, nfru, tr0, nt, E,
                       Ep, c.bra, c.ket, c.diagS, cvec, part);
  if (c.tr) c.tr->end(s);
  if (c.ev_adj1) (void)hipEventRecord(c.ev_adj1, s);
  hipLaunchKernelGGL(k_adjoint_finish<N>, dim3((K - nfru + 3) / 4), dim3(256), 0, s, K, nfru, nt, tr0, c.bra, part,
                     grad);
  return (int)hipGetLastError();
}

template <int N>
static int launch_penalty(const Ctx& c, int nfru, double pc, double t2, double* pen, double* total, double* grad,
                          hipStream_t s) {
  const int nf = c.K - nfru;
  hipLaunchKernelGGL(k_overlap_penalty<N>, dim3((nf + 3) / 4), dim3(256), 0, s, c.sc, c.K, nfru, c.ld, c.bra, c.ket,
                     c.diagS, c.S, pc, t2, pen, grad);
  hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(256), 0, s, pen, nf, total);
  return (int)hipGetLastError();
}

int prep(const Ctx& c, hipStream_t s) { return prep_range(c, 0, c.K, s); }
template <int N>
static int launch_prep_put(const Ctx& c, int f0, int count, const ParamBlock& pb, hipStream_t s) {
  hipLaunchKernelGGL(k_prep_put<N>, dim3((count + 63) / 64), dim3(64), 0, s, c.sc, f0, count, pb, c.params, c.bra,
                     c.ket);
  return (int)hipGetLastError();
}
int prep_put(const Ctx& c, int f0, int count, const double* hparams, hipStream_t s) {
  const int np = c.n * (c.n + 1);
  if (f0 < 0 || count < 1 || f0 + count > c.K || (size_t)np * count > PB_DOUBLES) return -1;
  ParamBlock pb;
  for (int e = 0; e < np * count; ++e) pb.v[e] = hparams[e];
  switch (c.n) {
    case 1: return launch_prep_put<1>(c, f0, count, pb, s);
    case 2: return launch_prep_put<2>(c, f0, count, pb, s);
    case 3: return launch_prep_put<3>(c, f0, count, pb, s);
    case 4: return launch_prep_put<4>(c, f0, count, pb, s);
    case 5: return launch_prep_put<5>(c, f0, count, pb, s);
  }
  return -1;
}
int prep_range(const Ctx& c, int f0, int f1, hipStream_t s) {
  if (f0 < 0 || f1 > c.K) return -1;
  if (f0 >= f1) return 0;
  switch (c.n) {
    case 1: return launch_prep<1>(c, f0, f1, s);
    case 2: return launch_prep<2>(c, f0, f1, s);
    case 3: return launch_prep<3>(c, f0, f1, s);
    case 4: return launch_prep<4>(c, f0, f1, s);
    case 5: return launch_prep<5>(c, f0, f1, s);
  }
  return -1;
}
int assemble(const Ctx& c, int r0, int r1, hipStream_t s) {
  switch (c.n) {
    case 1: return launch_assemble<1>(c, r0, r1, s);
    case 2: return launch_assemble<2>(c, r0, r1, s);
    case 3: return launch_assemble<3>(c, r0, r1, s);
    case 4: return launch_assemble<4>(c, r0, r1, s);
    case 5: return launch_assemble<5>(c, r0, r1, s);
  }
  return -1;
}
int assemble_deriv(const Ctx& c, int nfru, zd* D, hipStream_t s) {
  switch (c.n) {
    case 1: return launch_deriv<1>(c, nfru, D, s);
    case 2: return launch_deriv<2>(c, nfru, D, s);
    case 3: return launch_deriv<3>(c, nfru, D, s);
    case 4: return launch_deriv<4>(c, nfru, D, s);
    case 5: return launch_deriv<5>(c, nfru, D, s);
  }
  return -1;
}
int adjoint(const Ctx& c, int nfru, double E, const double* Ep, const zd* cvec, double* part, double* grad,
            hipStream_t s) {
  switch (c.n) {
    case 1: return launch_adjoint<1>(c, nfru, E, Ep, cvec, part, grad, s);
    case 2: return launch_adjoint<2>(c, nfru, E, Ep, cvec, part, grad, s);
    case 3: return launch_adjoint<3>(c, nfru, E, Ep, cvec, part, grad, s);
    case 4: return launch_adjoint<4>(c, nfru, E, Ep, cvec, part, grad, s);
    case 5: return launch_adjoint<5>(c, nfru, E, Ep, cvec, part, grad, s);
  }
  return -1;
}
int overlap_penalty(const Ctx& c, int nfru, double pc, double t2, double* pen, double* total, double* grad,
                    hipStream_t s) {
  if (nfru < 0 || nfru >= c.K) return -1;
  switch (c.n) {
    case 1: return launch_penalty<1>(c, nfru, pc, t2, pen, total, grad, s);
    case 2: return launch_penalty<2>(c, nfru, pc, t2, pen, total, grad, s);
    case 3: return launch_penalty<3>(c, nfru, pc, t2, pen, total, grad, s);
    case 4: return launch_penalty<4>(c, nfru, pc, t2, pen, total, grad, s);
    case 5: return launch_penalty<5>(c, nfru, pc, t2, pen, total, grad, s);
  }
  return -1;
}
size_t adjoint_partial_doubles(int n, int K) {
  const int nt = (K + TILE - 1) / TILE;
  const int m = n * (n + 1) / 2;
  return (size_t)K * (nt + 1) * (2 * m + 1);
}

size_t cand_rows_lds(int n, int T, int cpb) {
  if (cpb < 1 || 128 % cpb) return 0;
  const size_t bytes = sizeof(double) * ((size_t)(128 / cpb) * T * ket_stride(n) + (size_t)cpb * bra_stride(n));
  return bytes <= 64 * 1024 ? bytes : 0;
}
int candidate_rows(const Ctx& c, int Kp, int C, const double* cparams, double* cbra, double* cket,
                   double* cdiag, zd* Hb, zd* Sb, int ldb, int cpb, hipStream_t s, bool rows_only) {
  switch (c.n) {
    case 1: return launch_cand<1>(c, Kp, C, cparams, cbra, cket, cdiag, Hb, Sb, ldb, cpb, rows_only, s);
    case 2: return launch_cand<2>(c, Kp, C, cparams, cbra, cket, cdiag, Hb, Sb, ldb, cpb, rows_only, s);
    case 3: return launch_cand<3>(c, Kp, C, cparams, cbra, cket, cdiag, Hb, Sb, ldb, cpb, rows_only, s);
    case 4: return launch_cand<4>(c, Kp, C, cparams, cbra, cket, cdiag, Hb, Sb, ldb, cpb, rows_only, s);
    case 5: return launch_cand<5>(c, Kp, C, cparams, cbra, cket, cdiag, Hb, Sb, ldb, cpb, rows_only, s);
  }
  return -1;
}

int packed_pairs(int n, bool adjoint, const ProbDesc* dpd, int np, long long ntiles, int nf, int maxT,
                 hipStream_t s) {
  auto go = [&](auto tag) -> int {
    constexpr int N = decltype(tag)::value;
    if (!adjoint) {
      const size_t lds = sizeof(double) * TILE * (BraRec<N>::stride + maxT * KetRec<N>::stride);
      hipLaunchKernelGGL(k_offdiag_multi<N>, dim3((unsigned)ntiles), dim3(TILE, TILE), lds, s, dpd, np);
    } else {
      const size_t lds = sizeof(double) * TILE *
                         (BraRec<N>::stride + maxT * KetRec<N>::stride + TILE * AdjAcc<N>::nvals);
      hipLaunchKernelGGL(k_adjoint_multi<N>, dim3((unsigned)ntiles), dim3(TILE, TILE), lds, s, dpd, np);
      hipLaunchKernelGGL(k_adjoint_finish_multi<N>, dim3((nf + 3) / 4), dim3(256), 0, s, dpd, np, nf);
    }
    return (int)hipGetLastError();
  };
  switch (n) {
    case 1: return go(std::integral_constant<int, 1>());
    case 2: return go(std::integral_constant<int, 2>());
    case 3: return go(std::integral_constant<int, 3>());
    case 4: return go(std::integral_constant<int, 4>());
    case 5: return go(std::integral_constant<int, 5>());
  }
  return -1;
}

int assemble_diag(const Ctx& c, hipStream_t s) {
  auto go = [&](auto tag) -> int {
    constexpr int N = decltype(tag)::value;
    hipLaunchKernelGGL(k_diag<N>, dim3((c.K + 127) / 128), dim3(128), 0, s, c.sc, 0, c.K, c.ld, c.bra, c.ket,
                       c.diagS, c.diagH, c.H, c.S);
    return (int)hipGetLastError();
  };
  switch (c.n) {
    case 1: return go(std::integral_constant<int, 1>());
    case 2: return go(std::integral_constant<int, 2>());
    case 3: return go(std::integral_constant<int, 3>());
    case 4: return go(std::integral_constant<int, 4>());
    case 5: return go(std::integral_constant<int, 5>());
  }
  return -1;
}

}  // namespace ecg
