// Dense Hermitian linear algebra of the eigen step, FP64 complex:
//   * F = H - sigma S (lower triangle)
//   * blocked right-looking LDL^H of F without pivoting (the reference's
//     LDLHF, src/linalg.f90:316-455, reorganised in 64-wide blocks), with
//     the negative-pivot count (Sylvester inertia) and zero-pivot detection
//   * the solve x = F^{-1} b through L, D, L^H (LDLHS, src/linalg.f90:618-775)
//   * y = A x and the quadratic forms used by inverse iteration
//     (MHMV, VMMHMV, src/linalg.f90:1019-1107, 1180-1255)
// L overwrites the strict lower triangle of F.  X = L^{-1} is built by
// block rows alongside the factorisation and kept as X^H in the strict
// upper triangle, so that a solve is two GEMV passes over that triangle.
#include <algorithm>

#include "ecg_internal.hpp"
#include "ecg_schedule.hpp"

namespace ecg {

// ---------------------------------------------------------------- shift
// The first pass of a factorisation reads the lower triangle of
// M = H - sigma S straight from the assembled H and S (src.H non-null):
// every lower element is read once there (block column 0 by the first
// diagonal block and panel, block 1's diagonal tile by the first panel's
// diagonal strips, every other lower tile by the first trailing update)
// and written to F, so M is never formed as a matrix of its own.  With
// src.H null the element comes from F (later passes, the tail update).
__device__ inline zd ld_src(const ShiftSrc& src, const zd* __restrict__ F, size_t i, bool ok) {
  if (!src.H) return ld_or0(F, i, ok);
  const zd h = ld_or0(src.H, i, ok), s = ld_or0(src.S, i, ok);
  return mkz(fma(-src.sigma, s.x, h.x), fma(-src.sigma, s.y, h.y));
}

// ------------------------------------------------ workgroup flags
// Device-side waits are bounded in time: after WAIT_LIMIT_TICKS of the
// constant-rate wall clock (2 s at its 100 MHz) the waiting wave records
// status[2] and proceeds, so every wave exits; the host reports
// ECG_ERR_SYNC.
constexpr unsigned long long WAIT_LIMIT_TICKS = 200000000ull;

__device__ inline void publish_wg(int* flag, int epoch) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// The poll itself is relaxed; once both flags are seen, an acquire fence at
// agent scope pairs with the producer's release.  The eight XCDs of the
// chip each have their own L2, which is not coherent with the others':
// the fence invalidates this CU's vector cache and the non-coherent lines
// of its XCD's L2, so the loads after the barrier read what the producer
// (possibly on another XCD) wrote back before publishing.  The workgroup
// barrier then carries the acquire to the other threads.
__device__ inline void wait_wg(const int* f1, const int* f2, int epoch, int* status) {
  if (threadIdx.x == 0) {
    const unsigned long long t0 = wall_clock64();
    while (__hip_atomic_load(f1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch ||
           __hip_atomic_load(f2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
      __builtin_amdgcn_s_sleep(2);
      if (wall_clock64() - t0 > WAIT_LIMIT_TICKS) {
        atomicExch(&status[2], 1);
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// --------------------------------------------------- diagonal block LDL^H
// One workgroup of 256 threads (4 waves) factors the jb x jb diagonal block
// at (j0, j0), held in one 64 x 64 LDS array, in 16-wide sub-blocks s:
//   A. the four waves factor the 16 x 16 diagonal piece with the piece in
//      registers (one element per lane) and one LDS broadcast of four pivot
//      rows per 4-pivot step; the same row operations applied to the
//      identity give X_ss = L_ss^{-1} (Gauss-Jordan);
//   B. the sub-panel below: L = A X_ss^H D^{-1}, one 16-row tile per wave
//      on MFMA;
//   C. the trailing update A -= (L D) L^H on 16 x 16 lower tiles (MFMA).
// The off-diagonal blocks of X = L^{-1} are computed in the MFMA slots
// that B and C leave idle (x_side_job), with two closing rounds.
// LDS use: lower triangle = the block, then L (the pivots on the diagonal
// while factoring); strict upper triangle = X transposed (X(i, j), i > j,
// at A[j][i]; X has a unit diagonal).
// Outputs: unit-lower L (strict lower part of F), D, X^H (strict upper part
// of F: the diagonal block of X^H = L^{-H}), and X and X^H (column-major
// 64 x 64 each, used by the panel solve and the inverse's block rows).
constexpr int SB = 16;   // sub-block width
typedef double dbl4 __attribute__((ext_vector_type(4)));

struct DiagLds {
  zd A[NB][NB + 1];
  zd rowA[2][4][SB];   // pivot-block rows broadcast (double-buffered by step parity)
  zd rowX[2][4][SB];
  double d[NB];
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 16 x 16 complex tile product on MFMA: acc(r, c) += sum_{k<16} a(r, k) b(k, c),
// with a(r, k) and b(k, c) supplied by functors (LDS reads).  Accumulator
// element v of a lane is (row 4 v + lane / 16, column lane % 16).
struct Acc16 {
  dbl4 re, im;
};
__device__ inline void acc16_zero(Acc16& a) {
  a.re = dbl4{0.0, 0.0, 0.0, 0.0};
  a.im = dbl4{0.0, 0.0, 0.0, 0.0};
}
template <class FA, class FB>
__device__ __attribute__((always_inline)) inline void mfma16(const FA& fa, const FB& fb, Acc16& acc) {
  const int lane = threadIdx.x & 63, r = lane & 15, k = lane >> 4;
  zd av[4], bv[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    av[s] = fa(r, 4 * s + k);
    bv[s] = fb(4 * s + k, r);
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s].x, bv[s].x, acc.re, 0, 0, 0);
    acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[s].y, bv[s].y, acc.re, 0, 0, 0);
    acc.im = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s].x, bv[s].y, acc.im, 0, 0, 0);
    acc.im = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s].y, bv[s].x, acc.im, 0, 0, 0);
  }
}
__device__ inline int acc16_row(int v) { return 4 * v + ((threadIdx.x & 63) >> 4); }
__device__ inline int acc16_col() { return threadIdx.x & 15; }

// 2 x 2 Hermitian pivot pair [a, conj(b); b, c] (a, c real): d_0 = a,
// d_1 = det / a with det = a c - |b|^2.  Both reciprocals come from one,
// r = 1 / (a det): 1/d_0 = det r, 1/d_1 = a^2 r (hardware reciprocal and
// two Newton steps, below 1 ulp).  A zero pivot (recorded by the caller)
// leaves its column unscaled; other exceptional magnitudes use division.
__device__ inline void pivot_pair(double a, zd b, double c, double& rd0, double& rd1, double& d1) {
  const double det = fma(a, c, -fma(b.x, b.x, b.y * b.y));
  const double prod = a * det;
  if (prod != 0.0 && fabs(prod) > 1e-300 && fabs(prod) < 1e300) {
    double r = __builtin_amdgcn_rcp(prod);
    r = fma(fma(-prod, r, 1.0), r, r);
    r = fma(fma(-prod, r, 1.0), r, r);
    rd0 = det * r;
    rd1 = a * a * r;
    d1 = det * rd0;
  } else {
    rd0 = (a != 0.0) ? 1.0 / a : 0.0;
    d1 = (a != 0.0) ? det * rd0 : c;
    rd1 = (d1 != 0.0) ? 1.0 / d1 : 0.0;
  }
}

// One of four complex values by index, selected per component (a select of
// the struct itself goes through scratch memory).
__device__ inline zd zsel4(int k, zd v0, zd v1, zd v2, zd v3) {
  return mkz(k == 0 ? v0.x : (k == 1 ? v1.x : (k == 2 ? v2.x : v3.x)),
             k == 0 ? v0.y : (k == 1 ? v1.y : (k == 2 ? v2.y : v3.y)));
}

// Pivot d of global column col (1-based): stored, counted if negative, the
// first zero recorded.
__device__ inline void pivot_record(double d, bool valid, double* dst, int* status, int col) {
  if (!valid) return;
  *dst = d;
  if (d < 0.0) atomicAdd(&status[0], 1);
  if (d == 0.0) atomicCAS(&status[1], 0, col);
}

// X(i, j) of the block from the transposed store (unit diagonal, zero above)
__device__ inline zd x_at(const DiagLds& L, int i, int j) {
  const zd v = L.A[j][i];
  return mkz(i > j ? v.x : (i == j ? 1.0 : 0.0), i > j ? v.y : 0.0);
}

// X_ts (t > s, 16 x 16) = -X_tt T where T = sum_{u=s}^{t-1} L_tu X_us has
// been staged at X_ts's own place; result returned in acc (not yet stored).
__device__ __attribute__((always_inline)) inline void x_block_finish(const DiagLds& L, int t, int s, Acc16& acc) {
  acc16_zero(acc);
  const int ti = t * SB, si = s * SB;
  mfma16([&](int r, int k) { return x_at(L, ti + r, ti + k); },
         [&](int k, int c) { return L.A[si + c][ti + k]; }, acc);
}

// Side jobs computing the off-diagonal 16 x 16 blocks of X = L^{-1} in the
// MFMA slots that phases B and C leave idle, then two closing rounds:
//   X_ts = -X_tt T_ts,  T_ts = sum_{u=s}^{t-1} L_tu X_us,
// T staged (transposed) where X_ts goes.  Per phase each job writes its
// own block and reads only blocks completed in earlier phases.
//   phase C, s=0: wave 2 T10 = L10 X00;  wave 3 T20 = L20 X00
//   phase B, s=1: wave 2 X10;            wave 3 T30 = L30 X00
//   phase C, s=1: wave 3 T21 = L21 X11
//   phase B, s=2: wave 1 T20 += L21 X10; wave 2 T31 = L31 X11; wave 3 T30 += L31 X10
//   phase C, s=2: wave 1 X21; wave 2 X20; wave 3 T32 = L32 X22
//   closing 1:    wave 0 X32; wave 1 T31 += L32 X21; wave 2 T30 += L32 X20
//   closing 2:    wave 1 X31; wave 2 X30
// Jobs of blocks t >= nsub (a short last diagonal block) are skipped.
struct XJob {
  int kind;   // 0 none, 1 T (+)= L_tu X_us, 2 X_ts = -X_tt T_ts
  int t, s, u;
  bool acc;
};
__device__ inline XJob x_side_job(int phase, int s, int wave) {
  // phase 0: B, 1: C, 2: closing 1, 3: closing 2
  switch (phase * 16 + s * 4 + wave) {
    case 1 * 16 + 0 * 4 + 2: return {1, 1, 0, 0, false};
    case 1 * 16 + 0 * 4 + 3: return {1, 2, 0, 0, false};
    case 0 * 16 + 1 * 4 + 2: return {2, 1, 0, 0, false};
    case 0 * 16 + 1 * 4 + 3: return {1, 3, 0, 0, false};
    case 1 * 16 + 1 * 4 + 3: return {1, 2, 1, 1, false};
    case 0 * 16 + 2 * 4 + 1: return {1, 2, 0, 1, true};
    case 0 * 16 + 2 * 4 + 2: return {1, 3, 1, 1, false};
    case 0 * 16 + 2 * 4 + 3: return {1, 3, 0, 1, true};
    case 1 * 16 + 2 * 4 + 1: return {2, 2, 1, 0, false};
    case 1 * 16 + 2 * 4 + 2: return {2, 2, 0, 0, false};
    case 1 * 16 + 2 * 4 + 3: return {1, 3, 2, 2, false};
    case 2 * 16 + 0: return {2, 3, 2, 0, false};
    case 2 * 16 + 1: return {1, 3, 1, 2, true};
    case 2 * 16 + 2: return {1, 3, 0, 2, true};
    case 3 * 16 + 1: return {2, 3, 1, 0, false};
    case 3 * 16 + 2: return {2, 3, 0, 0, false};
  }
  return {0, 0, 0, 0, false};
}
__device__ __attribute__((always_inline)) inline void x_side_run(DiagLds& L, int nsub, const XJob j) {
  if (j.kind == 0 || j.t >= nsub) return;
  const int ti = j.t * SB, si = j.s * SB;
  Acc16 acc;
  if (j.kind == 1) {
    const int ui = j.u * SB;
    if (j.acc) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const zd o = L.A[si + acc16_col()][ti + acc16_row(v)];
        acc.re[v] = o.x;
        acc.im[v] = o.y;
      }
    } else {
      acc16_zero(acc);
    }
    mfma16([&](int r, int k) { return L.A[ti + r][ui + k]; },
           [&](int k, int c) { return x_at(L, ui + k, si + c); }, acc);
#pragma unroll
    for (int v = 0; v < 4; ++v) L.A[si + acc16_col()][ti + acc16_row(v)] = mkz(acc.re[v], acc.im[v]);
  } else {
    x_block_finish(L, j.t, j.s, acc);
#pragma unroll
    for (int v = 0; v < 4; ++v) L.A[si + acc16_col()][ti + acc16_row(v)] = mkz(-acc.re[v], -acc.im[v]);
  }
}

__device__ __attribute__((always_inline)) inline void ldl_diag_block(int K, int ld, int j0, zd* __restrict__ F,
                                                                     double* __restrict__ D,
                                                                     zd* __restrict__ Dinv,
                                                                     int* __restrict__ status, DiagLds& L,
                                                                     bool a_in_lds, int* pub = nullptr,
                                                                     int epoch = 0, ShiftSrc src = ShiftSrc()) {
  const int jb = min(NB, K - j0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (!a_in_lds) {
    zd v[NB * NB / 256];
#pragma unroll
    for (int u = 0; u < NB * NB / 256; ++u) {
      const int e = tid + 256 * u, j = e >> 6, i = e & (NB - 1);
      v[u] = ld_src(src, F, (size_t)(j0 + j) * ld + j0 + i, i < jb && j < jb && i >= j);
    }
#pragma unroll
    for (int u = 0; u < NB * NB / 256; ++u) {
      const int e = tid + 256 * u;
      L.A[e & (NB - 1)][e >> 6] = v[u];
    }
  }
  __syncthreads();
  const int nsub = (jb + SB - 1) / SB;
  for (int s = 0; s < nsub; ++s) {
    const int c = s * SB;
    const int sb = min(SB, jb - c);
    // ---- A: 16 x 16 LDL^H and X_ss, all four waves (lane = one element:
    // row 4 wave + lane / 16, column lane % 16), four pivots P = p..p+3 per
    // broadcast: wave g holds rows P of step g and publishes them (A and X
    // rows) through LDS; every lane factors the 4 x 4 pivot block itself
    // (two 2 x 2 pivot pairs) and applies the block step to its element:
    //   L(i, P) = A(i, P) L_P^{-H} D_P^{-1},  U(P, j) = L_P^{-1} A(P, j),
    //   A(i, j) -= L(i, P) U(P, j),   X(i, j) -= L(i, P) (L_P^{-1} X(P, j)).
    {
      const int i = 4 * wave + (lane >> 4), j = lane & 15;
      zd a, x;
      {
        const zd lo = L.A[c + max(i, j)][c + min(i, j)];
        const bool in = (i < sb && j < sb);
        const double id = (i == j) ? 1.0 : 0.0;
        a = in ? mkz(lo.x, j <= i ? lo.y : -lo.y) : mkz(id, 0.0);
        x = mkz(id, 0.0);
      }
#pragma unroll
      for (int g = 0; g < SB / 4; ++g) {
        const int p = 4 * g;
        zd(*rA)[SB] = L.rowA[g & 1];
        zd(*rX)[SB] = L.rowX[g & 1];
        if (wave == g) {
          rA[i - p][j] = a;
          rX[i - p][j] = x;
        }
        __syncthreads();
        // pivot block: pair (0, 1), then the Schur complement pair (2, 3)
        const double a00 = rA[0][p].x, a11 = rA[1][p + 1].x, a22 = rA[2][p + 2].x, a33 = rA[3][p + 3].x;
        const zd a10 = rA[1][p], a20 = rA[2][p], a21 = rA[2][p + 1], a30 = rA[3][p], a31 = rA[3][p + 1],
                 a32 = rA[3][p + 2];
        double rd0, rd1, d1;
        pivot_pair(a00, a10, a11, rd0, rd1, d1);
        const zd l10 = a10 * rd0, l20 = a20 * rd0, l30 = a30 * rd0;
        const zd a21p = zfms(a20, zconj(l10), a21), a31p = zfms(a30, zconj(l10), a31);
        const zd l21 = a21p * rd1, l31 = a31p * rd1;
        const double s22 = a22 - fma(a20.x, l20.x, a20.y * l20.y) - fma(a21p.x, l21.x, a21p.y * l21.y);
        const double s33 = a33 - fma(a30.x, l30.x, a30.y * l30.y) - fma(a31p.x, l31.x, a31p.y * l31.y);
        const zd s32 = zfms(a31p, zconj(l21), zfms(a30, zconj(l20), a32));
        double rd2, rd3, d3;
        pivot_pair(s22, s32, s33, rd2, rd3, d3);
        const zd l32 = s32 * rd2;
        // L_P^{-1} applied to the block's A and X columns j
        zd u0 = rA[0][j], u1 = rA[1][j], u2 = rA[2][j], u3 = rA[3][j];
        zd x0 = rX[0][j], x1 = rX[1][j], x2 = rX[2][j], x3 = rX[3][j];
        u1 = zfms(l10, u0, u1);
        x1 = zfms(l10, x0, x1);
        u2 = zfms(l21, u1, zfms(l20, u0, u2));
        x2 = zfms(l21, x1, zfms(l20, x0, x2));
        u3 = zfms(l32, u2, zfms(l31, u1, zfms(l30, u0, u3)));
        x3 = zfms(l32, x2, zfms(l31, x1, zfms(l30, x0, x3)));
        if (i >= p + 4) {
          // z = A(i, P) L_P^{-H} with A(i, k) = conj(A(k, i)); L(i, P) = z D_P^{-1}
          const zd y0 = zconj(rA[0][i]), y1 = zconj(rA[1][i]), y2 = zconj(rA[2][i]), y3 = zconj(rA[3][i]);
          const zd z1 = zfms(y0, zconj(l10), y1);
          const zd z2 = zfms(z1, zconj(l21), zfms(y0, zconj(l20), y2));
          const zd z3 = zfms(z2, zconj(l32), zfms(z1, zconj(l31), zfms(y0, zconj(l30), y3)));
          const zd L0 = y0 * rd0, L1 = z1 * rd1, L2 = z2 * rd2, L3 = z3 * rd3;
          if (j >= p + 4) a = zfms(L3, u3, zfms(L2, u2, zfms(L1, u1, zfms(L0, u0, a))));
          else if (j >= p) a = zsel4(j - p, L0, L1, L2, L3);
          x = zfms(L3, x3, zfms(L2, x2, zfms(L1, x1, zfms(L0, x0, x))));
        } else if (i >= p) {
          const int r = i - p;
          x = zsel4(r, x0, x1, x2, x3);
          if (j >= p && j < i) {
            const int cc = j - p;
            // L_P(r, cc), r > cc: (1,0) l10; (2,0) l20, (2,1) l21; (3,0..2) l30, l31, l32
            const zd l2 = (cc == 0) ? l20 : l21;
            a = zsel4(r, l10, l10, l2, zsel4(cc, l30, l31, l32, l32));
          }
        }
        if (tid == 0) {
          pivot_record(a00, p + 0 < sb, L.d + c + p + 0, status, j0 + c + p + 1);
          pivot_record(d1, p + 1 < sb, L.d + c + p + 1, status, j0 + c + p + 2);
          pivot_record(s22, p + 2 < sb, L.d + c + p + 2, status, j0 + c + p + 3);
          pivot_record(d3, p + 3 < sb, L.d + c + p + 3, status, j0 + c + p + 4);
        }
      }
      if (i < sb && j < i) {
        L.A[c + i][c + j] = a;
        L.A[c + j][c + i] = x;
      }
    }
    __syncthreads();
    const int rows = jb - c - sb;
    const int nt = (rows + SB - 1) / SB;
    // ---- B: sub-panel L(r, c+j) = sum_{k <= j} A(r, c+k) conj(X_ss(j, k)) / d_j
    // (waves without a row tile take their X side job)
    if (wave >= nt) {
      x_side_run(L, nsub, x_side_job(0, s, wave));
    } else {
      const int r0 = c + SB * (wave + 1);
      Acc16 acc;
      acc16_zero(acc);
      mfma16([&](int r, int k) { return (r0 + r < jb) ? L.A[r0 + r][c + k] : mkz(0.0, 0.0); },
             [&](int k, int cc) { return zconj(x_at(L, c + cc, c + k)); }, acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = r0 + acc16_row(v), cc = acc16_col();
        if (r < jb) L.A[r][c + cc] = mkz(acc.re[v], acc.im[v]) * (1.0 / L.d[c + cc]);
      }
    }
    __syncthreads();
    // ---- C: trailing update on lower 16 x 16 tiles (ta >= tb)
    for (int tt = wave; tt < nt * (nt + 1) / 2; tt += 4) {
      int ta = 0;
      while ((ta + 1) * (ta + 2) / 2 <= tt) ++ta;
      const int tb = tt - ta * (ta + 1) / 2;
      const int ri = c + SB * (ta + 1), rj = c + SB * (tb + 1);
      Acc16 acc;
      acc16_zero(acc);
      mfma16([&](int r, int k) { return (ri + r < jb) ? L.A[ri + r][c + k] * L.d[c + k] : mkz(0.0, 0.0); },
             [&](int k, int cc) { return (rj + cc < jb) ? zconj(L.A[rj + cc][c + k]) : mkz(0.0, 0.0); }, acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = ri + acc16_row(v), cc = rj + acc16_col();
        if (r < jb && r >= cc) {
          const zd o = L.A[r][cc];
          L.A[r][cc] = mkz(o.x - acc.re[v], o.y - acc.im[v]);
        }
      }
    }
    x_side_run(L, nsub, x_side_job(1, s, wave));
    __syncthreads();
  }
  // ---- X = L^{-1}: the two closing rounds of the side-job schedule
  x_side_run(L, nsub, x_side_job(2, 0, wave));
  __syncthreads();
  x_side_run(L, nsub, x_side_job(3, 0, wave));
  __syncthreads();
  // ---- outputs: X and D first (all the next panel needs), then the flag
  // (look-ahead use), then X^H (also into F's strict upper triangle) and L
  zd* Xg = Dinv + (size_t)(j0 / NB) * 2 * NB * NB;
  zd* XHg = Xg + NB * NB;
  {
    zd lo[NB * NB / 256], up[NB * NB / 256];
#pragma unroll
    for (int u = 0; u < NB * NB / 256; ++u) {
      const int e = tid + 256 * u, j = e >> 6, i = e & (NB - 1);   // column-major walk: coalesced
      lo[u] = L.A[i][j];
      up[u] = L.A[j][i];
    }
#pragma unroll
    for (int u = 0; u < NB * NB / 256; ++u) {
      const int e = tid + 256 * u, j = e >> 6, i = e & (NB - 1);
      const bool in = (i < jb && j < jb);
      const double dg = (in && i == j) ? 1.0 : 0.0;
      Xg[(size_t)j * NB + i] = (in && i > j) ? up[u] : mkz(dg, 0.0);
    }
    for (int p = tid; p < jb; p += blockDim.x) D[j0 + p] = L.d[p];
    if (pub) publish_wg(pub, epoch);
#pragma unroll
    for (int u = 0; u < NB * NB / 256; ++u) {
      const int e = tid + 256 * u, j = e >> 6, i = e & (NB - 1);
      const bool in = (i < jb && j < jb);
      const double dg = (in && i == j) ? 1.0 : 0.0;
      // X^H(i, j) = conj(X(j, i)), X(j, i) (j > i) at A[i][j]; also the
      // strict upper triangle of the block in F
      const zd xh = zconj(lo[u]);
      XHg[(size_t)j * NB + i] = (in && i < j) ? xh : mkz(dg, 0.0);
      if (in && i < j) F[(size_t)(j0 + j) * ld + j0 + i] = xh;
    }
#pragma unroll
    for (int u = 0; u < NB * NB / 256; ++u) {
      const int e = tid + 256 * u, j = e >> 6, i = e & (NB - 1);
      if (i < jb && j < jb && i > j) F[(size_t)(j0 + j) * ld + j0 + i] = lo[u];
    }
  }
}

__global__ void __launch_bounds__(256) k_ldl_diag(int K, int ld, int j0, zd* __restrict__ F,
                                                  double* __restrict__ D, zd* __restrict__ Dinv,
                                                  int* __restrict__ status, ShiftSrc src) {
  __shared__ DiagLds L;
  ldl_diag_block(K, ld, j0, F, D, Dinv, status, L, false, nullptr, 0, src);
}

// ----------------------------------------------------- MFMA 64x64 tile
// acc(r, c) = sum_p A(r, p) conj(B(c, p)), r, c < 64, p < kd, with
// A(r, p) = Ap[p * lda + r] (r < ra) and B(c, p) = Bp[p * ldb + c] (c < rb).
// 4 waves, each a 32 x 32 quarter as 2 x 2 tiles of v_mfma_f64_16x16x4f64,
// four real products per complex one.  Accumulator v of a lane holds row
// 4*v + lane/16, column lane%16 of its 16x16 tile (measured lane layout of
// the instruction on gfx950).
struct Tile64 {
  dbl4 re[2][2], im[2][2];
};

template <int KS = NB / 4>   // k-steps of 4 (kd <= 4 KS)
__device__ inline void mfma_tile_abh(const zd* __restrict__ Ap, int lda, int ra, const zd* __restrict__ Bp,
                                     int ldb, int rb, int kd, Tile64& t) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rt = (wave & 1) * 32, ct = (wave >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      t.re[a][q] = dbl4{0.0, 0.0, 0.0, 0.0};
      t.im[a][q] = dbl4{0.0, 0.0, 0.0, 0.0};
    }
  // all operand fragments are loaded up front (KS k-steps x 4 complex per
  // lane) so the global loads are in flight together
  zd fa[KS][2], fb[KS][2];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int p = 4 * s + lk;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int r = rt + a * 16 + lr;
      fa[s][a] = ld_or0(Ap, (size_t)p * lda + r, p < kd && r < ra);
      const int c = ct + a * 16 + lr;
      fb[s][a] = ld_or0(Bp, (size_t)p * ldb + c, p < kd && c < rb);
    }
  }
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    double ar[2], ai[2], br[2], bi[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      ar[a] = fa[s][a].x;
      ai[a] = fa[s][a].y;
      br[a] = fb[s][a].x;
      bi[a] = fb[s][a].y;
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        // (Ar + i Ai)(Br - i Bi) = (Ar Br + Ai Bi) + i (Ai Br - Ar Bi)
        t.re[a][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[q], t.re[a][q], 0, 0, 0);
        t.re[a][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], bi[q], t.re[a][q], 0, 0, 0);
        t.im[a][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[q], t.im[a][q], 0, 0, 0);
        t.im[a][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar[a], bi[q], t.im[a][q], 0, 0, 0);
      }
  }
}

// Position (row, col) in the 64x64 tile of accumulator element (a, q, v).
__device__ inline void tile_pos(int a, int q, int v, int& r, int& c) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  r = (wave & 1) * 32 + a * 16 + 4 * v + (lane >> 4);
  c = (wave >> 1) * 32 + q * 16 + (lane & 15);
}

// ------------------------------------------------------------ panel solve
// 16-row strip i0..i0+15 below the diagonal block at j0: W = A X^H (so
// that W L_JJ^H = A), L = W D^{-1}; L goes to F, W to the panel scratch
// used by the trailing update (and, for a diagonal strip, to rows 0..15 of
// wl->A).  Wave w computes columns 16w..16w+15 as one 16 x 16 MFMA tile:
// all 16 k-steps' operand fragments are loaded first (one memory latency),
// then the 64 MFMAs.
__device__ __attribute__((always_inline)) inline void panel_strip(int K, int ld, int j0, int i0, zd* __restrict__ F,
                                                                  const double* __restrict__ D,
                                                                  const zd* __restrict__ Dinv, zd* __restrict__ W,
                                                                  DiagLds* wl = nullptr, ShiftSrc src = ShiftSrc()) {
  const int wave = threadIdx.x >> 6;
  const int c0 = SB * wave;
  const zd* X = Dinv + (size_t)(j0 / NB) * 2 * NB * NB;   // X(i, j) at X[j * NB + i]
  const size_t a0 = (size_t)j0 * ld + i0;                 // A(r, k) at F[a0 + k * ld + r]
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  zd av[NB / 4], bv[NB / 4];
#pragma unroll
  for (int t = 0; t < NB / 4; ++t) {
    const int kk = 4 * t + lk;
    av[t] = ld_src(src, F, a0 + (size_t)kk * ld + lr, i0 + lr < K);
    bv[t] = zconj(X[(size_t)kk * NB + c0 + lr]);
  }
  // two accumulator chains (even and odd k-steps) halve the dependent
  // MFMA latency of the strip, which is on the factorisation's critical path
  Acc16 acc, ac1;
  acc16_zero(acc);
  acc16_zero(ac1);
#pragma unroll
  for (int t = 0; t < NB / 4; t += 2) {
    acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t].x, bv[t].x, acc.re, 0, 0, 0);
    acc.re = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[t].y, bv[t].y, acc.re, 0, 0, 0);
    acc.im = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t].x, bv[t].y, acc.im, 0, 0, 0);
    acc.im = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t].y, bv[t].x, acc.im, 0, 0, 0);
    ac1.re = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t + 1].x, bv[t + 1].x, ac1.re, 0, 0, 0);
    ac1.re = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[t + 1].y, bv[t + 1].y, ac1.re, 0, 0, 0);
    ac1.im = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t + 1].x, bv[t + 1].y, ac1.im, 0, 0, 0);
    ac1.im = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t + 1].y, bv[t + 1].x, ac1.im, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    acc.re[v] += ac1.re[v];
    acc.im[v] += ac1.im[v];
  }
  // every wave has read all 64 columns of the strip before any wave
  // overwrites its 16 columns with L below
  __syncthreads();
  // the panel's block is a full one (a strip exists only below it), so
  // D[j0 + c] is in range for every lane
  const double rdc = 1.0 / D[j0 + c0 + acc16_col()];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = i0 + acc16_row(v), c = c0 + acc16_col();
    const zd w = mkz(acc.re[v], acc.im[v]);   // zero in rows r >= K
    if (wl) wl->A[acc16_row(v)][c] = w;
    if (r < K) {
      const zd l = w * rdc;
      W[(size_t)c * ld + r] = w;
      F[(size_t)(j0 + c) * ld + r] = l;
    }
  }
}

// Diagonal strip q's second part: rows 16q..16q+15 of the diagonal tile of
// the block at T1 = j0 + 64 (the rows of this strip), updated by the panel
// at j0:  A(T1 + R, T1 + C) -= sum_p W(T1 + R, p) conj(L(T1 + C, p)),
// C <= R.  W of the strip's rows is in rows 0..15 of Ld.A; L of rows
// T1 + 16 w.. (strip w <= q, published) is read from F.  Wave w takes
// columns 16w..16w+15; two accumulator chains (even and odd k-steps) halve
// the dependent MFMA latency.
__device__ __attribute__((always_inline)) inline void diag_quarter_update(int K, int ld, int j0, int q,
                                                                          zd* __restrict__ F, const DiagLds& Ld,
                                                                          ShiftSrc src = ShiftSrc()) {
  const int wave = threadIdx.x >> 6;
  if (wave > q) return;   // those columns lie above the quarter's rows
  const int T1 = j0 + NB;
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const int R0 = SB * q, C0 = SB * wave;
  zd old[4], av[NB / 4], bv[NB / 4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int rr = R0 + acc16_row(v), c = C0 + acc16_col();
    old[v] = ld_src(src, F, (size_t)(T1 + c) * ld + T1 + rr, T1 + rr < K && c <= rr);
  }
#pragma unroll
  for (int st = 0; st < NB / 4; ++st) {
    const int pp = 4 * st + lk;
    av[st] = Ld.A[lr][pp];
    bv[st] = zconj(ld_or0(F, (size_t)(j0 + pp) * ld + T1 + C0 + lr, T1 + C0 + lr < K));
  }
  Acc16 a0, a1;
  acc16_zero(a0);
  acc16_zero(a1);
#pragma unroll
  for (int st = 0; st < NB / 4; st += 2) {
    a0.re = __builtin_amdgcn_mfma_f64_16x16x4f64(av[st].x, bv[st].x, a0.re, 0, 0, 0);
    a0.re = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[st].y, bv[st].y, a0.re, 0, 0, 0);
    a0.im = __builtin_amdgcn_mfma_f64_16x16x4f64(av[st].x, bv[st].y, a0.im, 0, 0, 0);
    a0.im = __builtin_amdgcn_mfma_f64_16x16x4f64(av[st].y, bv[st].x, a0.im, 0, 0, 0);
    a1.re = __builtin_amdgcn_mfma_f64_16x16x4f64(av[st + 1].x, bv[st + 1].x, a1.re, 0, 0, 0);
    a1.re = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[st + 1].y, bv[st + 1].y, a1.re, 0, 0, 0);
    a1.im = __builtin_amdgcn_mfma_f64_16x16x4f64(av[st + 1].x, bv[st + 1].y, a1.im, 0, 0, 0);
    a1.im = __builtin_amdgcn_mfma_f64_16x16x4f64(av[st + 1].y, bv[st + 1].x, a1.im, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int rr = R0 + acc16_row(v), c = C0 + acc16_col();
    if (T1 + rr < K && c <= rr)
      F[(size_t)(T1 + c) * ld + T1 + rr] = mkz(old[v].x - (a0.re[v] + a1.re[v]), old[v].y - (a0.im[v] + a1.im[v]));
  }
}

// Waits on the flags t.wait[from..] (two per poll).
__device__ inline void wait_list(const int* pflag, const int* wait, int from, int epoch, int* status) {
  for (int i = from; i < UPD_WAITS && wait[i] >= 0; i += 2) {
    const int f2 = (i + 1 < UPD_WAITS && wait[i + 1] >= 0) ? wait[i + 1] : wait[i];
    wait_wg(&pflag[wait[i]], &pflag[f2], epoch, status);
  }
}

// A strip of a panel (panel_task / upd_task): L and W of its rows, and for
// a diagonal strip, once the flags it waits on are seen, its quarter of
// the next diagonal tile.
__device__ __attribute__((always_inline)) inline void strip_task(int K, int ld, int j0, const UpdTask& t,
                                                                 int first_wait, zd* __restrict__ F,
                                                                 const double* __restrict__ D,
                                                                 const zd* __restrict__ Dinv, zd* __restrict__ W,
                                                                 int* __restrict__ pflag, int epoch,
                                                                 int* __restrict__ status, DiagLds& Ld,
                                                                 ShiftSrc src = ShiftSrc()) {
  const bool dg = t.diag >= 0;
  panel_strip(K, ld, j0, t.a, F, D, Dinv, W, dg ? &Ld : nullptr, src);
  if (!dg) return;
  publish_wg(&pflag[t.publish], epoch);   // also orders the LDS copy of W
  wait_list(pflag, t.wait, first_wait, epoch, status);
  diag_quarter_update(K, ld, j0, t.diag, F, Ld, src);
}

// Panel of the first block column (later panels are computed inside the
// trailing update of the previous column), its diagonal strips also
// updating block 1's diagonal tile.  Needs a full 64-wide block.
__global__ void __launch_bounds__(256) k_ldl_panel(int K, int ld, zd* __restrict__ F, const double* __restrict__ D,
                                                   const zd* __restrict__ Dinv, zd* __restrict__ W,
                                                   int* __restrict__ status, int* __restrict__ pflag, int qflag,
                                                   int epoch, ShiftSrc src) {
  __shared__ DiagLds Ld;
  const UpdTask t = panel_task((int)blockIdx.x, qflag);
  strip_task(K, ld, 0, t, 0, F, D, Dinv, W, pflag, epoch, status, Ld, src);
}

// ------------------------------------------------- inverse of L by blocks
// X = L^{-1} (unit lower) is built by block rows alongside the
// factorisation and stored as X^H in the strict upper triangle of F: block
// (c, i), c < i, holds X_ic^H; the diagonal blocks' upper parts hold
// X_ii^H (from the diagonal factor).  Block row i satisfies
//   X_ic = -X_ii sum_{u=c}^{i-1} L_iu X_uc,   c < i.
// The launch for block column k (k >= 1) uses panel k - 1, X_kk and block
// row k - 1 of X, all complete in earlier launches:
//   row k:      X_kc^H = (R_kc^H - X_{k-1,c}^H L_{k,k-1}^H) X_kk^H,  c < k
//   rows i > k: R_ic^H -= X_{k-1,c}^H L_{i,k-1}^H,                   c < k
// where R_ic = -sum_{u=c}^{k-2} L_iu X_uc is kept (as R^H) where X_ic^H
// will go; its first update (c = k - 1) assigns.  k_ldl_xlast finishes
// the last block row.  Every product is a 64 x 64 MFMA tile A B^H.
__device__ __attribute__((always_inline)) inline void xinv_tile(int K, int ld, int k, int i, int c,
                                                                zd* __restrict__ F, const zd* __restrict__ Dinv,
                                                                DiagLds& Ld) {
  const int ib = min(NB, K - i * NB);
  const bool first = (c == k - 1);
  // X_{k-1,c}^H: block (c, k-1) of F, or the diagonal block's X^H
  const zd* T = first ? Dinv + (size_t)(k - 1) * 2 * NB * NB + NB * NB : F + (size_t)(k - 1) * NB * ld + c * NB;
  const int ldt = first ? NB : ld;
  const zd* Lb = F + (size_t)(k - 1) * NB * ld + i * NB;   // L_{i,k-1}(s, p) at Lb[p * ld + s]
  zd* out = F + (size_t)i * NB * ld + c * NB;               // block (c, i): (r, s) at out[s * ld + r]
  // the partial sums R^H this product updates, loaded up front (they are
  // zero on the first update)
  zd old[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, sc;
        tile_pos(a, q, v, r, sc);
        old[a][q][v] = ld_or0(out, (size_t)sc * ld + r, !first && sc < ib);
      }
  Tile64 t;
  mfma_tile_abh(T, ldt, NB, Lb, ld, ib, NB, t);
  if (i > k) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          int r, sc;
          tile_pos(a, q, v, r, sc);
          if (sc < ib)
            out[(size_t)sc * ld + r] = mkz(old[a][q][v].x - t.re[a][q][v], old[a][q][v].y - t.im[a][q][v]);
        }
    return;
  }
  // row k: M = R^H - T L^H staged in LDS as M(r, p) at Ld.A[p][r], then
  // X_kc^H = M X_kk^H
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, sc;
        tile_pos(a, q, v, r, sc);
        const bool in = sc < ib;
        Ld.A[sc][r] = mkz(in ? old[a][q][v].x - t.re[a][q][v] : 0.0, in ? old[a][q][v].y - t.im[a][q][v] : 0.0);
      }
  __syncthreads();
  const zd* Xkk = Dinv + (size_t)k * 2 * NB * NB;   // X_kk(s, p) at Xkk[p * NB + s]
  mfma_tile_abh(&Ld.A[0][0], NB + 1, NB, Xkk, NB, ib, ib, t);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, sc;
        tile_pos(a, q, v, r, sc);
        if (sc < ib) out[(size_t)sc * ld + r] = mkz(t.re[a][q][v], t.im[a][q][v]);
      }
}

// Last block row of X (k = nb - 1), one workgroup per block column c.
__global__ void __launch_bounds__(256) k_ldl_xlast(int K, int ld, zd* __restrict__ F, const zd* __restrict__ Dinv) {
  __shared__ DiagLds Ld;
  const int k = (K + NB - 1) / NB - 1;
  xinv_tile(K, ld, k, k, blockIdx.x, F, Dinv, Ld);
}

// -------------------------------------------------------- trailing update
// One launch per block column j (diagonal block at j0, next block t0 =
// j + 1); the task map is in ecg_schedule.hpp:
//   workgroup 0 (look-ahead): factors block t0, whose diagonal tile the
//     previous launch's diagonal strips completed.
//   then the other lower block pairs (I >= K) of the trailing matrix,
//     C_IK -= W_I L_K^H; column tiles (bi, 0) and tile (1, 1) publish.
//   then block tiles of X = L^{-1} (xinv_tile; no waits).  They come
//     before the strips so that they are dispatched while the strips would
//     only wait.
//   last: 16-row strips of the next block column's panel.  Each waits
//     until the tile holding its rows and the look-ahead factor are
//     published, then runs the panel solve for its strip; the four strips
//     of block t0 + 1 then wait for tile (1, 1) and the lower ones' rows of
//     L and update their quarter of block t0 + 1's diagonal tile with this
//     panel (diag_quarter_update), so that the next launch's look-ahead
//     starts at once.
// Publication: the producer's threads fence at agent scope, the workgroup
// synchronises, thread 0 stores the flag (release).  A waiting workgroup's
// thread 0 polls the flags (relaxed, bounded in time: past the limit it
// records status[2] and proceeds, so every workgroup exits), then fences
// with acquire at agent scope before the workgroup barrier, so that the
// data written by a producer on another XCD is read from memory, not from
// a stale line of this XCD's L2 or this CU's cache.
// Waits only point to lower workgroup indices, dispatched no later: a
// workgroup can only wait on one that the dispatcher placed before it,
// which already holds its CU slot and never waits on a higher index, so
// the chain of waits ends.  A wait on a higher index has no such
// guarantee: the waiting workgroups can fill every slot the awaited one
// needs (a variant that let panel strips wait on sibling strips of higher
// index ran past its time limit, most likely for this reason).  The task
// map lives in ecg_schedule.hpp, and a host-side check
// (tests/sched_check.cpp) enumerates it for every launch of every basis
// size and verifies this rule.
// W is double-buffered by column parity: the tiles read column j's, the
// strips write column j+1's.
// Task w of the update launch for block column col of one factorisation.
__device__ __attribute__((always_inline)) inline void ldl_update_task(int K, int ld, int col, int w,
                                                                      const UpdShape& shape, zd* __restrict__ F,
                                                                      zd* __restrict__ W, double* __restrict__ D,
                                                                      zd* __restrict__ Dinv, int* __restrict__ status,
                                                                      int* __restrict__ pflag, int qflag, int epoch,
                                                                      const ShiftSrc& src, DiagLds& Ld) {
  const int j0 = col * NB;
  const zd* Wcur = W + (size_t)(col & 1) * ld * NB;
  zd* Wnext = W + (size_t)((col + 1) & 1) * ld * NB;
  const int jb = min(NB, K - j0);
  const int t0 = (j0 + jb) / NB;
  const UpdTask task = upd_task(w, K, j0, shape, qflag);
  if (task.kind == UPD_LOOKAHEAD) {
    ldl_diag_block(K, ld, t0 * NB, F, D, Dinv, status, Ld, false, &pflag[task.publish], epoch);
    return;
  }
  if (task.kind == UPD_XINV) {
    xinv_tile(K, ld, col, task.a, task.b, F, Dinv, Ld);
    return;
  }
  if (task.kind == UPD_STRIP) {
    wait_wg(&pflag[task.wait[0]], &pflag[task.wait[1]], epoch, status);
    strip_task(K, ld, t0 * NB, task, 2, F, D, Dinv, Wnext, pflag, epoch, status, Ld);
    return;
  }
  // UPD_TILE: lower block pair (bi, bk) of the trailing matrix
  const int bi = task.a, bk = task.b;
  const int I0 = (t0 + bi) * NB, K0 = (t0 + bk) * NB;
  const int ib = min(NB, K - I0), kb = min(NB, K - K0);
  // the tile's old values, loaded up front (in flight with the operands)
  zd old[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        tile_pos(a, q, v, r, c);
        old[a][q][v] = ld_src(src, F, (size_t)(K0 + c) * ld + I0 + r, r < ib && c < kb && (I0 + r) >= (K0 + c));
      }
  Tile64 t;
  mfma_tile_abh(Wcur + I0, ld, ib, F + (size_t)j0 * ld + K0, ld, kb, jb, t);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        tile_pos(a, q, v, r, c);
        if (r < ib && c < kb && (I0 + r) >= (K0 + c))
          F[(size_t)(K0 + c) * ld + I0 + r] = mkz(old[a][q][v].x - t.re[a][q][v], old[a][q][v].y - t.im[a][q][v]);
      }
  if (task.publish >= 0) publish_wg(&pflag[task.publish], epoch);
}

__global__ void __launch_bounds__(256) k_ldl_update(int K, int ld, int col, zd* __restrict__ F, zd* __restrict__ W,
                                                    double* __restrict__ D, zd* __restrict__ Dinv,
                                                    int* __restrict__ status, int* __restrict__ pflag, int qflag,
                                                    int epoch, ShiftSrc src) {
  __shared__ DiagLds Ld;
  const UpdShape shape = upd_shape(K, col);
  ldl_update_task(K, ld, col, (int)blockIdx.x, shape, F, W, D, Dinv, status, pflag, qflag, epoch, src, Ld);
}

// ------------------------------------------------- batched factorisations
// Independent problems of a batch factored in the same launches, on a
// one-dimensional grid of P x (largest task count) workgroups: workgroup
// id serves problem id % P, task id / P, so the problems' look-ahead
// factors come first and every problem's tasks interleave.  Workgroups
// beyond a problem's own task count return at once.  The dispatcher
// places workgroups in index order, so a task's waits, which point to
// lower task indices of its own problem, point to lower workgroup indices:
// the rule of the single-problem launches carries over.
__global__ void __launch_bounds__(64) k_clear_status_b(const FacDesc* __restrict__ d) {
  if (threadIdx.x < 3) d[blockIdx.x].status[threadIdx.x] = 0;
}

__global__ void __launch_bounds__(256) k_ldl_diag_b(const FacDesc* __restrict__ d) {
  __shared__ DiagLds L;
  const FacDesc& f = d[blockIdx.x];
  ldl_diag_block(f.K, f.ld, 0, f.F, f.D, f.Dinv, f.status, L, false, nullptr, 0, f.src);
}

__global__ void __launch_bounds__(256) k_ldl_panel_b(const FacDesc* __restrict__ d, int P, int epoch) {
  __shared__ DiagLds Ld;
  const FacDesc& f = d[blockIdx.x % P];
  const int sidx = (int)(blockIdx.x / P);
  if (f.K <= NB || sidx >= (f.K - NB + SB - 1) / SB) return;
  const UpdTask t = panel_task(sidx, f.ld / NB + 1);
  strip_task(f.K, f.ld, 0, t, 0, f.F, f.D, f.Dinv, f.W, f.pflag, epoch, f.status, Ld, f.src);
}

__global__ void __launch_bounds__(256) k_ldl_update_b(const FacDesc* __restrict__ d, int P, int col, int epoch) {
  __shared__ DiagLds Ld;
  const FacDesc& f = d[blockIdx.x % P];
  const int w = (int)(blockIdx.x / P);
  if (col * NB + NB >= f.K) return;
  const UpdShape shape = upd_shape(f.K, col);
  if (w >= shape.grid) return;
  const ShiftSrc none;
  ldl_update_task(f.K, f.ld, col, w, shape, f.F, f.W, f.D, f.Dinv, f.status, f.pflag, f.ld / NB + 1, epoch,
                  col == 0 ? f.src : none, Ld);
}

__global__ void __launch_bounds__(256) k_ldl_xlast_b(const FacDesc* __restrict__ d, int P) {
  __shared__ DiagLds Ld;
  const FacDesc& f = d[blockIdx.x % P];
  const int c = (int)(blockIdx.x / P);
  const int nb = (f.K + NB - 1) / NB;
  if (f.K <= NB || c >= nb - 1) return;
  xinv_tile(f.K, f.ld, nb - 1, nb - 1, c, f.F, f.Dinv, Ld);
}

int ldlh_factor_batch(const std::vector<FacDesc>& h, const FacDesc* d, int& epoch, hipStream_t s) {
  const int P = (int)h.size();
  if (P == 0) return 0;
  for (int p = 1; p < P; ++p)
    if (h[p].K > h[p - 1].K) return -1;   // problems must come largest first
  // problems still active in a launch are a prefix of the list
  auto active = [&](int minK) {
    int n = 0;
    while (n < P && h[n].K > minK) ++n;
    return n;
  };
  const int maxK = h[0].K;
  hipLaunchKernelGGL(k_clear_status_b, dim3(P), dim3(64), 0, s, d);
  hipLaunchKernelGGL(k_ldl_diag_b, dim3(P), dim3(256), 0, s, d);
  const int pp = active(NB);
  if (pp > 0) {
    ++epoch;
    hipLaunchKernelGGL(k_ldl_panel_b, dim3(pp * ((maxK - NB + SB - 1) / SB)), dim3(256), 0, s, d, pp, epoch);
  }
  for (int col = 0; col * NB + NB < maxK; ++col) {
    const int pc = active(col * NB + NB);
    int gy = 0;
    for (int p = 0; p < pc; ++p) gy = std::max(gy, upd_shape(h[p].K, col).grid);
    ++epoch;
    hipLaunchKernelGGL(k_ldl_update_b, dim3(pc * gy), dim3(256), 0, s, d, pc, col, epoch);
  }
  if (pp > 0) hipLaunchKernelGGL(k_ldl_xlast_b, dim3(pp * ((maxK + NB - 1) / NB - 1)), dim3(256), 0, s, d, pp);
  return (int)hipGetLastError();
}

__global__ void k_clear_status(int* status) {
  if (threadIdx.x < 3) status[threadIdx.x] = 0;
}

int ldlh_factor(Factor& f, hipStream_t s, const zd* H, const zd* S, double sigma) {
  const int K = f.K;
  const int qflag = f.ld / NB + 1;
  ShiftSrc src;   // first pass: H - sigma S (or F when H is null)
  src.H = H;
  src.S = S;
  src.sigma = sigma;
  const ShiftSrc none;
  hipLaunchKernelGGL(k_clear_status, dim3(1), dim3(64), 0, s, f.status);
  if (f.tr) f.tr->begin(TR_LDL_DIAG, s);
  hipLaunchKernelGGL(k_ldl_diag, dim3(1), dim3(256), 0, s, K, f.ld, 0, f.F, f.D, f.Dinv, f.status, src);
  if (f.tr) f.tr->end(s);
  static_assert(NB == SCHED_NB && SB == SCHED_SB, "task map and kernel disagree");
  if (K > NB) {
    const int epoch = ++f.epoch;
    if (f.tr) f.tr->begin(TR_LDL_PANEL, s);
    hipLaunchKernelGGL(k_ldl_panel, dim3((K - NB + SB - 1) / SB), dim3(256), 0, s, K, f.ld, f.F, f.D, f.Dinv, f.W,
                       f.status, f.pflag, qflag, epoch, src);
    if (f.tr) f.tr->end(s);
  }
  for (int col = 0; col * NB + NB < K; ++col) {
    const UpdShape shape = upd_shape(K, col);
    const int epoch = ++f.epoch;
    if (f.tr) f.tr->begin(TR_LDL_UPDATE, s);
    hipLaunchKernelGGL(k_ldl_update, dim3(shape.grid), dim3(256), 0, s, K, f.ld, col, f.F, f.W, f.D, f.Dinv, f.status,
                       f.pflag, qflag, epoch, col == 0 ? src : none);
    if (f.tr) f.tr->end(s);
  }
  if (K > NB) hipLaunchKernelGGL(k_ldl_xlast, dim3((K + NB - 1) / NB - 1), dim3(256), 0, s, K, f.ld, f.F, f.Dinv);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------ tail update
// Rows m..K-1 of M = H - sigma S have changed (r = K - m <= NB rows);
// rows < m and sigma are those of the last factorisation, whose L11, D1
// and X11 = L11^{-1} (X^H in the strict upper triangle) are still in F
// and D.  The factor is completed from row m, as the reference's LDLHF
// does from row Nmin (src/linalg.f90:316-455), in this storage:
//   Z = conj(L21) = D1^{-1} X11 conj(M21)^T, row by row: the forward pass
//       of the solve, z_i(c) = (b_i(c) + sum_{k<c} conj(F(k, c)) b_i(k))
//       / D_c with b_i = conj(M21(i, :))                 k_tail_fwd<NR>
//   S22 = M22 - L21 D1 L21^H                             k_tail_schur
//   L22 D2 L22^H = S22, X22 (one row: the pivot is S22)  k_ldl_diag at j0 = m
//   X21^H: the backward pass x_i = z_i + (X11^H z_i above the diagonal)
//       gives x_i = conj(L21 X11)(i, :), then
//       X^H(c, m + i) = -sum_{j <= i} X22^H(j, i) x_j(c)  k_tail_bwd<NR>, k_tail_xfin
// Each pass streams the m x m triangle of X11^H once per NR rows (HBM),
// with 8 r m^2 / 2 flops.  Scratch: the r rows of conj(M21) (Bt) and of
// Z (Zt), row-major with stride ld, in f.W; backward-pass partial sums in
// f.tp.
constexpr int TAIL_CHUNK = 128;   // columns of X11^H per backward-pass workgroup

// status[0] = negative pivots among D[0, m) (the tail's own are added by
// its diagonal factor), status[1] = status[2] = 0.
__global__ void __launch_bounds__(256) k_tail_status(int m, const double* __restrict__ D, int* __restrict__ status) {
  __shared__ int cnt[256];
  int n = 0;
  for (int p = threadIdx.x; p < m; p += blockDim.x) n += (D[p] < 0.0) ? 1 : 0;
  cnt[threadIdx.x] = n;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) cnt[threadIdx.x] += cnt[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    status[0] = cnt[0];
    status[1] = 0;
    status[2] = 0;
  }
}

// M(m + i, c) = H - sigma S for c <= m + i: conj of columns < m to Bt
// (row i), the rest into the lower triangle of F's tail block.
__global__ void __launch_bounds__(256) k_tail_form(int ld, int m, const zd* __restrict__ H, const zd* __restrict__ S,
                                                   double sigma, zd* __restrict__ F, zd* __restrict__ Bt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (c > m + i) return;
  const size_t o = (size_t)c * ld + m + i;
  const zd h = H[o], s = S[o];
  const zd v = mkz(fma(-sigma, s.x, h.x), fma(-sigma, s.y, h.y));
  if (c < m) Bt[(size_t)i * ld + c] = zconj(v);
  else F[o] = v;
}

// Forward pass for NR tail rows (row group blockIdx.y): one wave per
// column c < m, lanes over k < c (a contiguous read of column c of F),
// four loads of F in flight per lane.  Writes z_i(c) to Zt and L21(i, c)
// = conj(z_i(c)) to F.
template <int NR>
__global__ void __launch_bounds__(256) k_tail_fwd(int ld, int m, int r, const zd* __restrict__ F,
                                                  const double* __restrict__ D, const zd* __restrict__ Bt,
                                                  zd* __restrict__ Zt, zd* __restrict__ Fw) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= m) return;
  const int i0 = blockIdx.y * NR;
  const int nr = min(NR, r - i0);
  const int lane = threadIdx.x & 63;
  const zd* col = F + (size_t)c * ld;
  const zd* b = Bt + (size_t)i0 * ld;
  double ar[NR], ai[NR];
#pragma unroll
  for (int u = 0; u < NR; ++u) ar[u] = ai[u] = 0.0;
  int k = lane;
  for (; k + 192 < c; k += 256) {
    zd f[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) f[q] = col[k + 64 * q];
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      if (u < nr) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const zd bv = b[(size_t)u * ld + k + 64 * q];
          ar[u] = fma(f[q].x, bv.x, fma(f[q].y, bv.y, ar[u]));     // conj(f) b
          ai[u] = fma(f[q].x, bv.y, fma(-f[q].y, bv.x, ai[u]));
        }
      }
    }
  }
  for (; k < c; k += 64) {
    const zd f0 = col[k];
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      if (u < nr) {
        const zd bv = b[(size_t)u * ld + k];
        ar[u] = fma(f0.x, bv.x, fma(f0.y, bv.y, ar[u]));
        ai[u] = fma(f0.x, bv.y, fma(-f0.y, bv.x, ai[u]));
      }
    }
  }
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    for (int o = 32; o > 0; o >>= 1) {
      ar[u] += __shfl_xor(ar[u], o);
      ai[u] += __shfl_xor(ai[u], o);
    }
  }
  if (lane == 0) {
    const double inv = 1.0 / D[c];
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      if (u < nr) {
        const zd bc = b[(size_t)u * ld + c];
        const zd z = mkz((bc.x + ar[u]) * inv, (bc.y + ai[u]) * inv);
        Zt[(size_t)(i0 + u) * ld + c] = z;
        Fw[(size_t)c * ld + m + i0 + u] = zconj(z);
      }
    }
  }
}

// S22(i, j) -= sum_{c < m} conj(z_i(c)) D_c z_j(c) for i >= j: one
// workgroup per (i, j), fixed-order reduction.  A single tail row (r = 1)
// is its own pivot: D[m] and the status are written here.
__global__ void __launch_bounds__(256) k_tail_schur(int ld, int m, int r, const zd* __restrict__ Zt,
                                                    double* __restrict__ D, zd* __restrict__ F,
                                                    int* __restrict__ status) {
  __shared__ double sr[256], si[256];
  const int w = blockIdx.x;
  int i = (int)((sqrt(8.0 * w + 1.0) - 1.0) * 0.5);
  while ((i + 1) * (i + 2) / 2 <= w) ++i;
  while (i * (i + 1) / 2 > w) --i;
  const int j = w - i * (i + 1) / 2;
  const zd* zi = Zt + (size_t)i * ld;
  const zd* zj = Zt + (size_t)j * ld;
  double ar = 0.0, ai = 0.0;
  for (int c = threadIdx.x; c < m; c += blockDim.x) {
    const zd a = zi[c] * D[c], b = zj[c];
    ar = fma(a.x, b.x, fma(a.y, b.y, ar));      // conj(a) b
    ai = fma(a.x, b.y, fma(-a.y, b.x, ai));
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sr[threadIdx.x] += sr[threadIdx.x + o];
      si[threadIdx.x] += si[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    zd* o = F + (size_t)(m + j) * ld + m + i;
    const zd v = mkz(o->x - sr[0], o->y - si[0]);
    *o = v;
    if (r == 1) {
      D[m] = v.x;
      if (v.x < 0.0) atomicAdd(&status[0], 1);
      if (v.x == 0.0) atomicCAS(&status[1], 0, m + 1);
    }
  }
}

// Backward pass for NR tail rows: partial sums over chunk ch of
// x_i(c) = sum_{k > c} F(c, k) z_i(k), F(c, k) = F[k ld + c]; workgroup =
// 64 rows c x 4 column groups, eight loads of F in flight per lane,
// columns k <= c read and masked (in bounds: k < m).
template <int NR>
__global__ void __launch_bounds__(256) k_tail_bwd(int ld, int m, int r, int rp, const zd* __restrict__ F,
                                                  const zd* __restrict__ Zt, zd* __restrict__ part) {
  const int rb = blockIdx.x, ch = blockIdx.y;
  if ((ch + 1) * TAIL_CHUNK <= rb * 64 + 1) return;   // every column k of the chunk is <= every row c
  __shared__ zd red[4][64][NR];
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int c = rb * 64 + lane;
  const int i0 = blockIdx.z * NR;
  const int nr = min(NR, r - i0);
  const int k1 = min(m, (ch + 1) * TAIL_CHUNK);
  zd acc[NR];
#pragma unroll
  for (int u = 0; u < NR; ++u) acc[u] = mkz(0.0, 0.0);
  if (c < m) {
    int k = ch * TAIL_CHUNK + g;
    for (; k + 28 < k1; k += 32) {
      zd fv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) fv[q] = F[(size_t)(k + 4 * q) * ld + c];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const bool above = k + 4 * q > c;
#pragma unroll
        for (int u = 0; u < NR; ++u)
          if (u < nr) {
            const zd zv = Zt[(size_t)(i0 + u) * ld + k + 4 * q];
            if (above) acc[u] = zfma(fv[q], zv, acc[u]);
          }
      }
    }
    for (; k < k1; k += 4)
      if (k > c)
#pragma unroll
        for (int u = 0; u < NR; ++u)
          if (u < nr) acc[u] = zfma(F[(size_t)k * ld + c], Zt[(size_t)(i0 + u) * ld + k], acc[u]);
  }
#pragma unroll
  for (int u = 0; u < NR; ++u) red[g][lane][u] = acc[u];
  __syncthreads();
  if (g == 0 && c < m)
#pragma unroll
    for (int u = 0; u < NR; ++u)
      if (u < nr)
        part[((size_t)ch * rp + i0 + u) * ld + c] = red[0][lane][u] + red[1][lane][u] + red[2][lane][u] +
                                                    red[3][lane][u];
}

// x_i(c) = z_i(c) + the chunks (those k_tail_bwd wrote for c's row block),
// then X^H(c, m + i) = -(x_i(c) + sum_{j < i} X22^H(j, i) x_j(c)) with
// X22^H(j, i) = F[(m + i) ld + m + j].  One workgroup per 64 columns.
__global__ void __launch_bounds__(256) k_tail_xfin(int ld, int m, int r, int rp, int nch, const zd* __restrict__ Zt,
                                                   const zd* __restrict__ part, zd* __restrict__ F) {
  __shared__ zd X[NB][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lane;
  const bool in = c < m;
  const int rb = blockIdx.x;
  for (int i = wv; i < r; i += 4) {
    zd s = in ? Zt[(size_t)i * ld + c] : mkz(0.0, 0.0);
    for (int ch = 0; ch < nch; ++ch)
      if (in && (ch + 1) * TAIL_CHUNK > rb * 64 + 1) s = s + part[((size_t)ch * rp + i) * ld + c];
    X[i][lane] = s;
  }
  __syncthreads();
  for (int i = wv; i < r; i += 4) {
    zd x = X[i][lane];
    const zd* xh = F + (size_t)(m + i) * ld + m;
    for (int j = 0; j < i; ++j) x = zfma(xh[j], X[j][lane], x);
    if (in) F[(size_t)(m + i) * ld + c] = mkz(-x.x, -x.y);
  }
}

size_t tail_partial_entries(int ld) {
  return (size_t)((ld + TAIL_CHUNK - 1) / TAIL_CHUNK) * NB * (size_t)ld;
}

template <int NR>
static void tail_passes(const Ctx& c, Factor& f, int m, int r, int rp, int nch, zd* Bt, zd* Zt, hipStream_t s,
                        bool first) {
  const int ld = f.ld, ng = (r + NR - 1) / NR;
  if (first)
    hipLaunchKernelGGL(k_tail_fwd<NR>, dim3((m + 3) / 4, ng), dim3(256), 0, s, ld, m, r, f.F, f.D, Bt, Zt, f.F);
  else
    hipLaunchKernelGGL(k_tail_bwd<NR>, dim3((m + 63) / 64, nch, ng), dim3(256), 0, s, ld, m, r, rp, f.F, Zt, f.tp);
}

int ldlh_tail(const Ctx& c, Factor& f, double sigma, int m, hipStream_t s) {
  const int K = c.K, ld = f.ld, r = K - m;
  if (m < 1 || r < 0 || r > NB || !f.tp || ld != c.ld) return -1;
  const int rp = r;
  const int nch = (m + TAIL_CHUNK - 1) / TAIL_CHUNK;
  if ((size_t)nch * NB * ld > f.tp_cap) return -1;
  if (f.tr) f.tr->begin(TR_LDL_TAIL, s);
  hipLaunchKernelGGL(k_tail_status, dim3(1), dim3(256), 0, s, m, f.D, f.status);
  if (r > 0) {
    zd* Bt = f.W;
    zd* Zt = f.W + (size_t)NB * ld;
    hipLaunchKernelGGL(k_tail_form, dim3((K + 255) / 256, r), dim3(256), 0, s, ld, m, c.H, c.S, sigma, f.F, Bt);
    for (int pass = 0; pass < 2; ++pass) {
      const bool first = pass == 0;
      if (r == 1) tail_passes<1>(c, f, m, r, rp, nch, Bt, Zt, s, first);
      else if (r <= 4) tail_passes<4>(c, f, m, r, rp, nch, Bt, Zt, s, first);
      else tail_passes<8>(c, f, m, r, rp, nch, Bt, Zt, s, first);
      if (first) {
        hipLaunchKernelGGL(k_tail_schur, dim3(r * (r + 1) / 2), dim3(256), 0, s, ld, m, r, Zt, f.D, f.F, f.status);
        if (r > 1)
          hipLaunchKernelGGL(k_ldl_diag, dim3(1), dim3(256), 0, s, K, ld, m, f.F, f.D, f.Dinv, f.status, ShiftSrc());
      }
    }
    hipLaunchKernelGGL(k_tail_xfin, dim3((m + 63) / 64), dim3(256), 0, s, ld, m, r, rp, nch, Zt, f.tp, f.F);
  }
  if (f.tr) f.tr->end(s);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------ solve
// x = F^{-1} b = X^H D^{-1} X b with X = L^{-1} held as X^H in the strict
// upper triangle of F (unit diagonal): two passes over that triangle, each
// an HBM-bound GEMV with no dependency chain.
//   forward:  z_i = (b_i + sum_{j<i} conj(F(j, i)) b_j) / D_i, one wave per
//             column i (a contiguous read of the column);
//   backward: x_j = z_j + sum_{i>j} F(j, i) z_i, workgroups of 64 rows x
//             XS_CHUNK columns write partial sums that k_xsolve_reduce adds
//             in a fixed order.
constexpr int XS_CHUNK = 256;
static_assert(XS_CHUNK % 64 == 0, "chunks hold whole 64-row blocks");

__device__ __attribute__((always_inline)) inline void xsolve_fwd_body(int blk, int K, int ld,
                                                                      const zd* __restrict__ F,
                                                                      const double* __restrict__ D,
                                                                      const zd* __restrict__ b, zd* __restrict__ z) {
  const int i = blk * 4 + (threadIdx.x >> 6);
  if (i >= K) return;
  const int lane = threadIdx.x & 63;
  const zd* col = F + (size_t)i * ld;
  double ar = 0.0, ai = 0.0;
  int j = lane;
  for (; j + 192 < i; j += 256) {   // four columns' worth of loads in flight
    const zd f0 = col[j], f1 = col[j + 64], f2 = col[j + 128], f3 = col[j + 192];
    const zd b0 = b[j], b1 = b[j + 64], b2 = b[j + 128], b3 = b[j + 192];
    // conj(f) b
    ar = fma(f0.x, b0.x, fma(f0.y, b0.y, ar));
    ai = fma(f0.x, b0.y, fma(-f0.y, b0.x, ai));
    ar = fma(f1.x, b1.x, fma(f1.y, b1.y, ar));
    ai = fma(f1.x, b1.y, fma(-f1.y, b1.x, ai));
    ar = fma(f2.x, b2.x, fma(f2.y, b2.y, ar));
    ai = fma(f2.x, b2.y, fma(-f2.y, b2.x, ai));
    ar = fma(f3.x, b3.x, fma(f3.y, b3.y, ar));
    ai = fma(f3.x, b3.y, fma(-f3.y, b3.x, ai));
  }
  for (; j < i; j += 64) {
    const zd f0 = col[j], b0 = b[j];
    ar = fma(f0.x, b0.x, fma(f0.y, b0.y, ar));
    ai = fma(f0.x, b0.y, fma(-f0.y, b0.x, ai));
  }
  for (int o = 32; o > 0; o >>= 1) {
    ar += __shfl_xor(ar, o);
    ai += __shfl_xor(ai, o);
  }
  if (lane == 0) {
    const double inv = 1.0 / D[i];
    z[i] = mkz((b[i].x + ar) * inv, (b[i].y + ai) * inv);
  }
}
__global__ void __launch_bounds__(256) k_xsolve_fwd(int K, int ld, const zd* __restrict__ F,
                                                    const double* __restrict__ D, const zd* __restrict__ b,
                                                    zd* __restrict__ z) {
  xsolve_fwd_body((int)blockIdx.x, K, ld, F, D, b, z);
}

__device__ __attribute__((always_inline)) inline void xsolve_bwd_body(int rb, int ch, int K, int ld,
                                                                      const zd* __restrict__ F,
                                                                      const zd* __restrict__ z,
                                                                      zd* __restrict__ part) {
  if (ch < rb / (XS_CHUNK / 64)) return;   // the chunk's columns all lie left of these rows
  __shared__ zd red[4][64];
  const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int j = rb * 64 + r;
  const int c1 = min(K, (ch + 1) * XS_CHUNK);
  // eight columns' loads in flight per lane; columns i <= j (lower
  // triangle, in bounds) are read and masked
  zd acc0 = mkz(0.0, 0.0), acc1 = mkz(0.0, 0.0);
  if (j < K) {
    int i = ch * XS_CHUNK + g;
    for (; i + 28 < c1; i += 32) {
      zd fv[8], zv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        fv[u] = F[(size_t)(i + 4 * u) * ld + j];
        zv[u] = z[i + 4 * u];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const zd t = (i + 4 * u > j) ? zv[u] : mkz(0.0, 0.0);
        if (u & 1) acc1 = zfma(fv[u], t, acc1);
        else acc0 = zfma(fv[u], t, acc0);
      }
    }
    for (; i < c1; i += 4)
      if (i > j) acc0 = zfma(F[(size_t)i * ld + j], z[i], acc0);
  }
  red[g][r] = acc0 + acc1;
  __syncthreads();
  if (g == 0 && j < K) part[(size_t)ch * ld + j] = red[0][r] + red[1][r] + red[2][r] + red[3][r];
}
__global__ void __launch_bounds__(256) k_xsolve_bwd(int K, int ld, const zd* __restrict__ F,
                                                    const zd* __restrict__ z, zd* __restrict__ part) {
  xsolve_bwd_body((int)blockIdx.x, (int)blockIdx.y, K, ld, F, z, part);
}

__global__ void __launch_bounds__(256) k_xsolve_reduce(int K, int ld, int nch, const zd* __restrict__ part,
                                                       const zd* __restrict__ z, zd* __restrict__ x) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= K) return;
  zd s = z[j];
  for (int ch = j / XS_CHUNK; ch < nch; ++ch) s = s + part[(size_t)ch * ld + j];
  x[j] = s;
}

// The two passes: z (f.tvec) and the backward pass's partial sums (f.W,
// ld x nch <= ld x 2 NB); x = z + the partials' fixed-order sum is left to
// the caller (k_xsolve_reduce, or fused into the next kernel).
static int ldlh_passes(Factor& f, const zd* b, hipStream_t s, bool close_trace) {
  const int K = f.K;
  const int nrb = (K + 63) / 64, nch = (K + XS_CHUNK - 1) / XS_CHUNK;
  if (nch > 2 * NB) return -1;
  zd* z = f.tvec;
  if (f.tr) f.tr->begin(TR_TRSV_FWD, s);
  hipLaunchKernelGGL(k_xsolve_fwd, dim3((K + 3) / 4), dim3(256), 0, s, K, f.ld, f.F, f.D, b, z);
  if (f.tr) f.tr->end(s);
  if (f.tr) f.tr->begin(TR_TRSV_BWD, s);
  hipLaunchKernelGGL(k_xsolve_bwd, dim3(nrb, nch), dim3(256), 0, s, K, f.ld, f.F, z, f.W);
  if (f.tr && close_trace) f.tr->end(s);
  return (int)hipGetLastError();
}

// b and x distinct; f.tvec holds z, f.W the partial sums (ld x nch <= ld x 2 NB).
int ldlh_solve(Factor& f, const zd* b, zd* x, hipStream_t s) {
  const int K = f.K;
  const int nch = (K + XS_CHUNK - 1) / XS_CHUNK;
  int rc = ldlh_passes(f, b, s, false);
  if (rc) return rc;
  hipLaunchKernelGGL(k_xsolve_reduce, dim3((K + 255) / 256), dim3(256), 0, s, K, f.ld, nch, f.W, f.tvec, x);
  if (f.tr) f.tr->end(s);
  return (int)hipGetLastError();
}

// ------------------------------------------------------ vector kernels
// Block-wide reductions for a workgroup of up to 1024 threads (a multiple
// of 64): a butterfly within each wave, then the waves' results in index
// order through LDS (fixed order, two barriers).  sh: >= 16 doubles.
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ inline double block_sum(double v, double* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  for (int q = 0; q < nw; ++q) r += sh[q];
  __syncthreads();
  return r;
}
__device__ inline double block_max(double v, double* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_max(v);
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
  for (int q = 1; q < nw; ++q) r = fmax(r, sh[q]);
  __syncthreads();
  return r;
}

// One inverse-iteration bookkeeping step (GHEPIIS, src/linalg.f90:1892-1909):
// x /= max|Re|,|Im|; tc = x^H v / v^H v; out[0] = sqrt(2K) ||x - tc v||,
// out[1] = the scale used; out[4] keeps the previous out[0].
__global__ void __launch_bounds__(1024) k_invit_finish(int K, zd* __restrict__ x, const zd* __restrict__ v,
                                                       double* __restrict__ out) {
  __shared__ double sh[16];
  const int tid = threadIdx.x;
  double mx = 0.0;
  for (int i = tid; i < K; i += blockDim.x) mx = fmax(mx, fmax(fabs(x[i].x), fabs(x[i].y)));
  const double t1 = block_max(mx, sh);
  const double inv = 1.0 / t1;
  double ar = 0.0, ai = 0.0, rr = 0.0;
  for (int i = tid; i < K; i += blockDim.x) {
    const zd xi = x[i] * inv;
    x[i] = xi;
    const zd vi = v[i];
    ar += xi.x * vi.x + xi.y * vi.y;     // conj(x) v
    ai += xi.x * vi.y - xi.y * vi.x;
    rr += vi.x * vi.x + vi.y * vi.y;
  }
  ar = block_sum(ar, sh);
  ai = block_sum(ai, sh);
  rr = block_sum(rr, sh);
  const zd tc = mkz(ar / rr, ai / rr);
  double nd = 0.0;
  for (int i = tid; i < K; i += blockDim.x) {
    const zd d = x[i] - tc * v[i];
    nd += d.x * d.x + d.y * d.y;
  }
  nd = block_sum(nd, sh);
  if (tid == 0) {
    out[4] = out[0];
    out[0] = sqrt(2.0 * K) * sqrt(nd);
    out[1] = t1;
  }
}

// k_xsolve_reduce and k_invit_finish in one workgroup, for K <= 1024
// IF_REG: x_j = z_j + the backward pass's partial sums of the chunks right
// of j (k_xsolve_reduce's order), kept in registers through the same
// passes and block reductions as k_invit_finish (bitwise the same results).
constexpr int IF_REG = 8;

__device__ __attribute__((always_inline)) inline void invit_finish_red_body(int K, int ld, int nch,
                                                                            const zd* __restrict__ part,
                                                                            const zd* __restrict__ z,
                                                                            const zd* __restrict__ v,
                                                                            zd* __restrict__ x,
                                                                            double* __restrict__ out) {
  __shared__ double sh[16];
  const int tid = threadIdx.x;
  zd xr[IF_REG], vr[IF_REG];
  double mx = 0.0;
#pragma unroll
  for (int u = 0; u < IF_REG; ++u) {
    const int j = tid + 1024 * u;
    xr[u] = vr[u] = mkz(0.0, 0.0);
    if (j < K) {
      zd a = z[j];
      // the partials in k_xsolve_reduce's order, eight loads in flight
      for (int ch0 = j / XS_CHUNK; ch0 < nch; ch0 += 8) {
        zd t[8];
#pragma unroll
        for (int w = 0; w < 8; ++w) t[w] = ld_or0(part, (size_t)(ch0 + w) * ld + j, ch0 + w < nch);
#pragma unroll
        for (int w = 0; w < 8; ++w)
          if (ch0 + w < nch) a = a + t[w];
      }
      xr[u] = a;
      vr[u] = v[j];
      mx = fmax(mx, fmax(fabs(a.x), fabs(a.y)));
    }
  }
  const double t1 = block_max(mx, sh);
  const double inv = 1.0 / t1;
  double ar = 0.0, ai = 0.0, rr = 0.0;
#pragma unroll
  for (int u = 0; u < IF_REG; ++u) {
    const int j = tid + 1024 * u;
    if (j < K) {
      const zd xi = xr[u] * inv;
      xr[u] = xi;
      x[j] = xi;
      const zd vi = vr[u];
      ar += xi.x * vi.x + xi.y * vi.y;     // conj(x) v
      ai += xi.x * vi.y - xi.y * vi.x;
      rr += vi.x * vi.x + vi.y * vi.y;
    }
  }
  ar = block_sum(ar, sh);
  ai = block_sum(ai, sh);
  rr = block_sum(rr, sh);
  const zd tc = mkz(ar / rr, ai / rr);
  double nd = 0.0;
#pragma unroll
  for (int u = 0; u < IF_REG; ++u) {
    const int j = tid + 1024 * u;
    if (j < K) {
      const zd d = xr[u] - tc * vr[u];
      nd += d.x * d.x + d.y * d.y;
    }
  }
  nd = block_sum(nd, sh);
  if (tid == 0) {
    out[4] = out[0];
    out[0] = sqrt(2.0 * K) * sqrt(nd);
    out[1] = t1;
  }
}
__global__ void __launch_bounds__(1024) k_invit_finish_red(int K, int ld, int nch, const zd* __restrict__ part,
                                                           const zd* __restrict__ z, const zd* __restrict__ v,
                                                           zd* __restrict__ x, double* __restrict__ out) {
  invit_finish_red_body(K, ld, nch, part, z, v, x, out);
}

// w = S^H v with MHMV's formula (src/linalg.f90:1037-1044):
// w_i = sum_j conj(S(j, i)) v_j over the whole stored column i (contiguous),
// one wave per column, no partial sums.
__device__ __attribute__((always_inline)) inline void mhmv_body(int blk, int K, int ld, const zd* __restrict__ A,
                                                                const zd* __restrict__ v, zd* __restrict__ w) {
  const int i = blk * 4 + (threadIdx.x >> 6);
  if (i >= K) return;
  const int lane = threadIdx.x & 63;
  const zd* col = A + (size_t)i * ld;
  double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
  int j = lane;
  for (; j + 192 < K; j += 256) {   // four loads in flight, two accumulator pairs
    const zd f0 = col[j], f1 = col[j + 64], f2 = col[j + 128], f3 = col[j + 192];
    const zd b0 = v[j], b1 = v[j + 64], b2 = v[j + 128], b3 = v[j + 192];
    ar = fma(f0.x, b0.x, fma(f0.y, b0.y, ar));
    ai = fma(f0.x, b0.y, fma(-f0.y, b0.x, ai));
    br = fma(f1.x, b1.x, fma(f1.y, b1.y, br));
    bi = fma(f1.x, b1.y, fma(-f1.y, b1.x, bi));
    ar = fma(f2.x, b2.x, fma(f2.y, b2.y, ar));
    ai = fma(f2.x, b2.y, fma(-f2.y, b2.x, ai));
    br = fma(f3.x, b3.x, fma(f3.y, b3.y, br));
    bi = fma(f3.x, b3.y, fma(-f3.y, b3.x, bi));
  }
  for (; j < K; j += 64) {
    const zd f0 = col[j], b0 = v[j];
    ar = fma(f0.x, b0.x, fma(f0.y, b0.y, ar));
    ai = fma(f0.x, b0.y, fma(-f0.y, b0.x, ai));
  }
  ar += br;
  ai += bi;
  for (int o = 32; o > 0; o >>= 1) {
    ar += __shfl_xor(ar, o);
    ai += __shfl_xor(ai, o);
  }
  if (lane == 0) w[i] = mkz(ar, ai);
}
__global__ void __launch_bounds__(256) k_mhmv(int K, int ld, const zd* __restrict__ A, const zd* __restrict__ v,
                                              zd* __restrict__ w) {
  mhmv_body((int)blockIdx.x, K, ld, A, v, w);
}

__global__ void k_scale(int K, zd* __restrict__ x, const double* __restrict__ t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < K) x[i] = x[i] * (1.0 / sqrt(*t));
}

// One inverse-iteration step: x = F^{-1} (S v), then the GHEPIIS
// normalisation and convergence measure (written to f.scal[0..1]).
// S v is MHMV's column form (k_mhmv, into f.vec); the solve's partial
// sums use f.W, and for K <= 1024 IF_REG their reduction is fused into the
// bookkeeping kernel.
int invit_iteration(const Ctx& c, Factor& f, const zd* v, zd* x, hipStream_t s) {
  const int K = c.K;
  zd* w = f.vec;
  hipLaunchKernelGGL(k_mhmv, dim3((K + 3) / 4), dim3(256), 0, s, K, c.ld, c.S, v, w);
  int rc;
  if (K <= 1024 * IF_REG) {
    rc = ldlh_passes(f, w, s, true);
    if (rc) return rc;
    const int nch = (K + XS_CHUNK - 1) / XS_CHUNK;
    hipLaunchKernelGGL(k_invit_finish_red, dim3(1), dim3(1024), 0, s, K, f.ld, nch, f.W, f.tvec, v, x, f.scal);
    return (int)hipGetLastError();
  }
  rc = ldlh_solve(f, w, x, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_invit_finish, dim3(1), dim3(1024), 0, s, K, x, v, f.scal);
  return (int)hipGetLastError();
}

// Rayleigh quotient pieces: scal[2] = x^H S x, scal[3] = x^H (H - sigma S) x.
// Quadratic forms of the Rayleigh quotient (VMMHMV, src/linalg.f90:
// 1180-1255) from the lower triangles of S and M = H - sigma S in one pass:
// x^H A x = sum_i A_ii |x_i|^2 + 2 Re sum_{i>j} conj(x_i) A_ij x_j.
// Workgroup = 64 rows x QF_CHUNK columns (below the diagonal); each lane
// sums A_ij x_j along its row, then Re(conj(x_i) .) is reduced to two
// doubles per workgroup; k_quad_finish adds those and the diagonal in a
// fixed order: out[2] = x^H S x, out[3] = x^H M x.
constexpr int QF_CHUNK = 128;

__device__ __attribute__((always_inline)) inline void quad_part_body(int rb, int ch, int nch, int K, int ld,
                                                                     const zd* __restrict__ H,
                                                                     const zd* __restrict__ S, double sigma,
                                                                     const zd* __restrict__ x,
                                                                     double* __restrict__ part) {
  __shared__ double red[2][4];
  double* out = part + 2 * ((size_t)rb * nch + ch);
  const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int i = rb * 64 + r;
  const int c0 = ch * QF_CHUNK;
  if (c0 > rb * 64 + 63) {   // chunk entirely above the block's rows
    if (threadIdx.x == 0) out[0] = out[1] = 0.0;
    return;
  }
  const int c1 = min(min(K, c0 + QF_CHUNK), rb * 64 + 64);
  zd ts = mkz(0.0, 0.0), tm = mkz(0.0, 0.0);
  if (i < K)
    for (int j = c0 + g; j < c1; j += 4)
      if (j < i) {
        const size_t o = (size_t)j * ld + i;
        const zd sv = S[o], hv = H[o], xj = x[j];
        const zd mv = mkz(fma(-sigma, sv.x, hv.x), fma(-sigma, sv.y, hv.y));
        ts = zfma(sv, xj, ts);
        tm = zfma(mv, xj, tm);
      }
  double qs = 0.0, qm = 0.0;
  if (i < K) {
    const zd xi = x[i];
    qs = 2.0 * (xi.x * ts.x + xi.y * ts.y);
    qm = 2.0 * (xi.x * tm.x + xi.y * tm.y);
  }
  for (int o = 32; o > 0; o >>= 1) {
    qs += __shfl_xor(qs, o);
    qm += __shfl_xor(qm, o);
  }
  if (r == 0) {
    red[0][g] = qs;
    red[1][g] = qm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    out[1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}
__global__ void __launch_bounds__(256) k_quad_part(int K, int ld, const zd* __restrict__ H,
                                                   const zd* __restrict__ S, double sigma,
                                                   const zd* __restrict__ x, double* __restrict__ part) {
  quad_part_body((int)blockIdx.x, (int)blockIdx.y, (int)gridDim.y, K, ld, H, S, sigma, x, part);
}

__device__ __attribute__((always_inline)) inline void quad_finish_body(int K, int ld, const zd* __restrict__ H,
                                                                       const zd* __restrict__ S, double sigma,
                                                                       const zd* __restrict__ x, int nparts,
                                                                       const double* __restrict__ part,
                                                                       double* __restrict__ out) {
  __shared__ double sh[16];
  double qs = 0.0, qm = 0.0;
  for (int i = threadIdx.x; i < K; i += blockDim.x) {
    const size_t o = (size_t)i * ld + i;
    const double a2 = x[i].x * x[i].x + x[i].y * x[i].y;
    qs = fma(S[o].x, a2, qs);
    qm = fma(fma(-sigma, S[o].x, H[o].x), a2, qm);
  }
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) {
    qs += part[2 * p];
    qm += part[2 * p + 1];
  }
  qs = block_sum(qs, sh);
  qm = block_sum(qm, sh);
  if (threadIdx.x == 0) {
    out[2] = qs;
    out[3] = qm;
  }
}
__global__ void __launch_bounds__(1024) k_quad_finish(int K, int ld, const zd* __restrict__ H,
                                                      const zd* __restrict__ S, double sigma,
                                                      const zd* __restrict__ x, int nparts,
                                                      const double* __restrict__ part, double* __restrict__ out) {
  quad_finish_body(K, ld, H, S, sigma, x, nparts, part, out);
}

// scal[2] = x^H S x, scal[3] = x^H (H - sigma S) x; partials in f.W.
int rayleigh(const Ctx& c, Factor& f, const zd* x, double sigma, hipStream_t s) {
  const int nrb = (c.K + 63) / 64, nch = (c.K + QF_CHUNK - 1) / QF_CHUNK;
  double* part = reinterpret_cast<double*>(f.W);
  hipLaunchKernelGGL(k_quad_part, dim3(nrb, nch), dim3(256), 0, s, c.K, c.ld, c.H, c.S, sigma, x, part);
  hipLaunchKernelGGL(k_quad_finish, dim3(1), dim3(1024), 0, s, c.K, c.ld, c.H, c.S, sigma, x, nrb * nch, part,
                     f.scal);
  return (int)hipGetLastError();
}

int normalise(zd* x, int K, const double* t, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3((K + 255) / 256), dim3(256), 0, s, K, x, t);
  return (int)hipGetLastError();
}

// ---------------------------------------------- batched inverse iteration
// The inverse iterations of a batch's problems in shared launches, each
// kernel's per-problem work as in the single-problem launches above:
// workgroup id of a one-dimensional grid serves problem id % P, block
// id / P (blocks beyond a problem's own count return), and the kernels
// of one workgroup per problem take blockIdx.x as the problem.  Iteration
// it reads ev[it & 1] and writes ev[(it + 1) & 1].
__global__ void __launch_bounds__(256) k_vinit_b(const InvDesc* __restrict__ d, int P) {
  const InvDesc& q = d[blockIdx.x % P];
  const int i = (int)(blockIdx.x / P) * 256 + threadIdx.x;
  if (i < q.K) q.ev[0][i] = q.warm ? q.ev[2][i] : mkz(1.0, 0.0);
}
__global__ void __launch_bounds__(256) k_mhmv_b(const InvDesc* __restrict__ d, int P, int it) {
  const InvDesc& q = d[blockIdx.x % P];
  const int blk = (int)(blockIdx.x / P);
  if (blk * 4 >= q.K) return;
  mhmv_body(blk, q.K, q.ld, q.S, q.ev[it & 1], q.w);
}
__global__ void __launch_bounds__(256) k_xsolve_fwd_b(const InvDesc* __restrict__ d, int P) {
  const InvDesc& q = d[blockIdx.x % P];
  const int blk = (int)(blockIdx.x / P);
  if (blk * 4 >= q.K) return;
  xsolve_fwd_body(blk, q.K, q.ld, q.F, q.D, q.w, q.z);
}
__global__ void __launch_bounds__(256) k_xsolve_bwd_b(const InvDesc* __restrict__ d, int P) {
  const InvDesc& q = d[blockIdx.x % P];
  const int blk = (int)(blockIdx.x / P);
  const int nrb = (q.K + 63) / 64, nch = (q.K + XS_CHUNK - 1) / XS_CHUNK;
  if (blk >= nrb * nch) return;
  xsolve_bwd_body(blk % nrb, blk / nrb, q.K, q.ld, q.F, q.z, q.part);
}
__global__ void __launch_bounds__(1024) k_invit_finish_red_b(const InvDesc* __restrict__ d, int it) {
  const InvDesc& q = d[blockIdx.x];
  const int nch = (q.K + XS_CHUNK - 1) / XS_CHUNK;
  invit_finish_red_body(q.K, q.ld, nch, q.part, q.z, q.ev[it & 1], q.ev[(it + 1) & 1], q.scal);
}
__global__ void __launch_bounds__(256) k_quad_part_b(const InvDesc* __restrict__ d, int P, int it) {
  const InvDesc& q = d[blockIdx.x % P];
  const int blk = (int)(blockIdx.x / P);
  const int nrb = (q.K + 63) / 64, nch = (q.K + QF_CHUNK - 1) / QF_CHUNK;
  if (blk >= nrb * nch) return;
  // the single-problem launch's grid is (nrb, nch) and its partials are
  // stored by (rb, ch) row-major, which the finishing sum follows
  quad_part_body(blk / nch, blk % nch, nch, q.K, q.ld, q.H, q.S, q.sigma, q.ev[it & 1],
                 reinterpret_cast<double*>(q.part));
}
__global__ void __launch_bounds__(1024) k_quad_finish_b(const InvDesc* __restrict__ d, int it) {
  const InvDesc& q = d[blockIdx.x];
  const int nrb = (q.K + 63) / 64, nch = (q.K + QF_CHUNK - 1) / QF_CHUNK;
  quad_finish_body(q.K, q.ld, q.H, q.S, q.sigma, q.ev[it & 1], nrb * nch, reinterpret_cast<const double*>(q.part),
                   q.scal);
}
// normalise x^H S x = 1, report the scalars and the factorisation status
// into the problem's pinned host array (the single-problem k_report's
// layout: rep[0..4] = scal[0..4], rep[5..7] = status[0..2], rep[8] = E),
// keep E in scal[8], and copy the vector to ev[2]
__global__ void __launch_bounds__(256) k_invit_tail_b(const InvDesc* __restrict__ d, int P, int it) {
  const InvDesc& q = d[blockIdx.x % P];
  const int i = (int)(blockIdx.x / P) * 256 + threadIdx.x;
  zd* x = q.ev[it & 1];
  if (i < q.K) {
    const zd v = x[i] * (1.0 / sqrt(q.scal[2]));
    x[i] = v;
    q.ev[2][i] = v;
  }
  if (blockIdx.x / P == 0 && threadIdx.x == 0) {
    const double E = q.scal[3] / q.scal[2] + q.sigma;
    for (int k = 0; k < 5; ++k) __hip_atomic_store(&q.rep[k], q.scal[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    for (int k = 0; k < 3; ++k)
      __hip_atomic_store(&q.rep[5 + k], (double)q.status[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&q.rep[8], E, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
  }
}

int invit_batch(const std::vector<InvDesc>& h, const InvDesc* d, int nq, hipStream_t s) {
  const int P = (int)h.size();
  if (P == 0) return 0;
  int maxK = 0;
  for (const InvDesc& q : h) maxK = std::max(maxK, q.K);
  if (maxK > 1024 * IF_REG) return -1;   // the fused reduction keeps K <= 1024 IF_REG in registers
  const int nv = (maxK + 255) / 256, nc = (maxK + 3) / 4;
  const int nxb = ((maxK + 63) / 64) * ((maxK + XS_CHUNK - 1) / XS_CHUNK);
  const int nqb = ((maxK + 63) / 64) * ((maxK + QF_CHUNK - 1) / QF_CHUNK);
  hipLaunchKernelGGL(k_vinit_b, dim3(P * nv), dim3(256), 0, s, d, P);
  for (int it = 0; it < nq; ++it) {
    hipLaunchKernelGGL(k_mhmv_b, dim3(P * nc), dim3(256), 0, s, d, P, it);
    hipLaunchKernelGGL(k_xsolve_fwd_b, dim3(P * nc), dim3(256), 0, s, d, P);
    hipLaunchKernelGGL(k_xsolve_bwd_b, dim3(P * nxb), dim3(256), 0, s, d, P);
    hipLaunchKernelGGL(k_invit_finish_red_b, dim3(P), dim3(1024), 0, s, d, it);
  }
  hipLaunchKernelGGL(k_quad_part_b, dim3(P * nqb), dim3(256), 0, s, d, P, nq);
  hipLaunchKernelGGL(k_quad_finish_b, dim3(P), dim3(1024), 0, s, d, nq);
  hipLaunchKernelGGL(k_invit_tail_b, dim3(P * nv), dim3(256), 0, s, d, P, nq);
  return (int)hipGetLastError();
}

}  // namespace ecg
