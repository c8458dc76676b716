// Dense Hermitian linear algebra of the eigen step, FP64 complex:
//   * F = H - sigma S (lower triangle)
//   * blocked right-looking LDL^H of F without pivoting (the reference's
//     LDLHF, src/linalg.f90:316-455, reorganised in 64-wide blocks), with
//     the negative-pivot count (Sylvester inertia) and zero-pivot detection
//   * the solve x = F^{-1} b through L, D, L^H (LDLHS, src/linalg.f90:618-775)
//   * y = A x and the quadratic forms used by inverse iteration
//     (MHMV, VMMHMV, src/linalg.f90:1019-1107, 1180-1255)
// L overwrites the strict lower triangle of F; L^H is mirrored into the
// strict upper triangle so both solves read memory along columns.
#include "ecg_internal.hpp"

namespace ecg {

// ---------------------------------------------------------------- shift
__global__ void k_form_shifted(int K, int ld, const zd* __restrict__ H, const zd* __restrict__ S,
                               double sigma, zd* __restrict__ F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= K || i < j) return;
  const size_t o = (size_t)j * ld + i;
  const zd h = H[o], s = S[o];
  F[o] = mkz(fma(-sigma, s.x, h.x), fma(-sigma, s.y, h.y));
}

int form_shifted(const Ctx& c, Factor& f, double sigma, hipStream_t s) {
  hipLaunchKernelGGL(k_form_shifted, dim3((c.K + 255) / 256, c.K), dim3(256), 0, s, c.K, c.ld, c.H, c.S,
                     sigma, f.F);
  return (int)hipGetLastError();
}

// --------------------------------------------------- diagonal block LDL^H
// One workgroup of 256 threads factors the jb x jb diagonal block at
// (j0, j0), held in one 64 x 64 LDS array, blocked by 16:
//   for each 16-wide sub-block s: wave 0 factors the 16 x 16 diagonal piece
//   and inverts its unit-lower factor (wave-synchronous, no workgroup
//   barriers); all waves then form the sub-panel (W = A X_ss^H, L = W D^-1)
//   and update the trailing part;
//   afterwards wave s builds block column s of X = L^{-1}:
//   X_ts = -X_tt sum_{u=s}^{t-1} L_tu X_us.
// LDS use: lower triangle = the block, then L (pivots on the diagonal while
// factoring); strict upper triangle = X transposed (X(i, j), i > j, at
// A[j][i]; X has a unit diagonal), and while factoring, block row 0 of the
// upper triangle holds the current sub-panel W transposed.
// Outputs: unit-lower L (strict part, and L^H mirrored into the upper
// triangle of F), D, and X and X^H (column-major 64 x 64 each, used by the
// panel solve and the triangular solves).
constexpr int SB = 16;   // sub-block width

struct DiagLds {
  zd A[NB][NB + 1];
  double d[NB];
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __attribute__((always_inline)) inline void ldl_diag_block(int K, int ld, int j0, zd* __restrict__ F,
                                                                     double* __restrict__ D,
                                                                     zd* __restrict__ Dinv,
                                                                     int* __restrict__ status, DiagLds& L,
                                                                     bool a_in_lds) {
  const int jb = min(NB, K - j0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (!a_in_lds) {
    for (int e = tid; e < NB * NB; e += blockDim.x) {
      const int j = e >> 6, i = e & (NB - 1);
      L.A[i][j] = (i < jb && j < jb && i >= j) ? F[(size_t)(j0 + j) * ld + j0 + i] : mkz(0.0, 0.0);
    }
  }
  __syncthreads();
  const int nsub = (jb + SB - 1) / SB;
  for (int s = 0; s < nsub; ++s) {
    const int c = s * SB;
    const int sb = min(SB, jb - c);
    if (wave == 0) {
      // 16 x 16 LDL^H of A[c.., c..]
      for (int p = 0; p < sb; ++p) {
        const double d = L.A[c + p][c + p].x;
        const double rd = (d != 0.0) ? 1.0 / d : 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int e = lane + 64 * q;
          const int i = e >> 4, j = e & (SB - 1);
          if (j > p && j <= i && i < sb)
            L.A[c + i][c + j] = zfms(L.A[c + i][c + p] * rd, zconj(L.A[c + j][c + p]), L.A[c + i][c + j]);
        }
        wave_sync();
        if (lane > p && lane < sb) L.A[c + lane][c + p] = L.A[c + lane][c + p] * rd;
        if (lane == 0) {
          L.d[c + p] = d;
          if (d < 0.0) atomicAdd(&status[0], 1);
          if (d == 0.0) atomicCAS(&status[1], 0, j0 + c + p + 1);
        }
        wave_sync();
      }
      // X_ss = inverse of the unit-lower piece, by row operations; X_ss(i, j)
      // (i > j) lives at A[c + j][c + i], which starts at zero
      for (int p = 0; p < sb; ++p) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int e = lane + 64 * q;
          const int i = e >> 4, j = e & (SB - 1);
          if (i > p && i < sb && j <= p) {
            const zd a = L.A[c + j][c + p];
            const zd xpj = mkz(j == p ? 1.0 : a.x, j == p ? 0.0 : a.y);
            L.A[c + j][c + i] = zfms(L.A[c + i][c + p], xpj, L.A[c + j][c + i]);
          }
        }
        wave_sync();
      }
    }
    __syncthreads();
    // sub-panel: W(ii, j) = sum_{p <= j} A(c+sb+ii, c+p) conj(X_ss(j, p)), kept at A[j][SB + ii]
    const int rows = jb - c - sb;
    for (int e = tid; e < rows * SB; e += blockDim.x) {
      const int ii = e >> 4, j = e & (SB - 1);
      const int i = c + sb + ii;
      if (j < sb) {
        zd w = L.A[i][c + j];
        for (int p = 0; p < j; ++p) w = zfma(L.A[i][c + p], zconj(L.A[c + p][c + j]), w);
        L.A[j][SB + ii] = w;
      }
    }
    __syncthreads();
    for (int e = tid; e < rows * SB; e += blockDim.x) {
      const int ii = e >> 4, j = e & (SB - 1);
      if (j < sb) L.A[c + sb + ii][c + j] = L.A[j][SB + ii] * (1.0 / L.d[c + j]);
    }
    __syncthreads();
    // trailing update: A(i, j) -= sum_p W(i, p) conj(L(j, c+p)), c+sb <= j <= i < jb
    for (int e = tid; e < rows * rows; e += blockDim.x) {
      const int ii = e / rows, jj = e - ii * rows;
      if (jj <= ii) {
        const int i = c + sb + ii, j = c + sb + jj;
        zd acc = L.A[i][j];
        for (int p = 0; p < sb; ++p) acc = zfms(L.A[p][SB + ii], zconj(L.A[j][c + p]), acc);
        L.A[i][j] = acc;
      }
    }
    __syncthreads();
  }
  // X = L^{-1}: wave s fills the strict upper block row s (X block column s
  // transposed).  Only wave s writes block row s, and it reads block row s,
  // the lower triangle and the diagonal blocks' upper parts.
  if (wave < nsub) {
    const int s = wave;
    for (int t = s + 1; t < nsub; ++t) {
      zd T[4];
      // T(i, j) = sum_{k = m}^{t*SB-1} L(t*SB+i, k) X(k, m), m = s*SB + j (X(m, m) = 1)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = lane + 64 * q;
        const int i = e >> 4, j = e & (SB - 1);
        const int m = s * SB + j, row = t * SB + i;
        zd acc = L.A[row][m];
        for (int k = m + 1; k < t * SB; ++k) acc = zfma(L.A[row][k], L.A[m][k], acc);
        T[q] = acc;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = lane + 64 * q;
        L.A[s * SB + (e & (SB - 1))][t * SB + (e >> 4)] = T[q];
      }
      wave_sync();
      // X_ts(i, j) = -(T(i, j) + sum_{k < i} X_tt(i, k) T(k, j))
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = lane + 64 * q;
        const int i = e >> 4, j = e & (SB - 1);
        const int m = s * SB + j;
        zd acc = L.A[m][t * SB + i];
        for (int k = 0; k < i; ++k) acc = zfma(L.A[t * SB + k][t * SB + i], L.A[m][t * SB + k], acc);
        T[q] = mkz(-acc.x, -acc.y);
      }
      wave_sync();
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = lane + 64 * q;
        L.A[s * SB + (e & (SB - 1))][t * SB + (e >> 4)] = T[q];
      }
      wave_sync();
    }
  }
  __syncthreads();
  // outputs
  zd* Xg = Dinv + (size_t)(j0 / NB) * 2 * NB * NB;
  zd* XHg = Xg + NB * NB;
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int j = e >> 6, i = e & (NB - 1);       // column-major walk: coalesced writes
    const bool in = (i < jb && j < jb);
    if (in && i > j) F[(size_t)(j0 + j) * ld + j0 + i] = L.A[i][j];
    const zd xu = L.A[j][i], xl = L.A[i][j];
    const double dg = (in && i == j) ? 1.0 : 0.0;
    Xg[(size_t)j * NB + i] = (in && i > j) ? xu : mkz(dg, 0.0);
    XHg[(size_t)j * NB + i] = (in && i < j) ? zconj(xl) : mkz(dg, 0.0);
  }
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int i = e >> 6, j = e & (NB - 1);       // L^H into the upper triangle: row-major walk
    if (i < jb && j < jb && i > j) F[(size_t)(j0 + i) * ld + j0 + j] = zconj(L.A[i][j]);
  }
  for (int p = tid; p < jb; p += blockDim.x) D[j0 + p] = L.d[p];
}

__global__ void __launch_bounds__(256) k_ldl_diag(int K, int ld, int j0, zd* __restrict__ F,
                                                  double* __restrict__ D, zd* __restrict__ Dinv,
                                                  int* __restrict__ status) {
  __shared__ DiagLds L;
  ldl_diag_block(K, ld, j0, F, D, Dinv, status, L, false);
}

// ----------------------------------------------------- MFMA 64x64 tile
// acc(r, c) = sum_p A(r, p) conj(B(c, p)), r, c < 64, p < kd, with
// A(r, p) = Ap[p * lda + r] (r < ra) and B(c, p) = Bp[p * ldb + c] (c < rb).
// 4 waves, each a 32 x 32 quarter as 2 x 2 tiles of v_mfma_f64_16x16x4f64,
// four real products per complex one.  Accumulator v of a lane holds row
// 4*v + lane/16, column lane%16 of its 16x16 tile (measured lane layout of
// the instruction on gfx950).
typedef double dbl4 __attribute__((ext_vector_type(4)));

struct Tile64 {
  dbl4 re[2][2], im[2][2];
};

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
  // all operand fragments are loaded up front (16 k-steps x 4 complex per
  // lane) so the global loads are in flight together
  zd fa[NB / 4][2], fb[NB / 4][2];
#pragma unroll
  for (int s = 0; s < NB / 4; ++s) {
    const int p = 4 * s + lk;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int r = rt + a * 16 + lr;
      fa[s][a] = (p < kd && r < ra) ? Ap[(size_t)p * lda + r] : mkz(0.0, 0.0);
      const int c = ct + a * 16 + lr;
      fb[s][a] = (p < kd && c < rb) ? Bp[(size_t)p * ldb + c] : mkz(0.0, 0.0);
    }
  }
#pragma unroll
  for (int s = 0; s < NB / 4; ++s) {
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
// Row block i0..i0+ib-1 below the diagonal block: W = A X^H (so that
// W L_JJ^H = A), L = W D^{-1}; L goes to F (and L^H to its upper
// triangle), W to the panel scratch used by the trailing update.
__global__ void __launch_bounds__(256) k_ldl_panel(int K, int ld, int j0, zd* __restrict__ F,
                                                   const double* __restrict__ D,
                                                   const zd* __restrict__ Dinv, zd* __restrict__ W) {
  const int jb = min(NB, K - j0);
  const int i0 = j0 + jb + blockIdx.x * NB;
  const int ib = min(NB, K - i0);
  Tile64 t;
  mfma_tile_abh(F + (size_t)j0 * ld + i0, ld, ib, Dinv + (size_t)(j0 / NB) * 2 * NB * NB, NB, jb, jb, t);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        tile_pos(a, q, v, r, c);
        if (r < ib && c < jb) {
          const zd w = mkz(t.re[a][q][v], t.im[a][q][v]);
          const zd l = w * (1.0 / D[j0 + c]);
          W[(size_t)c * ld + i0 + r] = w;
          F[(size_t)(j0 + c) * ld + i0 + r] = l;
          F[(size_t)(i0 + r) * ld + j0 + c] = zconj(l);
        }
      }
}

// -------------------------------------------------------- trailing update
// Lower block pairs (I >= Kb) of the trailing matrix: C_IK -= W_I L_K^H.
// The workgroup that updates the first trailing diagonal block then
// factors it (look-ahead), so the next block column needs no separate
// diagonal launch.
__global__ void __launch_bounds__(256) k_ldl_update(int K, int ld, int j0, zd* __restrict__ F,
                                                    const zd* __restrict__ W, double* __restrict__ D,
                                                    zd* __restrict__ Dinv, int* __restrict__ status) {
  __shared__ DiagLds Ld;
  const int jb = min(NB, K - j0);
  const int t0 = (j0 + jb) / NB;
  const int b = blockIdx.x;
  int bi = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= b) ++bi;
  while (bi * (bi + 1) / 2 > b) --bi;
  const int bk = b - bi * (bi + 1) / 2;
  const int I0 = (t0 + bi) * NB, K0 = (t0 + bk) * NB;
  const int ib = min(NB, K - I0), kb = min(NB, K - K0);
  Tile64 t;
  mfma_tile_abh(W + I0, ld, ib, F + (size_t)j0 * ld + K0, ld, kb, jb, t);
  const bool lookahead = (bi == 0 && bk == 0);
  if (lookahead)
    for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) Ld.A[e >> 6][e & (NB - 1)] = mkz(0.0, 0.0);
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        tile_pos(a, q, v, r, c);
        if (r < ib && c < kb && (I0 + r) >= (K0 + c)) {
          zd* o = F + (size_t)(K0 + c) * ld + I0 + r;
          const zd old = *o;
          const zd nw = mkz(old.x - t.re[a][q][v], old.y - t.im[a][q][v]);
          *o = nw;
          if (lookahead) Ld.A[r][c] = nw;
        }
      }
  if (lookahead) {
    // this workgroup owns the next diagonal block (its values are in LDS):
    // factor it now instead of in a separate launch
    __syncthreads();
    ldl_diag_block(K, ld, I0, F, D, Dinv, status, Ld, true);
  }
}

__global__ void k_clear_status(int* status) {
  if (threadIdx.x < 3) status[threadIdx.x] = 0;
}

int ldlh_factor(Factor& f, hipStream_t s) {
  const int K = f.K;
  hipLaunchKernelGGL(k_clear_status, dim3(1), dim3(64), 0, s, f.status);
  hipLaunchKernelGGL(k_ldl_diag, dim3(1), dim3(256), 0, s, K, f.ld, 0, f.F, f.D, f.Dinv, f.status);
  for (int j0 = 0; j0 < K; j0 += NB) {
    const int jb = (K - j0 < NB) ? K - j0 : NB;
    const int rest = K - (j0 + jb);
    if (rest <= 0) break;
    const int nrb = (rest + NB - 1) / NB;
    hipLaunchKernelGGL(k_ldl_panel, dim3(nrb), dim3(256), 0, s, K, f.ld, j0, f.F, f.D, f.Dinv, f.W);
    hipLaunchKernelGGL(k_ldl_update, dim3(nrb * (nrb + 1) / 2), dim3(256), 0, s, K, f.ld, j0, f.F, f.W, f.D,
                       f.Dinv, f.status);
  }
  return (int)hipGetLastError();
}

// ------------------------------------------------------------ solve
// Sync-free blocked triangular solves.  Workgroup w owns one 64-row block
// and consumes the solution blocks it depends on as their producers publish
// them.  Producer (wave 0): stores its 64 values with device-coherent atomic
// stores, then raises the block's flag to the current epoch with a release
// store.  Consumer (wave 0): polls flags with relaxed device-coherent loads,
// one lane per block, takes every leading run of ready blocks at once and
// reads their values with device-coherent atomic loads issued after the
// poll returned.  No acquire fence is used: it would invalidate the L2 that
// holds the factor being streamed.  The next L block is prefetched into
// registers before polling, and X_b (the inverse of the diagonal block) is
// staged in LDS at kernel start.  Polling is bounded: after SPIN_LIMIT
// rounds the wave records an error in status[2] and proceeds, so every wave
// always exits.  Dependencies only point to lower workgroup indices, which
// are dispatched no later than the consumer.
constexpr long SPIN_LIMIT = 1L << 20;
constexpr int RUN_MAX = 8;   // blocks folded per round

__device__ inline zd load_published(const zd* p) {
  const unsigned long long* q = reinterpret_cast<const unsigned long long*>(p);
  unsigned long long a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned long long b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return mkz(__longlong_as_double((long long)a), __longlong_as_double((long long)b));
}
__device__ inline void store_published(zd* p, zd v) {
  unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
  __hip_atomic_store(q, (unsigned long long)__double_as_longlong(v.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q + 1, (unsigned long long)__double_as_longlong(v.y), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
// wave 0, after store_published of its values
__device__ inline void publish_flag(int* flag, int epoch) {
  if ((threadIdx.x & 63) == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave 0: number (>= 1, <= RUN_MAX) of consecutive ready blocks among the
// `left` blocks whose flags are flags[first + dir * t], t = 0, 1, ...
// Returns `left` (capped) on a spin timeout.
__device__ inline int await_run(const int* flags, int first, int dir, int left, int epoch, int* status) {
  const int lane = threadIdx.x & 63;
  const int want = min(left, RUN_MAX);
  long n = 0;
  for (;;) {
    bool ready = false;
    if (lane < want)
      ready = __hip_atomic_load(&flags[first + dir * lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch;
    const unsigned long long mask = __ballot(ready);
    const int m = (~mask == 0ull) ? 64 : (__ffsll((long long)~mask) - 1);
    if (m > 0) return min(m, want);
    __builtin_amdgcn_s_sleep(1);
    if (++n > SPIN_LIMIT) {
      if (lane == 0) atomicExch(&status[2], 1);
      return want;
    }
  }
}

// Forward: L y = b (unit lower), then z = D^{-1} y.
__global__ void __launch_bounds__(256) k_trsv_fwd(int K, int ld, const zd* __restrict__ F,
                                                  const double* __restrict__ D, const zd* __restrict__ Dinv,
                                                  const zd* __restrict__ bvec, zd* y, zd* __restrict__ z,
                                                  const int* fin, int* fout, int epoch, int* status) {
  __shared__ zd Xs[NB][NB + 1];
  __shared__ zd ys[RUN_MAX][NB];
  __shared__ zd part[4][NB];
  __shared__ int run;
  const int bl = blockIdx.x;
  const int b0 = bl * NB, bs = min(NB, K - b0);
  const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
  {
    const zd* X = Dinv + (size_t)bl * 2 * NB * NB;
    for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) Xs[e & (NB - 1)][e >> 6] = X[e];   // Xs[i][p] = X(i,p)
  }
  zd acc = mkz(0.0, 0.0);
  int done = 0;
  zd pre[NB / 4];   // prefetched column block of L for the next dependency
  if (bl > 0 && r < bs) {
#pragma unroll
    for (int u = 0; u < NB / 4; ++u) pre[u] = F[(size_t)(4 * u + g) * ld + b0 + r];
  }
  while (done < bl) {
    if (g == 0) {
      const int m = await_run(fin, done, 1, bl - done, epoch, status);
      for (int t = 0; t < m; ++t) ys[t][r] = load_published(&y[(size_t)(done + t) * NB + r]);
      if (r == 0) run = m;
    }
    __syncthreads();
    const int m = run;
    if (r < bs) {
#pragma unroll
      for (int u = 0; u < NB / 4; ++u) acc = zfma(pre[u], ys[0][4 * u + g], acc);
      for (int t = 1; t < m; ++t) {
        const zd* col = F + (size_t)((done + t) * NB) * ld + b0 + r;
        for (int c = g; c < NB; c += 4) acc = zfma(col[(size_t)c * ld], ys[t][c], acc);
      }
      if (done + m < bl) {
        const zd* col = F + (size_t)((done + m) * NB) * ld + b0 + r;
#pragma unroll
        for (int u = 0; u < NB / 4; ++u) pre[u] = col[(size_t)(4 * u + g) * ld];
      }
    }
    done += m;
    __syncthreads();
  }
  part[g][r] = acc;
  __syncthreads();
  // r_b = b_b - sum; y_b = X_b r_b
  if (g == 0)
    ys[0][r] = (r < bs) ? bvec[b0 + r] - (part[0][r] + part[1][r] + part[2][r] + part[3][r]) : mkz(0.0, 0.0);
  __syncthreads();
  zd v = mkz(0.0, 0.0);
  for (int p = g; p <= r; p += 4) v = zfma(Xs[r][p], ys[0][p], v);
  part[g][r] = v;
  __syncthreads();
  if (g == 0) {
    if (r < bs) {
      const zd yv = part[0][r] + part[1][r] + part[2][r] + part[3][r];
      store_published(&y[b0 + r], yv);
      z[b0 + r] = yv * (1.0 / D[b0 + r]);
    }
    publish_flag(&fout[bl], epoch);
  }
}

// Backward: L^H x = z, L^H read from the strict upper triangle of F.
// Workgroup w handles block nb-1-w.
__global__ void __launch_bounds__(256) k_trsv_bwd(int K, int ld, const zd* __restrict__ F,
                                                  const zd* __restrict__ Dinv, const zd* __restrict__ z,
                                                  zd* x, const int* fin, int* fout, int epoch, int* status) {
  __shared__ zd Xs[NB][NB + 1];
  __shared__ zd xs[RUN_MAX][NB];
  __shared__ zd part[4][NB];
  __shared__ int run;
  const int nb = (K + NB - 1) / NB;
  const int bl = nb - 1 - (int)blockIdx.x;
  const int b0 = bl * NB, bs = min(NB, K - b0);
  const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
  {
    const zd* XH = Dinv + (size_t)bl * 2 * NB * NB + NB * NB;
    for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) Xs[e & (NB - 1)][e >> 6] = XH[e];   // Xs[i][p] = X^H(i,p)
  }
  zd acc = mkz(0.0, 0.0);
  int done = 0;                  // blocks nb-1, nb-2, ... already folded
  const int need = nb - 1 - bl;
  while (done < need) {
    if (g == 0) {
      const int m = await_run(fin, nb - 1 - done, -1, need - done, epoch, status);
      for (int t = 0; t < m; ++t) {
        const int J = nb - 1 - done - t;
        const int js = min(NB, K - J * NB);
        xs[t][r] = (r < js) ? load_published(&x[(size_t)J * NB + r]) : mkz(0.0, 0.0);
      }
      if (r == 0) run = m;
    }
    __syncthreads();
    const int m = run;
    if (r < bs)
      for (int t = 0; t < m; ++t) {
        const int J = nb - 1 - done - t;
        const int js = min(NB, K - J * NB);
        const zd* col = F + (size_t)(J * NB) * ld + b0 + r;
        for (int c = g; c < js; c += 4) acc = zfma(col[(size_t)c * ld], xs[t][c], acc);
      }
    done += m;
    __syncthreads();
  }
  part[g][r] = acc;
  __syncthreads();
  if (g == 0)
    xs[0][r] = (r < bs) ? z[b0 + r] - (part[0][r] + part[1][r] + part[2][r] + part[3][r]) : mkz(0.0, 0.0);
  __syncthreads();
  // x_b = X_b^H rb
  zd v = mkz(0.0, 0.0);
  for (int p = r + ((g - r % 4 + 4) % 4); p < bs; p += 4) v = zfma(Xs[r][p], xs[0][p], v);
  part[g][r] = v;
  __syncthreads();
  if (g == 0) {
    if (r < bs) store_published(&x[b0 + r], part[0][r] + part[1][r] + part[2][r] + part[3][r]);
    publish_flag(&fout[bl], epoch);
  }
}

int ldlh_solve(Factor& f, const zd* b, zd* x, hipStream_t s) {
  const int nb = (f.K + NB - 1) / NB;
  zd* y = f.vec + 3 * (size_t)f.ld;
  zd* z = f.tvec;
  int* ff = f.flags;
  int* fb = f.flags + f.ld / NB;
  const int e1 = ++f.epoch;
  hipLaunchKernelGGL(k_trsv_fwd, dim3(nb), dim3(256), 0, s, f.K, f.ld, f.F, f.D, f.Dinv, b, y, z, ff, ff, e1,
                     f.status);
  const int e2 = ++f.epoch;
  hipLaunchKernelGGL(k_trsv_bwd, dim3(nb), dim3(256), 0, s, f.K, f.ld, f.F, f.Dinv, z, x, fb, fb, e2, f.status);
  return (int)hipGetLastError();
}

// -------------------------------------------------------------- GEMV
// y = (A - shift A2) x for full column-major ld x K matrices (A2 may be
// null).  Workgroup = 64 rows x 4 column groups over a chunk of columns;
// partial sums per chunk are reduced by k_gemv_reduce in a fixed order.
constexpr int GEMV_CHUNK = 128;

__global__ void __launch_bounds__(256) k_gemv_part(int K, int ld, const zd* __restrict__ A,
                                                   const zd* __restrict__ A2, double shift,
                                                   const zd* __restrict__ x, zd* __restrict__ part) {
  __shared__ zd red[4][64];
  const int r = threadIdx.x % 64, g = threadIdx.x / 64;
  const int i = blockIdx.x * 64 + r;
  const int c0 = blockIdx.y * GEMV_CHUNK;
  const int c1 = min(K, c0 + GEMV_CHUNK);
  zd acc = mkz(0.0, 0.0);
  if (i < K) {
    if (A2) {
      for (int c = c0 + g; c < c1; c += 4) {
        const size_t o = (size_t)c * ld + i;
        const zd a = A[o], a2 = A2[o];
        const zd m = mkz(fma(-shift, a2.x, a.x), fma(-shift, a2.y, a.y));
        acc = zfma(m, x[c], acc);
      }
    } else {
      for (int c = c0 + g; c < c1; c += 4) acc = zfma(A[(size_t)c * ld + i], x[c], acc);
    }
  }
  red[g][r] = acc;
  __syncthreads();
  if (g == 0 && i < K) part[(size_t)blockIdx.y * ld + i] = red[0][r] + red[1][r] + red[2][r] + red[3][r];
}

__global__ void k_gemv_reduce(int K, int ld, int nch, const zd* __restrict__ part, zd* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K) return;
  zd s = mkz(0.0, 0.0);
  for (int q = 0; q < nch; ++q) s = s + part[(size_t)q * ld + i];
  y[i] = s;
}

// ------------------------------------------------------ vector kernels
// Block-wide reduction helpers for a 1024-thread workgroup.
__device__ inline double block_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (tid < o) sh[tid] += sh[tid + o];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}
__device__ inline double block_max(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (tid < o) sh[tid] = fmax(sh[tid], sh[tid + o]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// One inverse-iteration bookkeeping step (GHEPIIS, src/linalg.f90:1892-1909):
// x /= max|Re|,|Im|; tc = x^H v / v^H v; out[0] = sqrt(2K) ||x - tc v||,
// out[1] = the scale used.
__global__ void __launch_bounds__(1024) k_invit_finish(int K, zd* __restrict__ x, const zd* __restrict__ v,
                                                       double* __restrict__ out) {
  __shared__ double sh[1024];
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
    out[0] = sqrt(2.0 * K) * sqrt(nd);
    out[1] = t1;
  }
}

// out[2] = Re(x^H y1), out[3] = Re(x^H y2)
__global__ void __launch_bounds__(1024) k_dots(int K, const zd* __restrict__ x, const zd* __restrict__ y1,
                                               const zd* __restrict__ y2, double* __restrict__ out) {
  __shared__ double sh[1024];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < K; i += blockDim.x) {
    a += x[i].x * y1[i].x + x[i].y * y1[i].y;
    b += x[i].x * y2[i].x + x[i].y * y2[i].y;
  }
  a = block_sum(a, sh);
  b = block_sum(b, sh);
  if (threadIdx.x == 0) {
    out[2] = a;
    out[3] = b;
  }
}

__global__ void k_scale(int K, zd* __restrict__ x, const double* __restrict__ t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < K) x[i] = x[i] * (1.0 / sqrt(*t));
}

static int gemv(const zd* A, const zd* A2, double shift, int ld, int K, const zd* x, zd* y, zd* part,
                hipStream_t s) {
  const int nch = (K + GEMV_CHUNK - 1) / GEMV_CHUNK;
  hipLaunchKernelGGL(k_gemv_part, dim3((K + 63) / 64, nch), dim3(256), 0, s, K, ld, A, A2, shift, x, part);
  hipLaunchKernelGGL(k_gemv_reduce, dim3((K + 255) / 256), dim3(256), 0, s, K, ld, nch, part, y);
  return (int)hipGetLastError();
}

// One inverse-iteration step: x = F^{-1} (S v), then the GHEPIIS
// normalisation and convergence measure (written to f.scal[0..1]).
// Work: f.vec holds 4 vectors of ld entries; the GEMV partials use
// f.W (ld x NB >= ld x ceil(K/128)).
int invit_iteration(const Ctx& c, Factor& f, const zd* v, zd* x, hipStream_t s) {
  zd* w = f.vec;
  int rc = gemv(c.S, nullptr, 0.0, c.ld, c.K, v, w, f.W, s);
  if (rc) return rc;
  rc = ldlh_solve(f, w, x, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_invit_finish, dim3(1), dim3(1024), 0, s, c.K, x, v, f.scal);
  return (int)hipGetLastError();
}

// Rayleigh quotient pieces: scal[2] = x^H S x, scal[3] = x^H (H - sigma S) x.
int rayleigh(const Ctx& c, Factor& f, const zd* x, double sigma, hipStream_t s) {
  zd* y1 = f.vec + c.ld;
  zd* y2 = f.vec + 2 * (size_t)c.ld;
  int rc = gemv(c.S, nullptr, 0.0, c.ld, c.K, x, y1, f.W, s);
  if (rc) return rc;
  rc = gemv(c.H, c.S, sigma, c.ld, c.K, x, y2, f.W, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_dots, dim3(1), dim3(1024), 0, s, c.K, x, y1, y2, f.scal);
  return (int)hipGetLastError();
}

int normalise(zd* x, int K, const double* t, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3((K + 255) / 256), dim3(256), 0, s, K, x, t);
  return (int)hipGetLastError();
}

}  // namespace ecg
