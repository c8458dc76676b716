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
// One workgroup factors the jb x jb diagonal block at (j0, j0) in LDS,
// stores unit-lower L (strict part, and L^H in the upper triangle) and D,
// and the inverse X = L^{-1} of the block (used by the panel solve and the
// triangular solves).  Thread e-loops run over the full 64 x 64 index space
// with cheap bit arithmetic; two barriers per pivot.
__device__ void ldl_diag_block(int K, int ld, int j0, zd* __restrict__ F, double* __restrict__ D,
                               zd* __restrict__ Dinv, int* __restrict__ status, zd (*A)[NB + 1],
                               zd (*X)[NB + 1], double* dsh, bool a_in_lds) {
  const int jb = min(NB, K - j0);
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int i = e & (NB - 1), j = e >> 6;
    if (!a_in_lds)
      A[i][j] = (i < jb && j < jb && i >= j) ? F[(size_t)(j0 + j) * ld + j0 + i] : mkz(0.0, 0.0);
    X[i][j] = (i == j) ? mkz(1.0, 0.0) : mkz(0.0, 0.0);
  }
  __syncthreads();
  for (int p = 0; p < jb; ++p) {
    const double d = A[p][p].x;
    const double rd = (d != 0.0) ? 1.0 / d : 0.0;
    if (tid == 0) {
      dsh[p] = d;
      if (d < 0.0) atomicAdd(&status[0], 1);
      if (d == 0.0) atomicCAS(&status[1], 0, j0 + p + 1);
    }
    // trailing update with the unscaled column: A(i,j) -= A(i,p) conj(A(j,p)) / d
    for (int e = tid; e < NB * NB; e += blockDim.x) {
      const int i = e & (NB - 1), j = e >> 6;
      if (j > p && i >= j && i < jb) A[i][j] = zfms(A[i][p] * rd, zconj(A[j][p]), A[i][j]);
    }
    __syncthreads();
    if (tid > p && tid < jb) A[tid][p] = A[tid][p] * rd;
    __syncthreads();
  }
  // X = L^{-1} by row operations in pivot order
  for (int p = 0; p < jb; ++p) {
    for (int e = tid; e < NB * NB; e += blockDim.x) {
      const int i = e & (NB - 1), j = e >> 6;
      if (i > p && i < jb && j <= p) X[i][j] = zfms(A[i][p], X[p][j], X[i][j]);
    }
    __syncthreads();
  }
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int i = e & (NB - 1), j = e >> 6;
    if (i < jb && j < jb) {
      if (i > j) {
        const zd l = A[i][j];
        F[(size_t)(j0 + j) * ld + j0 + i] = l;
        F[(size_t)(j0 + i) * ld + j0 + j] = zconj(l);
      }
      Dinv[(size_t)(j0 / NB) * NB * NB + (size_t)j * NB + i] = (i >= j) ? X[i][j] : mkz(0.0, 0.0);
    }
  }
  for (int p = tid; p < jb; p += blockDim.x) D[j0 + p] = dsh[p];
}

__global__ void __launch_bounds__(256) k_ldl_diag(int K, int ld, int j0, zd* __restrict__ F,
                                                  double* __restrict__ D, zd* __restrict__ Dinv,
                                                  int* __restrict__ status) {
  __shared__ zd A[NB][NB + 1];
  __shared__ zd X[NB][NB + 1];
  __shared__ double dsh[NB];
  ldl_diag_block(K, ld, j0, F, D, Dinv, status, A, X, dsh, false);
}

// ------------------------------------------------------------ panel solve
// Row block i0..i0+ib-1 below the diagonal block: W = A X^H (so that
// W L_JJ^H = A), L = W D^{-1}; L goes to F (and L^H to its upper
// triangle), W to the panel scratch used by the trailing update.
__global__ void __launch_bounds__(256) k_ldl_panel(int K, int ld, int j0, zd* __restrict__ F,
                                                   const double* __restrict__ D,
                                                   const zd* __restrict__ Dinv, zd* __restrict__ W) {
  __shared__ zd A[NB][NB + 1];
  __shared__ zd X[NB][NB + 1];
  const int jb = min(NB, K - j0);
  const int i0 = j0 + jb + blockIdx.x * NB;
  const int ib = min(NB, K - i0);
  const int tid = threadIdx.x;
  const zd* Xg = Dinv + (size_t)(j0 / NB) * NB * NB;
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int i = e % NB, j = e / NB;
    A[i][j] = (i < ib && j < jb) ? F[(size_t)(j0 + j) * ld + i0 + i] : mkz(0.0, 0.0);
    X[i][j] = (i < jb && j < jb) ? Xg[(size_t)j * NB + i] : mkz(0.0, 0.0);
  }
  __syncthreads();
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int r = e % NB, c = e / NB;
    if (r >= ib || c >= jb) continue;
    zd w = mkz(0.0, 0.0);
    for (int p = 0; p <= c; ++p) w = zfma(A[r][p], zconj(X[c][p]), w);
    const zd l = w * (1.0 / D[j0 + c]);
    W[(size_t)c * ld + i0 + r] = w;
    F[(size_t)(j0 + c) * ld + i0 + r] = l;
    F[(size_t)(i0 + r) * ld + j0 + c] = zconj(l);
  }
}

// -------------------------------------------------------- trailing update
// Lower block pairs (I >= Kb) of the trailing matrix: C_IK -= W_I L_K^H on
// FP64 MFMA (v_mfma_f64_16x16x4f64, four real products per complex one).
// 4 waves per 64 x 64 output tile, each wave a 32 x 32 quarter (2 x 2 MFMA
// tiles).  Operand fragments are read straight from the panel scratch W
// and the factor's panel column (L2-resident).  The workgroup that updates
// the first trailing diagonal block then factors it (look-ahead), so the
// next block column needs no separate diagonal launch.
typedef double dbl4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_ldl_update(int K, int ld, int j0, zd* __restrict__ F,
                                                    const zd* __restrict__ W, double* __restrict__ D,
                                                    zd* __restrict__ Dinv, int* __restrict__ status) {
  __shared__ zd A[NB][NB + 1];
  __shared__ zd X[NB][NB + 1];
  __shared__ double dsh[NB];
  const int jb = min(NB, K - j0);
  const int t0 = (j0 + jb) / NB;
  const int b = blockIdx.x;
  int bi = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= b) ++bi;
  while (bi * (bi + 1) / 2 > b) --bi;
  const int bk = b - bi * (bi + 1) / 2;
  const int I0 = (t0 + bi) * NB, K0 = (t0 + bk) * NB;
  const int ib = min(NB, K - I0), kb = min(NB, K - K0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rt = (wave & 1) * 32, ct = (wave >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  dbl4 cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      cr[a][q] = dbl4{0.0, 0.0, 0.0, 0.0};
      ci[a][q] = dbl4{0.0, 0.0, 0.0, 0.0};
    }
  for (int p0 = 0; p0 < jb; p0 += 4) {
    const int p = p0 + lk;
    double war[2], wai[2], lrr[2], lri[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int r = rt + a * 16 + lr;
      zd w = (p < jb && r < ib) ? W[(size_t)p * ld + I0 + r] : mkz(0.0, 0.0);
      war[a] = w.x;
      wai[a] = w.y;
      const int c = ct + a * 16 + lr;
      zd l = (p < jb && c < kb) ? F[(size_t)(j0 + p) * ld + K0 + c] : mkz(0.0, 0.0);
      lrr[a] = l.x;
      lri[a] = l.y;
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        // (Wr + i Wi)(Lr - i Li): re = Wr Lr + Wi Li, im = Wi Lr - Wr Li; subtract
        cr[a][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-war[a], lrr[q], cr[a][q], 0, 0, 0);
        cr[a][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-wai[a], lri[q], cr[a][q], 0, 0, 0);
        ci[a][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-wai[a], lrr[q], ci[a][q], 0, 0, 0);
        ci[a][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(war[a], lri[q], ci[a][q], 0, 0, 0);
      }
  }
  // accumulator (v) of lane: row 4*(lane/16)+v, column lane%16 of the 16x16 tile
  const bool lookahead = (bi == 0 && bk == 0);
  if (lookahead)
    for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) A[e & (NB - 1)][e >> 6] = mkz(0.0, 0.0);
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = rt + a * 16 + 4 * lk + v, c = ct + q * 16 + lr;
        if (r < ib && c < kb && (I0 + r) >= (K0 + c)) {
          zd* o = F + (size_t)(K0 + c) * ld + I0 + r;
          const zd old = *o;
          const zd nw = mkz(old.x + cr[a][q][v], old.y + ci[a][q][v]);
          *o = nw;
          if (lookahead) A[r][c] = nw;
        }
      }
  if (lookahead) {
    // this workgroup owns the next diagonal block (its values are in LDS):
    // factor it now instead of in a separate launch
    __syncthreads();
    ldl_diag_block(K, ld, I0, F, D, Dinv, status, A, X, dsh, true);
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
// them: a producer's wave 0 stores its 64 values, fences at device scope and
// raises the block's flag to the current epoch; a consumer's wave 0 spins on
// the flag (bounded: after SPIN_LIMIT polls it records an error in
// status[2] and proceeds, so every wave always exits), fences, and stages the
// values in LDS.  Dependencies only point to lower workgroup indices, which
// are dispatched no later than the consumer, so progress is guaranteed.
constexpr long SPIN_LIMIT = 1L << 20;

__device__ inline void publish_flag(int* flag, int epoch) {
  __threadfence();
  if ((threadIdx.x & 63) == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// wave 0 only; returns after the flag shows `epoch` (or the limit is hit)
__device__ inline void await_flag(int* flag, int epoch, int* status) {
  if ((threadIdx.x & 63) == 0) {
    long n = 0;
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
      __builtin_amdgcn_s_sleep(1);
      if (++n > SPIN_LIMIT) {
        atomicExch(&status[2], 1);
        break;
      }
    }
  }
  __threadfence();
}

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

// Forward: L y = b (unit lower), then z = D^{-1} y.
__global__ void __launch_bounds__(256) k_trsv_fwd(int K, int ld, const zd* __restrict__ F,
                                                  const double* __restrict__ D, const zd* __restrict__ Dinv,
                                                  const zd* __restrict__ bvec, zd* y, zd* __restrict__ z,
                                                  int* flags, int epoch, int* status) {
  __shared__ zd ys[NB];
  __shared__ zd part[4][NB];
  const int bl = blockIdx.x;
  const int b0 = bl * NB, bs = min(NB, K - b0);
  const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
  zd acc = mkz(0.0, 0.0);
  for (int J = 0; J < bl; ++J) {
    if (g == 0) {
      await_flag(&flags[J], epoch, status);
      ys[r] = load_published(&y[(size_t)J * NB + r]);
    }
    __syncthreads();
    if (r < bs)
      for (int c = g; c < NB; c += 4) acc = zfma(F[(size_t)(J * NB + c) * ld + b0 + r], ys[c], acc);
    __syncthreads();
  }
  part[g][r] = acc;
  __syncthreads();
  // r_b = b_b - sum, then y_b = X_b r_b (X_b = L_bb^{-1}, column-major 64 x 64)
  if (g == 0)
    ys[r] = (r < bs) ? bvec[b0 + r] - (part[0][r] + part[1][r] + part[2][r] + part[3][r]) : mkz(0.0, 0.0);
  __syncthreads();
  if (g == 0) {
    if (r < bs) {
      const zd* X = Dinv + (size_t)bl * NB * NB;
      zd v = mkz(0.0, 0.0);
      for (int p = 0; p <= r; ++p) v = zfma(X[(size_t)p * NB + r], ys[p], v);
      store_published(&y[b0 + r], v);
      z[b0 + r] = v * (1.0 / D[b0 + r]);
    }
    publish_flag(&flags[bl], epoch);
  }
}

// Backward: L^H x = z, L^H read from the strict upper triangle of F.
// Workgroup w handles block nb-1-w.
__global__ void __launch_bounds__(256) k_trsv_bwd(int K, int ld, const zd* __restrict__ F,
                                                  const zd* __restrict__ Dinv, const zd* __restrict__ z,
                                                  zd* x, int* flags, int epoch, int* status) {
  __shared__ zd xs[NB];
  __shared__ zd part[4][NB];
  const int nb = (K + NB - 1) / NB;
  const int bl = nb - 1 - (int)blockIdx.x;
  const int b0 = bl * NB, bs = min(NB, K - b0);
  const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
  zd acc = mkz(0.0, 0.0);
  for (int J = nb - 1; J > bl; --J) {
    const int js = min(NB, K - J * NB);
    if (g == 0) {
      await_flag(&flags[J], epoch, status);
      xs[r] = (r < js) ? load_published(&x[(size_t)J * NB + r]) : mkz(0.0, 0.0);
    }
    __syncthreads();
    if (r < bs)
      for (int c = g; c < js; c += 4) acc = zfma(F[(size_t)(J * NB + c) * ld + b0 + r], xs[c], acc);
    __syncthreads();
  }
  part[g][r] = acc;
  __syncthreads();
  if (g == 0)
    xs[r] = (r < bs) ? z[b0 + r] - (part[0][r] + part[1][r] + part[2][r] + part[3][r]) : mkz(0.0, 0.0);
  __syncthreads();
  if (g == 0) {
    if (r < bs) {
      // x_b = X_b^H rb: sum_{p >= r} conj(X(p, r)) rb_p
      const zd* X = Dinv + (size_t)bl * NB * NB;
      zd v = mkz(0.0, 0.0);
      for (int p = r; p < bs; ++p) v = zfma(zconj(X[(size_t)r * NB + p]), xs[p], v);
      store_published(&x[b0 + r], v);
    }
    publish_flag(&flags[bl], epoch);
  }
}

int ldlh_solve(Factor& f, const zd* b, zd* x, hipStream_t s) {
  const int nb = (f.K + NB - 1) / NB;
  zd* y = f.vec + 3 * (size_t)f.ld;
  zd* z = f.tvec;
  const int e1 = ++f.epoch;
  hipLaunchKernelGGL(k_trsv_fwd, dim3(nb), dim3(256), 0, s, f.K, f.ld, f.F, f.D, f.Dinv, b, y, z, f.flags, e1,
                     f.status);
  const int e2 = ++f.epoch;
  hipLaunchKernelGGL(k_trsv_bwd, dim3(nb), dim3(256), 0, s, f.K, f.ld, f.F, f.Dinv, z, x, f.flags + f.ld / NB,
                     e2, f.status);
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
