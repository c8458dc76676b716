// Batched candidate screening for the last basis position.
//
// The reference scores trial functions one at a time (GenerateTrialParam
// and the trial loops, src/workproc.f90:1375-1487, 3397-3413, 4054-4070):
// for each trial it recomputes row K of H and S, updates the LDL^H factor
// from row K (LDLHF, src/linalg.f90:316) and runs inverse iteration
// (GHEPIIS, src/linalg.f90:1781).  Here all candidates of a step share the
// fixed leading block M' = H' - sigma S' (functions 1..K-1):
//   set-up (once per session):  LDL^H of H - sigma S, Z = L'^{-1} (the
//     leading block of the L^{-1} the factorisation builds),
//     G = Z^H D'^{-1} Z = M'^{-1}, P = G S';
//   per step, for C candidates (columns of one matrix):
//     border columns m = h - sigma s, s of each candidate (pair kernel),
//     [w_m | w_s] = G [m | s]                     (one GEMM, N = 2C),
//     Schur complement delta = mu - m^H w_m        (the candidate's pivot),
//     inverse iteration x = M_j^{-1} S_j v with the bordered solve
//       u = P v' + w_s v_K,  x_K = (s^H v' + v_K - m^H u) / delta,
//       x' = u - w_m x_K                            (one GEMM per iteration),
//     GHEPIIS's scaling and convergence measure per candidate,
//     Rayleigh quotient E = x^H H_j x / x^H S_j x (two GEMMs).
// Every GEMM is Kp x Kp times Kp x C on MFMA; nothing in a step synchronises
// with the host, so a step can be captured in a hipGraph and replayed.
#include "ecg_internal.hpp"

namespace ecg {

// ------------------------------------------------------------------ ZGEMM
// C = alpha op(A) B + beta C, complex, column-major; op(A) = A (TRANS = 0)
// or A^H (TRANS = 1).  Workgroup tile TM x 32 (TM = 64 or 32), k-chunks
// of 32 staged in LDS with coalesced loads (the next chunk's loads in
// flight during the current chunk's MFMAs); wave w computes rows
// (TM / 2)(w & 1) + 16 a (a < TM / 32) and columns 16 (w >> 1) as 16 x 16
// f64 MFMA tiles, three real products per complex one.  TM = 32 doubles
// the workgroups (more waves per SIMD to hide latency) at the price of
// twice the reads of B; the screening session picks TM by timing both on
// its own shapes (screen_tune).
typedef double gdbl4 __attribute__((ext_vector_type(4)));

template <int TRANS, int TM>
__global__ void __launch_bounds__(256) k_zgemm(int M, int N, int Kd, zd alpha, const zd* __restrict__ A, int lda,
                                               const zd* __restrict__ B, int ldb, zd beta, zd* __restrict__ Cm,
                                               int ldc, int lower) {
  constexpr int KC = 32;           // k-chunk
  constexpr int NA = TM / 32;      // 16-row tiles per wave
  __shared__ zd As[KC][TM + 1];   // As[p][r] = op(A)(r0 + r, p0 + p)
  __shared__ zd Bs[32][KC + 1];   // Bs[c][p] = B(p0 + p, c0 + c)
  const int r0 = blockIdx.x * TM, c0 = blockIdx.y * 32;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int rt = (wave & 1) * (TM / 2), ct = (wave >> 1) * 16;
  // three real products per complex one: for x b, p1 += xr br, p2 += xi bi,
  // p3 += (xr + xi)(br + bi); x b = (p1 - p2) + i (p3 - p1 - p2).  The
  // step is bound by the MFMA rate here (a quarter fewer MFMAs than the
  // four-product form); both forms' rounding errors are bounded by a few
  // units of sum |x||b|.
  gdbl4 p1[NA], p2[NA], p3[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) p1[a] = p2[a] = p3[a] = gdbl4{0.0, 0.0, 0.0, 0.0};
  // chunk loads into registers; the next chunk's loads are issued before
  // the current chunk's MFMAs, so their latency overlaps the arithmetic
  zd va[TM * KC / 256], vb[KC / 8];
  auto load_chunk = [&](int p0) {
#pragma unroll
    for (int u = 0; u < TM * KC / 256; ++u) {
      const int e = tid + 256 * u;
      int r, p;
      if (TRANS == 0) {
        r = e % TM;
        p = e / TM;
      } else {
        p = e & (KC - 1);
        r = e / KC;
      }
      const bool in = (r0 + r < M) && (p0 + p < Kd);
      // guarded loads as branches: here the branch-free form (ld_or0)
      // measured 41 % slower over the screening step (2.71 against
      // 1.92 ms at K = 2000, 256 candidates)
      zd v = mkz(0.0, 0.0);
      if (in) {
        if (TRANS == 0) v = A[(size_t)(p0 + p) * lda + r0 + r];
        else v = zconj(A[(size_t)(r0 + r) * lda + p0 + p]);
      }
      va[u] = v;
    }
#pragma unroll
    for (int u = 0; u < KC / 8; ++u) {
      const int e = tid + 256 * u, p = e & (KC - 1), c = e / KC;
      vb[u] = (c0 + c < N && p0 + p < Kd) ? B[(size_t)(c0 + c) * ldb + p0 + p] : mkz(0.0, 0.0);
    }
  };
  // lower: A and B are zero above their diagonals, so for the tile's rows
  // r >= r0 and columns c >= c0 the terms p < max(r0, c0) vanish
  const int pb = lower ? (max(r0, c0) / KC) * KC : 0;
  load_chunk(pb);
  for (int p0 = pb; p0 < Kd; p0 += KC) {
    __syncthreads();   // the previous chunk's reads are done
#pragma unroll
    for (int u = 0; u < TM * KC / 256; ++u) {
      const int e = tid + 256 * u;
      if (TRANS == 0) As[e / TM][e % TM] = va[u];
      else As[e & (KC - 1)][e / KC] = va[u];
    }
#pragma unroll
    for (int u = 0; u < KC / 8; ++u) {
      const int e = tid + 256 * u;
      Bs[e / KC][e & (KC - 1)] = vb[u];
    }
    __syncthreads();
    if (p0 + KC < Kd) load_chunk(p0 + KC);
#pragma unroll
    for (int kk = 0; kk < KC / 4; ++kk) {
      const int p = 4 * kk + lk;
      const zd b = Bs[ct + lr][p];
      const double bs = b.x + b.y;
#pragma unroll
      for (int a = 0; a < NA; ++a) {
        const zd x = As[p][rt + 16 * a + lr];
        p1[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, b.x, p1[a], 0, 0, 0);
        p2[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, b.y, p2[a], 0, 0, 0);
        p3[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x + x.y, bs, p3[a], 0, 0, 0);
      }
    }
  }
  const bool bz = (beta.x == 0.0 && beta.y == 0.0);
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = r0 + rt + 16 * a + 4 * v + lk, c = c0 + ct + lr;
      if (r < M && c < N) {
        const zd acc = mkz(p1[a][v] - p2[a][v], p3[a][v] - (p1[a][v] + p2[a][v]));
        zd* o = Cm + (size_t)c * ldc + r;
        zd out = alpha * acc;
        if (!bz) out = zfma(beta, *o, out);
        *o = out;
      }
    }
}

// C = alpha A B + beta C for few columns (N <= GV_N, A M x Kd column-major):
// the MFMA tiles would leave most of each 32-column tile and most CUs idle,
// and the product is a stream of A, so it is laid out for bandwidth. A
// workgroup of 16 waves takes 8 rows (one 128-byte line of each column of A);
// lane l of wave w reads row l % 8 of columns p = 8 w + l / 8 (mod 128), four
// columns 128 apart per load batch so that each lane keeps four loads in
// flight. The 128 partial sums of a row are added in a fixed order: a
// butterfly over the eight column slots of a wave, then the 16 waves in
// sequence through LDS.
constexpr int GV_N = 8;
constexpr int GV_ROWS = 8;
__global__ void __launch_bounds__(1024) k_zgemv(int M, int N, int Kd, zd alpha, const zd* __restrict__ A, int lda,
                                                const zd* __restrict__ B, int ldb, zd beta, zd* __restrict__ Cm,
                                                int ldc) {
  __shared__ zd red[16][GV_ROWS][GV_N + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 7, ks = 8 * wave + (lane >> 3);
  const int row = blockIdx.x * GV_ROWS + r;
  zd acc[GV_N];
#pragma unroll
  for (int n = 0; n < GV_N; ++n) acc[n] = mkz(0.0, 0.0);
  if (row < M) {
    for (int p0 = ks; p0 < Kd; p0 += 512) {
      zd a[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int p = p0 + 128 * u;
        a[u] = p < Kd ? A[(size_t)p * lda + row] : mkz(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int p = p0 + 128 * u;
        if (p < Kd) {
#pragma unroll
          for (int n = 0; n < GV_N; ++n)
            if (n < N) acc[n] = zfma(a[u], B[(size_t)n * ldb + p], acc[n]);
        }
      }
    }
  }
#pragma unroll
  for (int n = 0; n < GV_N; ++n) {
#pragma unroll
    for (int off = 8; off < 64; off <<= 1) {
      acc[n].x += __shfl_xor(acc[n].x, off);
      acc[n].y += __shfl_xor(acc[n].y, off);
    }
  }
  if (lane < GV_ROWS) {
#pragma unroll
    for (int n = 0; n < GV_N; ++n) red[wave][lane][n] = acc[n];
  }
  __syncthreads();
  if (threadIdx.x < GV_ROWS * GV_N) {
    const int rr = threadIdx.x % GV_ROWS, n = threadIdx.x / GV_ROWS;
    const int i = blockIdx.x * GV_ROWS + rr;
    if (n < N && i < M) {
      zd t = mkz(0.0, 0.0);
      for (int k = 0; k < 16; ++k) t = t + red[k][rr][n];
      zd* c = Cm + (size_t)n * ldc + i;
      const bool zb = beta.x == 0.0 && beta.y == 0.0;
      *c = zb ? alpha * t : zfma(beta, *c, alpha * t);
    }
  }
}

int zgemm(int trans, int M, int N, int Kd, zd alpha, const zd* A, int lda, const zd* B, int ldb, zd beta, zd* Cm,
          int ldc, hipStream_t s, int tm, int lower) {
  if (M <= 0 || N <= 0) return 0;
  if (!trans && !lower && N <= GV_N) {
    hipLaunchKernelGGL(k_zgemv, dim3((M + GV_ROWS - 1) / GV_ROWS), dim3(1024), 0, s, M, N, Kd, alpha, A, lda, B, ldb, beta, Cm, ldc);
    return (int)hipGetLastError();
  }
  if (tm != 32) tm = 64;
  const dim3 grid((M + tm - 1) / tm, (N + 31) / 32);
  if (trans) {
    if (tm == 32) hipLaunchKernelGGL((k_zgemm<1, 32>), grid, dim3(256), 0, s, M, N, Kd, alpha, A, lda, B, ldb, beta, Cm, ldc, lower);
    else hipLaunchKernelGGL((k_zgemm<1, 64>), grid, dim3(256), 0, s, M, N, Kd, alpha, A, lda, B, ldb, beta, Cm, ldc, lower);
  } else {
    if (tm == 32) hipLaunchKernelGGL((k_zgemm<0, 32>), grid, dim3(256), 0, s, M, N, Kd, alpha, A, lda, B, ldb, beta, Cm, ldc, lower);
    else hipLaunchKernelGGL((k_zgemm<0, 64>), grid, dim3(256), 0, s, M, N, Kd, alpha, A, lda, B, ldb, beta, Cm, ldc, lower);
  }
  return (int)hipGetLastError();
}

// The screening step's two shapes for C candidates, each the faster of
// its variants timed on the session's own operands: the per-step GEMMs'
// row tile (64 or 32 rows; timed on P V, Kp x C x Kp into U, which the
// step overwrites) and the border rows' candidates per workgroup (1, 2, 4,
// 8; timed on the candidates in cparams, writing the step's own row
// buffers); an alternative replaces the first variant only when it is
// more than 3 % faster.  Kept for this C and basis sizes in the same
// 512-wide band;
// runs outside any graph capture (it synchronises).
// Keeps the GPU busy for `ticks` of the 100 MHz wall clock (bounded), so
// that launches queued behind it run back to back even when the host
// queues them slower than the GPU runs them.
__global__ void k_hold(long long ticks) {
  const long long t0 = wall_clock64();
  while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(8);
}

int screen_tune(const Ctx& c, Screen& sc, int C, hipStream_t s) {
  if (sc.tuned_c == C && sc.tuned_kb == (sc.Kp >> 9)) return 0;
  const int Kp = sc.Kp, ld = c.ld;
  const zd one = mkz(1.0, 0.0), zero = mkz(0.0, 0.0);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -1;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -1;
  }
  auto timed = [&](auto&& run, float& ms) -> int {
    int rc = run();   // warm-up
    if (!rc) hipLaunchKernelGGL(k_hold, dim3(1), dim3(64), 0, s, 30000LL);   // 300 us
    if (!rc && hipEventRecord(e0, s) != hipSuccess) rc = -1;
    for (int r = 0; r < 5 && !rc; ++r) rc = run();
    if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = -1;
    return rc;
  };
  int rc = 0;
  float best = 0.0f;
  for (int tm : {64, 32}) {
    float ms = 0.0f;
    rc = timed([&] { return zgemm(0, Kp, C, Kp, one, sc.P, ld, sc.V, ld, zero, sc.U, ld, s, tm); }, ms);
    if (rc) break;
    if (tm == 64 || ms < 0.97f * best) {   // an alternative must win by more than the timing noise
      best = ms;
      sc.tm = tm;
    }
  }
  // the candidates' records once, then the row kernel alone per variant
  if (!rc)
    rc = candidate_rows(c, Kp, C, sc.cparams, sc.cbra, sc.cket, sc.cdiag, sc.Hb, sc.Bm + (size_t)C * ld, ld, 1, s);
  for (int cpb : {1, 2, 4, 8}) {
    if (rc) break;
    if (cpb > 1 && cand_rows_lds(c.n, c.sc.T, cpb) == 0) continue;
    float ms = 0.0f;
    rc = timed([&] {
      return candidate_rows(c, Kp, C, sc.cparams, sc.cbra, sc.cket, sc.cdiag, sc.Hb, sc.Bm + (size_t)C * ld, ld, cpb,
                            s, true);
    }, ms);
    if (rc) break;
    if (cpb == 1 || ms < 0.97f * best) {
      best = ms;
      sc.cpb = cpb;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!rc) {
    sc.tuned_c = C;
    sc.tuned_kb = sc.Kp >> 9;
  }
  return rc;
}

// ------------------------------------------------------------ set-up
// Y(i, j) = Z(i, j) / D_i
__global__ void k_row_scale(int n, int ld, const double* __restrict__ D, const zd* __restrict__ Z,
                            zd* __restrict__ Y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < n) Y[(size_t)j * ld + i] = Z[(size_t)j * ld + i] * (1.0 / D[i]);
}

// Z (Kp x Kp, column-major, ld) = X = L^{-1} from X^H held in F's strict
// upper triangle (F(r, c) = conj(X(c, r)), r < c): Z(i, j) = conj(F(j, i))
// for i > j, 1 on the diagonal, 0 above.  A 64 x 64 tile per workgroup,
// transposed through LDS so that both the reads of F and the writes of Z
// are coalesced.
__global__ void __launch_bounds__(256) k_z_from_xh(int Kp, int ld, const zd* __restrict__ F, zd* __restrict__ Z) {
  __shared__ zd t[64][65];
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;   // Z rows i0.., columns j0..
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  if (i0 + 63 > j0) {   // the tile reaches below the diagonal
    for (int r = ty; r < 64; r += 4) {
      const int i = i0 + r, j = j0 + tx;
      t[r][tx] = (i < Kp && j < i) ? F[(size_t)i * ld + j] : mkz(0.0, 0.0);
    }
  }
  __syncthreads();
  for (int cc = ty; cc < 64; cc += 4) {
    const int j = j0 + cc, i = i0 + tx;
    if (i < Kp && j < Kp) Z[(size_t)j * ld + i] = (i > j) ? zconj(t[tx][cc]) : mkz(i == j ? 1.0 : 0.0, 0.0);
  }
}

int screen_prepare(const Ctx& c, Factor& f, Screen& sc, double sigma, hipStream_t s) {
  const int Kp = c.K - 1, ld = c.ld;
  sc.Kp = Kp;
  sc.sigma = sigma;
  int rc = ldlh_factor(f, s, c.H, c.S, sigma);
  if (rc) return rc;
  // Z = L'^{-1}: the leading Kp x Kp block of the factorisation's own
  // X = L^{-1} (X^H in F's strict upper triangle)
  hipLaunchKernelGGL(k_z_from_xh, dim3((Kp + 63) / 64, (Kp + 63) / 64), dim3(256), 0, s, Kp, ld, f.F, sc.Z);
  // G = Z^H (D^{-1} Z), P = G S'
  hipLaunchKernelGGL(k_row_scale, dim3((Kp + 255) / 256, Kp), dim3(256), 0, s, Kp, ld, f.D, sc.Z, sc.P);
  rc = zgemm(1, Kp, Kp, Kp, mkz(1.0, 0.0), sc.Z, ld, sc.P, ld, mkz(0.0, 0.0), sc.G, ld, s, 64, 1);
  if (!rc) rc = zgemm(0, Kp, Kp, Kp, mkz(1.0, 0.0), sc.G, ld, c.S, ld, mkz(0.0, 0.0), sc.P, ld, s);
  return rc;
}

// ------------------------------------------------- per-candidate kernels
// per-candidate scalars cs[8 j + .]
enum { CS_E = 0, CS_ND = 1, CS_DELTA = 2, CS_XSX = 3, CS_MU = 4, CS_HKK = 5 };

__device__ inline double wg_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
  return t;
}
__device__ inline double wg_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t = fmax(t, sh[i]);
  return t;
}

// loop mode: candidates of the current step into cparams
__global__ void k_cand_fetch(int n, const double* __restrict__ call, const double* __restrict__ state,
                             double* __restrict__ cparams) {
  const size_t step = (size_t)state[1];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    cparams[i] = call[step * n + i];
}

// m = h - sigma s, mu = H_KK - sigma; start vectors
__global__ void __launch_bounds__(256) k_cand_border(int Kp, int ld, int C, double sigma,
                                                     const zd* __restrict__ Hb, zd* __restrict__ Bm,
                                                     const double* __restrict__ cdiag, double* __restrict__ cs,
                                                     const zd* __restrict__ v0, zd* __restrict__ V,
                                                     zd* __restrict__ vK) {
  const int j = blockIdx.x;
  const zd* h = Hb + (size_t)j * ld;
  const zd* sv = Bm + (size_t)(C + j) * ld;
  zd* m = Bm + (size_t)j * ld;
  zd* v = V + (size_t)j * ld;
  for (int l = threadIdx.x; l < Kp; l += blockDim.x) {
    m[l] = zfms(mkz(sigma, 0.0), sv[l], h[l]);
    v[l] = v0[l];
  }
  if (threadIdx.x == 0) {
    cs[8 * j + CS_HKK] = cdiag[2 * j + 1];
    cs[8 * j + CS_MU] = cdiag[2 * j + 1] - sigma;
    vK[j] = v0[Kp];
  }
}

// delta_j = mu_j - Re(m_j^H w_m,j)
__global__ void __launch_bounds__(256) k_cand_schur(int Kp, int ld, const zd* __restrict__ Bm,
                                                    const zd* __restrict__ Wm, double* __restrict__ cs) {
  __shared__ double sh[8];
  const int j = blockIdx.x;
  const zd* m = Bm + (size_t)j * ld;
  const zd* w = Wm + (size_t)j * ld;
  double a = 0.0;
  for (int l = threadIdx.x; l < Kp; l += blockDim.x) a += m[l].x * w[l].x + m[l].y * w[l].y;
  a = wg_sum(a, sh);
  if (threadIdx.x == 0) cs[8 * j + CS_DELTA] = cs[8 * j + CS_MU] - a;
}

// One bordered inverse-iteration update per candidate (U = P V on entry):
// x = M_j^{-1} S_j v, scaled by max(|Re|, |Im|), GHEPIIS's convergence
// measure sqrt(2K) ||x - tc v||, tc = x^H v / v^H v; V, vK := x.
__global__ void __launch_bounds__(256) k_cand_iter(int Kp, int ld, int C, zd* __restrict__ U,
                                                   const zd* __restrict__ Wm, const zd* __restrict__ Bm,
                                                   zd* __restrict__ V, zd* __restrict__ vK,
                                                   double* __restrict__ cs) {
  __shared__ double sh[8];
  const int j = blockIdx.x;
  const zd* m = Bm + (size_t)j * ld;
  const zd* sv = Bm + (size_t)(C + j) * ld;
  const zd* wm = Wm + (size_t)j * ld;
  const zd* ws = Wm + (size_t)(C + j) * ld;
  zd* u = U + (size_t)j * ld;
  zd* v = V + (size_t)j * ld;
  const zd vk = vK[j];
  // r_K = s^H v' + v_K;  u = P v' + w_s v_K;  m^H u
  double rr = 0.0, ri = 0.0, mr = 0.0, mi = 0.0;
  for (int l = threadIdx.x; l < Kp; l += blockDim.x) {
    const zd vl = v[l], sl = sv[l];
    rr += sl.x * vl.x + sl.y * vl.y;
    ri += sl.x * vl.y - sl.y * vl.x;
    const zd ul = zfma(ws[l], vk, u[l]);
    u[l] = ul;
    const zd ml = m[l];
    mr += ml.x * ul.x + ml.y * ul.y;
    mi += ml.x * ul.y - ml.y * ul.x;
  }
  rr = wg_sum(rr, sh);
  ri = wg_sum(ri, sh);
  mr = wg_sum(mr, sh);
  mi = wg_sum(mi, sh);
  const double delta = cs[8 * j + CS_DELTA];
  const zd xk = mkz((rr + vk.x - mr) / delta, (ri + vk.y - mi) / delta);
  double mx = fmax(fabs(xk.x), fabs(xk.y));
  for (int l = threadIdx.x; l < Kp; l += blockDim.x) {
    const zd xl = zfms(wm[l], xk, u[l]);
    u[l] = xl;
    mx = fmax(mx, fmax(fabs(xl.x), fabs(xl.y)));
  }
  const double t1 = wg_max(mx, sh);
  const double inv = 1.0 / t1;
  // tc = (x^H v)/(v^H v) over all K entries
  double ar = 0.0, ai = 0.0, vv = 0.0;
  for (int l = threadIdx.x; l < Kp; l += blockDim.x) {
    const zd xl = u[l] * inv, vl = v[l];
    ar += xl.x * vl.x + xl.y * vl.y;
    ai += xl.x * vl.y - xl.y * vl.x;
    vv += vl.x * vl.x + vl.y * vl.y;
  }
  ar = wg_sum(ar, sh);
  ai = wg_sum(ai, sh);
  vv = wg_sum(vv, sh);
  const zd xks = xk * inv;
  ar += xks.x * vk.x + xks.y * vk.y;
  ai += xks.x * vk.y - xks.y * vk.x;
  vv += vk.x * vk.x + vk.y * vk.y;
  const zd tc = mkz(ar / vv, ai / vv);
  double nd = 0.0;
  for (int l = threadIdx.x; l < Kp; l += blockDim.x) {
    const zd xl = u[l] * inv;
    const zd d = zfms(tc, v[l], xl);
    nd += d.x * d.x + d.y * d.y;
    v[l] = xl;
  }
  nd = wg_sum(nd, sh);
  if (threadIdx.x == 0) {
    const zd d = zfms(tc, vk, xks);
    nd += d.x * d.x + d.y * d.y;
    cs[8 * j + CS_ND] = sqrt(2.0 * (Kp + 1)) * sqrt(nd);
    vK[j] = xks;
  }
}

// E_j = x^H H_j x / x^H S_j x with Y1 = S' x', Y2 = H' x'
__global__ void __launch_bounds__(256) k_cand_rayleigh(int Kp, int ld, int C, const zd* __restrict__ V,
                                                       const zd* __restrict__ vK, const zd* __restrict__ Y1,
                                                       const zd* __restrict__ Y2, const zd* __restrict__ Hb,
                                                       const zd* __restrict__ Bm, double* __restrict__ cs) {
  __shared__ double sh[8];
  const int j = blockIdx.x;
  const zd* x = V + (size_t)j * ld;
  const zd* y1 = Y1 + (size_t)j * ld;
  const zd* y2 = Y2 + (size_t)j * ld;
  const zd* h = Hb + (size_t)j * ld;
  const zd* sv = Bm + (size_t)(C + j) * ld;
  const zd xk = vK[j];
  double s1 = 0.0, h1 = 0.0, sb = 0.0, hb = 0.0;
  for (int l = threadIdx.x; l < Kp; l += blockDim.x) {
    const zd xl = x[l];
    s1 += xl.x * y1[l].x + xl.y * y1[l].y;
    h1 += xl.x * y2[l].x + xl.y * y2[l].y;
    // Re(conj(x_l) s_l x_K), Re(conj(x_l) h_l x_K)
    const zd sx = sv[l] * xk, hx = h[l] * xk;
    sb += xl.x * sx.x + xl.y * sx.y;
    hb += xl.x * hx.x + xl.y * hx.y;
  }
  s1 = wg_sum(s1, sh);
  h1 = wg_sum(h1, sh);
  sb = wg_sum(sb, sh);
  hb = wg_sum(hb, sh);
  if (threadIdx.x == 0) {
    const double xk2 = xk.x * xk.x + xk.y * xk.y;
    const double xsx = s1 + 2.0 * sb + xk2;
    const double xhx = h1 + 2.0 * hb + cs[8 * j + CS_HKK] * xk2;
    cs[8 * j + CS_E] = xhx / xsx;
    cs[8 * j + CS_XSX] = xsx;
  }
}

// Loop bookkeeping (one workgroup): best candidate of the step; when it
// lowers the current energy its parameters become function K and its
// eigenvector (normalised x^H S x = 1) the next start vector.
__global__ void __launch_bounds__(256) k_cand_select(int Kp, int ld, int C, int npt,
                                                     const double* __restrict__ cs,
                                                     const double* __restrict__ cparams,
                                                     double* __restrict__ params, const zd* __restrict__ V,
                                                     const zd* __restrict__ vK, zd* __restrict__ v0,
                                                     double* __restrict__ state, double* __restrict__ rec) {
  __shared__ double shE[256];
  __shared__ int shI[256];
  double e = INFINITY;
  int bi = -1;
  for (int j = threadIdx.x; j < C; j += blockDim.x) {
    const double ej = cs[8 * j + CS_E];
    if (ej < e) {   // NaN never wins
      e = ej;
      bi = j;
    }
  }
  shE[threadIdx.x] = e;
  shI[threadIdx.x] = bi;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const double e2 = shE[threadIdx.x + o];
      const int i2 = shI[threadIdx.x + o];
      if (i2 >= 0 && (e2 < shE[threadIdx.x] || (e2 == shE[threadIdx.x] && i2 < shI[threadIdx.x]))) {
        shE[threadIdx.x] = e2;
        shI[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  const double eb = shE[0];
  const int ib = shI[0];
  const double ecur = state[0];
  const bool accept = (ib >= 0) && (eb < ecur);
  if (accept) {
    for (int i = threadIdx.x; i < npt; i += blockDim.x) params[(size_t)Kp * npt + i] = cparams[(size_t)ib * npt + i];
    const double sc = 1.0 / sqrt(cs[8 * ib + CS_XSX]);
    for (int l = threadIdx.x; l < Kp; l += blockDim.x) v0[l] = V[(size_t)ib * ld + l] * sc;
    if (threadIdx.x == 0) v0[Kp] = vK[ib] * sc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int step = (int)state[1];
    rec[2 * step] = eb;
    rec[2 * step + 1] = accept ? (double)ib : -1.0 - (double)(ib < 0 ? 0 : ib);
    if (accept) state[0] = eb;
    state[1] = step + 1.0;
  }
}

int screen_batch(const Ctx& c, const Factor& f, Screen& sc, int C, int iters, bool from_loop, int npt,
                 hipStream_t s) {
  const int Kp = sc.Kp, ld = sc.ld;
  const zd one = mkz(1.0, 0.0), zero = mkz(0.0, 0.0);
  if (from_loop)
    hipLaunchKernelGGL(k_cand_fetch, dim3((npt * C + 255) / 256), dim3(256), 0, s, npt * C, sc.call, sc.state,
                       sc.cparams);
  int rc = candidate_rows(c, Kp, C, sc.cparams, sc.cbra, sc.cket, sc.cdiag, sc.Hb, sc.Bm + (size_t)C * ld, ld, sc.cpb,
                          s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cand_border, dim3(C), dim3(256), 0, s, Kp, ld, C, sc.sigma, sc.Hb, sc.Bm, sc.cdiag, sc.cs,
                     sc.v0, sc.V, sc.vK);
  rc = zgemm(0, Kp, 2 * C, Kp, one, sc.G, ld, sc.Bm, ld, zero, sc.Wm, ld, s, sc.tm);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cand_schur, dim3(C), dim3(256), 0, s, Kp, ld, sc.Bm, sc.Wm, sc.cs);
  for (int it = 0; it < iters; ++it) {
    rc = zgemm(0, Kp, C, Kp, one, sc.P, ld, sc.V, ld, zero, sc.U, ld, s, sc.tm);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cand_iter, dim3(C), dim3(256), 0, s, Kp, ld, C, sc.U, sc.Wm, sc.Bm, sc.V, sc.vK, sc.cs);
  }
  rc = zgemm(0, Kp, C, Kp, one, c.S, ld, sc.V, ld, zero, sc.Y1, ld, s, sc.tm);
  if (!rc) rc = zgemm(0, Kp, C, Kp, one, c.H, ld, sc.V, ld, zero, sc.Y2, ld, s, sc.tm);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cand_rayleigh, dim3(C), dim3(256), 0, s, Kp, ld, C, sc.V, sc.vK, sc.Y1, sc.Y2, sc.Hb, sc.Bm,
                     sc.cs);
  (void)f;
  return (int)hipGetLastError();
}

int screen_select(const Ctx& c, Screen& sc, int C, int npt, hipStream_t s) {
  hipLaunchKernelGGL(k_cand_select, dim3(1), dim3(256), 0, s, sc.Kp, sc.ld, C, npt, sc.cs, sc.cparams, c.params,
                     sc.V, sc.vK, sc.v0, sc.state, sc.rec);
  return (int)hipGetLastError();
}

}  // namespace ecg
