// Task map of one trailing-update launch of the blocked LDL^H (k_ldl_update),
// shared by the kernel (which decodes its workgroup's task from it) and a
// host-side check (tests/sched_check.cpp) that enumerates every launch of
// every basis size and verifies the wait rule: each flag a workgroup waits
// on is published in the same launch by a workgroup of LOWER index.  The
// dispatcher places workgroups in index order, so the awaited one already
// holds a CU slot and, waiting only on lower indices itself, finishes; a
// wait on a higher index has no such guarantee (the waiting workgroups can
// fill every slot the awaited one needs).
#pragma once

#if defined(__HIPCC__)
#define ECG_HD __host__ __device__
#else
#define ECG_HD
#endif

namespace ecg {

constexpr int SCHED_NB = 64;   // block size of the factorisation
constexpr int SCHED_SB = 16;   // panel strip height
constexpr int UPD_NQ = 4;      // row quarters of the look-ahead tile's update

// Launch shape for block column col (diagonal block at j0 = 64 col):
// ntiles lower block pairs of the trailing matrix, nstrip 16-row strips of
// the next panel, nx tiles of X = L^{-1}; grid = NQ + ntiles + nstrip + nx.
struct UpdShape {
  int ntiles, nstrip, nx, grid;
};
ECG_HD inline UpdShape upd_shape(int K, int col) {
  const int NB = SCHED_NB, SB = SCHED_SB;
  const int j0 = col * NB;
  const int rest = K - (j0 + NB);   // rows below block column col
  const int nrb = (rest + NB - 1) / NB;
  UpdShape s;
  s.ntiles = nrb * (nrb + 1) / 2;
  const int rest2 = rest - NB;      // rows below the next block column
  s.nstrip = rest2 > 0 ? (rest2 + SB - 1) / SB : 0;
  const int nb = (K + NB - 1) / NB;
  s.nx = (col >= 1) ? col + (nb - 1 - col) * col : 0;
  s.grid = UPD_NQ + s.ntiles + s.nstrip + s.nx;
  return s;
}

// Flags of one launch: 0 = the look-ahead diagonal factor (X and D
// published), 1 + bi = column tile (bi, 0) of the trailing matrix updated,
// qflag + q = quarter q of the next diagonal tile updated.
enum UpdKind { UPD_QUARTER = 0, UPD_LOOKAHEAD = 1, UPD_TILE = 2, UPD_XINV = 3, UPD_STRIP = 4 };
struct UpdTask {
  int kind;
  int a, b;          // quarter: (q, -); tile: (bi, bk); L^{-1} tile: (block row i, block column c); strip: (i0, bi)
  int wait[4];       // flags awaited (-1: none)
  int publish;       // flag published (-1: none)
};

// Task of workgroup w of the launch for the diagonal block at j0.
ECG_HD inline UpdTask upd_task(int w, int K, int j0, const UpdShape& s, int qflag) {
  const int NB = SCHED_NB, SB = SCHED_SB;
  const int jb = (K - j0 < NB) ? K - j0 : NB;
  const int t0 = (j0 + jb) / NB;
  UpdTask t;
  t.a = t.b = 0;
  for (int i = 0; i < 4; ++i) t.wait[i] = -1;
  t.publish = -1;
  if (w < UPD_NQ) {
    t.kind = UPD_QUARTER;
    t.a = w;
    t.publish = qflag + w;
    return t;
  }
  if (w == UPD_NQ) {
    t.kind = UPD_LOOKAHEAD;
    for (int q = 0; q < UPD_NQ; ++q) t.wait[q] = qflag + q;
    t.publish = 0;
    return t;
  }
  if (w >= UPD_NQ + s.ntiles && w < UPD_NQ + s.ntiles + s.nx) {
    const int k = j0 / NB, x = w - UPD_NQ - s.ntiles;
    t.kind = UPD_XINV;
    t.a = (x < k) ? k : k + 1 + (x - k) / k;
    t.b = (x < k) ? x : (x - k) % k;
    return t;
  }
  if (w >= UPD_NQ + s.ntiles) {
    const int i0 = (t0 + 1) * NB + (w - UPD_NQ - s.ntiles - s.nx) * SB;
    t.kind = UPD_STRIP;
    t.a = i0;
    t.b = (i0 - t0 * NB) / NB;
    t.wait[0] = 0;
    t.wait[1] = 1 + t.b;
    return t;
  }
  // lower block pairs of the trailing matrix, numbered row by row; pair 0
  // (the next diagonal tile) is the look-ahead's
  const int tb = w - UPD_NQ;
  int bi = 0;
  while ((bi + 1) * (bi + 2) / 2 <= tb) ++bi;
  t.kind = UPD_TILE;
  t.a = bi;
  t.b = tb - bi * (bi + 1) / 2;
  if (t.b == 0) t.publish = 1 + bi;
  return t;
}

}  // namespace ecg
