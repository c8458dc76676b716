// Task maps of the blocked LDL^H's launches, shared by the kernels (which
// decode their workgroup's task from them) and a host-side check
// (tests/sched_check.cpp) that enumerates every launch of every basis size
// and verifies the wait rule: each flag a workgroup waits on is published
// in the same launch by a workgroup of LOWER index.  The dispatcher places
// workgroups in index order, so the awaited one already holds a CU slot
// and, waiting only on lower indices itself, finishes; a wait on a higher
// index has no such guarantee (the waiting workgroups can fill every slot
// the awaited one needs).
//
// One trailing-update launch per block column col (diagonal block at
// j0 = 64 col, next block t0 = col + 1):
//   workgroup 0: the look-ahead factor of block t0, whose diagonal tile is
//     complete when the launch starts (the previous launch's diagonal
//     strips applied the last panel to it);
//   then the lower block pairs (bi, bk) of the trailing matrix from block
//     t0 (pair (0, 0), block t0's own tile, excluded) updated by panel col;
//   then the tiles of X = L^{-1} (no waits);
//   last, the 16-row strips of panel t0 below block t0.  The first four
//     (the rows of block t0 + 1, "diagonal strips") also apply panel t0 to
//     the diagonal tile of block t0 + 1, quarter by quarter, so that the
//     next launch's look-ahead can start at once.
// The first panel (below block 0) is its own launch of strips only, with
// the same diagonal strips.
#pragma once

#if defined(__HIPCC__)
#define ECG_HD __host__ __device__
#else
#define ECG_HD
#endif

namespace ecg {

constexpr int SCHED_NB = 64;   // block size of the factorisation
constexpr int SCHED_SB = 16;   // panel strip height
constexpr int NDIAG = SCHED_NB / SCHED_SB;   // diagonal strips per block
constexpr int UPD_WAITS = 8;   // flags one task can wait on

// Flags (entries of the flag array, qflag = ld / 64 + 1):
//   0            look-ahead factor published (X, D of block t0)
//   1 + bi       column tile (bi, 0) of the trailing matrix updated, bi >= 1
//   qflag + q    diagonal strip q stored its rows of L (q < 4)
//   qflag + 4    tile (1, 1) (block t0 + 1's diagonal tile) updated by panel col
// The array holds qflag + 5 = ld / 64 + 6 entries.
ECG_HD inline int sched_flags(int ld) { return ld / SCHED_NB + 6; }

// Launch shape for block column col: ntiles lower block pairs of the
// trailing matrix besides block t0's own tile, nstrip 16-row strips of the
// next panel, nx tiles of X = L^{-1}; grid = 1 + ntiles + nx + nstrip.
struct UpdShape {
  int ntiles, nstrip, nx, grid;
};
ECG_HD inline UpdShape upd_shape(int K, int col) {
  const int NB = SCHED_NB, SB = SCHED_SB;
  const int j0 = col * NB;
  const int rest = K - (j0 + NB);   // rows below block column col
  const int nrb = (rest + NB - 1) / NB;
  UpdShape s;
  s.ntiles = nrb * (nrb + 1) / 2 - 1;
  const int rest2 = rest - NB;      // rows below the next block column
  s.nstrip = rest2 > 0 ? (rest2 + SB - 1) / SB : 0;
  const int nb = (K + NB - 1) / NB;
  s.nx = (col >= 1) ? col + (nb - 1 - col) * col : 0;
  s.grid = 1 + s.ntiles + s.nx + s.nstrip;
  return s;
}

enum UpdKind { UPD_LOOKAHEAD = 1, UPD_TILE = 2, UPD_XINV = 3, UPD_STRIP = 4 };
struct UpdTask {
  int kind;
  int a, b;              // tile: (bi, bk); L^{-1} tile: (block row i, block column c); strip: (i0, bi)
  int diag;              // strip: quarter q of block t0 + 1 it also updates, or -1
  int wait[UPD_WAITS];   // flags awaited (-1: none)
  int publish;           // flag published (-1: none)
};

ECG_HD inline void task_clear(UpdTask& t) {
  t.a = t.b = 0;
  t.diag = -1;
  for (int i = 0; i < UPD_WAITS; ++i) t.wait[i] = -1;
  t.publish = -1;
}

// Strip s of a panel whose rows start at i1 (the block after the panel's
// diagonal block): the first NDIAG strips are diagonal strips; strip q of
// them waits for the lower diagonal strips' rows of L (and, in an update
// launch, for tile (1, 1) and the column tile it reads).
ECG_HD inline void strip_diag_waits(UpdTask& t, int q, int qflag, int first) {
  int n = first;
  for (int u = 0; u < q; ++u) t.wait[n++] = qflag + u;
  t.diag = q;
  t.publish = qflag + q;
}

// Task of workgroup w of the update launch for the diagonal block at j0.
ECG_HD inline UpdTask upd_task(int w, int K, int j0, const UpdShape& s, int qflag) {
  const int NB = SCHED_NB, SB = SCHED_SB;
  const int jb = (K - j0 < NB) ? K - j0 : NB;
  const int t0 = (j0 + jb) / NB;
  UpdTask t;
  task_clear(t);
  if (w == 0) {
    t.kind = UPD_LOOKAHEAD;
    t.publish = 0;
    return t;
  }
  if (w <= s.ntiles) {
    // lower block pairs of the trailing matrix, numbered row by row; pair
    // 0 (block t0's diagonal tile) is complete already
    const int tb = w;
    int bi = 0;
    while ((bi + 1) * (bi + 2) / 2 <= tb) ++bi;
    t.kind = UPD_TILE;
    t.a = bi;
    t.b = tb - bi * (bi + 1) / 2;
    if (t.b == 0) t.publish = 1 + bi;
    else if (bi == 1 && t.b == 1) t.publish = qflag + 4;
    return t;
  }
  if (w <= s.ntiles + s.nx) {
    const int k = j0 / NB, x = w - 1 - s.ntiles;
    t.kind = UPD_XINV;
    t.a = (x < k) ? k : k + 1 + (x - k) / k;
    t.b = (x < k) ? x : (x - k) % k;
    return t;
  }
  const int sidx = w - 1 - s.ntiles - s.nx;
  const int i0 = (t0 + 1) * NB + sidx * SB;
  t.kind = UPD_STRIP;
  t.a = i0;
  t.b = (i0 - t0 * NB) / NB;
  t.wait[0] = 0;
  t.wait[1] = 1 + t.b;
  if (sidx < NDIAG) {
    t.wait[2] = qflag + 4;
    strip_diag_waits(t, sidx, qflag, 3);
  }
  return t;
}

// Strip s of the first panel's launch (rows NB + 16 s; k_ldl_panel).
ECG_HD inline UpdTask panel_task(int s, int qflag) {
  UpdTask t;
  task_clear(t);
  t.kind = UPD_STRIP;
  t.a = SCHED_NB + s * SCHED_SB;
  t.b = 1;
  if (s < NDIAG) strip_diag_waits(t, s, qflag, 0);
  return t;
}

}  // namespace ecg
