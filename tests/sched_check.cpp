// Host-side check of the blocked LDL^H launches' task maps
// (atom-mol-nonbo_amd/csrc/ecg_schedule.hpp, the decode k_ldl_update and
// k_ldl_panel run): for every basis size K in [1, KMAX] and every launch,
// each flag a workgroup waits on must be published exactly once in the
// same launch, by a workgroup of lower index, inside the flag array; the
// tasks must cover the trailing matrix's lower block pairs (all but block
// t0's diagonal tile), the next panel's strips, the diagonal strips' four
// quarters of the next diagonal tile and the L^{-1} tiles exactly once.
// Prints "OK <launches checked>" or the first violation and exits
// non-zero.  With a second argument "plant", the look-ahead of every
// update launch is made to wait on column tile (1, 0) (a higher index):
// the check must then fail.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "../atom-mol-nonbo_amd/csrc/ecg_schedule.hpp"

using namespace ecg;

static int check_waits(int K, int col, const std::vector<UpdTask>& tasks, int nflags) {
  std::map<int, int> pub;   // flag -> publishing workgroup
  for (int w = 0; w < (int)tasks.size(); ++w) {
    const int f = tasks[w].publish;
    if (f < 0) continue;
    if (pub.count(f)) {
      std::printf("K=%d col=%d: flag %d published twice\n", K, col, f);
      return 1;
    }
    if (f >= nflags) {
      std::printf("K=%d col=%d: flag %d outside the flag array\n", K, col, f);
      return 1;
    }
    pub[f] = w;
  }
  for (int w = 0; w < (int)tasks.size(); ++w)
    for (int i = 0; i < UPD_WAITS; ++i) {
      const int f = tasks[w].wait[i];
      if (f < 0) continue;
      auto it = pub.find(f);
      if (it == pub.end()) {
        std::printf("K=%d col=%d: workgroup %d waits on flag %d that nobody publishes\n", K, col, w, f);
        return 1;
      }
      if (it->second >= w) {
        std::printf("K=%d col=%d: workgroup %d waits on workgroup %d (not lower)\n", K, col, w, it->second);
        return 1;
      }
    }
  return 0;
}

// the strips of a panel whose rows start at i1: one per 16 rows below, the
// first NDIAG (block i1's rows) diagonal with quarters 0, 1, ...
static int check_strips(int K, int col, const std::vector<UpdTask>& tasks, int i1) {
  const int SB = SCHED_SB;
  std::set<int> strips, quarters;
  for (const UpdTask& t : tasks) {
    if (t.kind != UPD_STRIP) continue;
    strips.insert(t.a);
    if (t.a < i1 || t.a >= K || (t.a - i1) % SB) {
      std::printf("K=%d col=%d: bad strip %d\n", K, col, t.a);
      return 1;
    }
    const bool diag = t.a < i1 + SCHED_NB;
    if (diag != (t.diag >= 0) || (diag && t.diag != (t.a - i1) / SB)) {
      std::printf("K=%d col=%d: strip %d has diagonal quarter %d\n", K, col, t.a, t.diag);
      return 1;
    }
    if (diag) quarters.insert(t.diag);
  }
  const int rest = K - i1;
  const int want = rest > 0 ? (rest + SB - 1) / SB : 0;
  if ((int)strips.size() != want) {
    std::printf("K=%d col=%d: strip cover wrong\n", K, col);
    return 1;
  }
  const int wantq = want < NDIAG ? want : NDIAG;
  if ((int)quarters.size() != wantq) {
    std::printf("K=%d col=%d: diagonal quarter cover wrong\n", K, col);
    return 1;
  }
  return 0;
}

// Batched launches (k_ldl_panel_b, k_ldl_update_b): workgroup id serves
// problem id % P, task w = id / P, i.e. id = w * P + p.  For seeded sets of uneven
// basis sizes, every wait must name a flag of the same problem published
// in the same launch at a lower linear index.  With plant, the look-ahead
// of problem 0 waits on problem 1's look-ahead (another problem's flag).
static int check_batches(int nsets, bool plant) {
  const int NB = SCHED_NB;
  unsigned long long st = 20261204ull;
  long launches = 0;
  for (int set = 0; set < nsets; ++set) {
    const int P = 1 + (int)(st % 40);
    std::vector<int> Ks(P);
    for (int p = 0; p < P; ++p) {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      Ks[p] = 1 + (int)((st >> 33) % 2600);
    }
    std::sort(Ks.begin(), Ks.end(), [](int a, int b) { return a > b; });   // the launches' order
    const int maxK = Ks[0];
    for (int col = -1; col * NB + NB < maxK; ++col) {
      // (problem, task) of every workgroup, indexed by linear id
      int gy = 0;
      for (int K : Ks) {
        if (col < 0 && K > NB) gy = std::max(gy, (K - NB + SCHED_SB - 1) / SCHED_SB);
        if (col >= 0 && col * NB + NB < K) gy = std::max(gy, upd_shape(K, col).grid);
      }
      std::vector<std::vector<UpdTask>> tasks(P);
      std::map<std::pair<int, int>, long> pub;   // (problem, flag) -> linear id
      // the launch covers the active prefix of the list only
      int PA = 0;
      while (PA < P && Ks[PA] > (col < 0 ? NB : col * NB + NB)) ++PA;
      for (int p = 0; p < P; ++p) {
        const int K = Ks[p], qflag = (K + 63) / 64 * 64 / NB + 1;
        if (col < 0) {
          if (K > NB)
            for (int w = 0; w < (K - NB + SCHED_SB - 1) / SCHED_SB; ++w) tasks[p].push_back(panel_task(w, qflag));
        } else if (col * NB + NB < K) {
          const UpdShape s = upd_shape(K, col);
          for (int w = 0; w < s.grid; ++w) tasks[p].push_back(upd_task(w, K, col * NB, s, qflag));
        }
        if ((int)tasks[p].size() > gy) {
          std::printf("set %d col %d: problem %d has more tasks than the grid\n", set, col, p);
          return 1;
        }
        for (int w = 0; w < (int)tasks[p].size(); ++w)
          if (tasks[p][w].publish >= 0) pub[{p, tasks[p][w].publish}] = (long)w * PA + p;
      }
      for (int p = 0; p < P; ++p)
        for (int w = 0; w < (int)tasks[p].size(); ++w) {
          UpdTask t = tasks[p][w];
          int owner = p;
          if (plant && p == 0 && w == 0 && P > 1 && !tasks[1].empty() && tasks[1][0].publish >= 0) {
            owner = 1;
            t.wait[0] = tasks[1][0].publish;
          }
          for (int i = 0; i < UPD_WAITS; ++i) {
            const int f = t.wait[i];
            if (f < 0) continue;
            auto it = pub.find({owner, f});
            if (it == pub.end()) {
              std::printf("set %d col %d: problem %d task %d waits on an unpublished flag %d\n", set, col, p, w, f);
              return 1;
            }
            if (owner != p) {
              std::printf("set %d col %d: problem %d task %d waits on problem %d (not its own)\n", set, col, p,
                          w, owner);
              return 1;
            }
            if (p >= PA || it->second >= (long)w * PA + p) {
              std::printf("set %d col %d: problem %d task %d waits on linear id %ld (not lower)\n", set, col, p,
                          w, it->second);
              return 1;
            }
          }
        }
      ++launches;
    }
    st = st * 6364136223846793005ull + 1442695040888963407ull;
  }
  std::printf("OK %ld\n", launches);
  return 0;
}

int main(int argc, char** argv) {
  if (argc > 1 && std::string(argv[1]) == "batch")
    return check_batches(argc > 2 ? std::atoi(argv[2]) : 200, argc > 3);
  const int KMAX = argc > 1 ? std::atoi(argv[1]) : 4096;
  const bool plant = argc > 2;
  const int NB = SCHED_NB;
  long launches = 0;
  for (int K = 1; K <= KMAX; ++K) {
    const int nb = (K + NB - 1) / NB;
    const int ld = (K + 63) / 64 * 64;
    const int qflag = ld / NB + 1;
    const int nflags = sched_flags(ld);
    // the first panel's launch (strips below block 0)
    if (K > NB) {
      const int ns = (K - NB + SCHED_SB - 1) / SCHED_SB;
      std::vector<UpdTask> tasks(ns);
      for (int w = 0; w < ns; ++w) tasks[w] = panel_task(w, qflag);
      if (check_waits(K, -1, tasks, nflags) || check_strips(K, -1, tasks, NB)) return 1;
      ++launches;
    }
    for (int j0 = 0, col = 0; j0 + NB < K; j0 += NB, ++col) {
      const UpdShape s = upd_shape(K, col);
      std::vector<UpdTask> tasks(s.grid);
      for (int w = 0; w < s.grid; ++w) tasks[w] = upd_task(w, K, j0, s, qflag);
      if (plant && s.ntiles >= 1) tasks[0].wait[0] = 1 + 1;
      if (check_waits(K, col, tasks, nflags)) return 1;
      std::set<std::pair<int, int>> tiles, xtiles;
      int look = 0;
      for (const UpdTask& t : tasks) {
        switch (t.kind) {
          case UPD_LOOKAHEAD: ++look; break;
          case UPD_TILE: tiles.insert({t.a, t.b}); break;
          case UPD_XINV: xtiles.insert({t.a, t.b}); break;
          default: break;
        }
      }
      // coverage: lower block pairs (bi >= bk) of the trailing matrix but
      // (0, 0), the strips of the next panel, the L^{-1} tiles of block
      // rows k.. by block columns c < k (row k: all c < k; rows i > k: all
      // c < k)
      const int t0 = col + 1, nrb = nb - t0;
      if (look != 1 || (int)tiles.size() != nrb * (nrb + 1) / 2 - 1 || tiles.count({0, 0})) {
        std::printf("K=%d col=%d: tile cover wrong\n", K, col);
        return 1;
      }
      for (auto& p : tiles)
        if (p.first < p.second || p.first >= nrb) {
          std::printf("K=%d col=%d: bad tile (%d,%d)\n", K, col, p.first, p.second);
          return 1;
        }
      if (check_strips(K, col, tasks, (t0 + 1) * NB)) return 1;
      const int k = col;
      std::set<std::pair<int, int>> want;
      if (k >= 1)
        for (int i = k; i < nb; ++i)
          for (int c = 0; c < k; ++c) want.insert({i, c});
      if (xtiles != want) {
        std::printf("K=%d col=%d: L^-1 tile cover wrong (%zu vs %zu)\n", K, col, xtiles.size(), want.size());
        return 1;
      }
      ++launches;
    }
  }
  std::printf("OK %ld\n", launches);
  return 0;
}
