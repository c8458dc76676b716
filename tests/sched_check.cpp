// Host-side check of the trailing-update launches' task maps
// (atom-mol-nonbo_amd/csrc/ecg_schedule.hpp, the decode k_ldl_update runs):
// for every basis size K in [1, KMAX] and every launch of the blocked
// LDL^H, each flag a workgroup waits on must be published exactly once in
// the same launch, by a workgroup of lower index; the tasks must cover the
// trailing matrix's lower block pairs, the next panel's strips and the
// L^{-1} tiles exactly once.  Prints "OK <launches checked>" or the first
// violation and exits non-zero.  With a second argument "plant", quarter 0
// of every launch is made to wait on the look-ahead factor (a higher
// index): the check must then fail.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <vector>

#include "../atom-mol-nonbo_amd/csrc/ecg_schedule.hpp"

using namespace ecg;

int main(int argc, char** argv) {
  const int KMAX = argc > 1 ? std::atoi(argv[1]) : 4096;
  const bool plant = argc > 2;
  const int NB = SCHED_NB, SB = SCHED_SB;
  long launches = 0;
  for (int K = 1; K <= KMAX; ++K) {
    const int nb = (K + NB - 1) / NB;
    const int ld = (K + 63) / 64 * 64;
    const int qflag = ld / NB + 1;
    for (int j0 = 0, col = 0; j0 + NB < K; j0 += NB, ++col) {
      const UpdShape s = upd_shape(K, col);
      std::map<int, int> pub;   // flag -> publishing workgroup
      std::vector<UpdTask> tasks(s.grid);
      for (int w = 0; w < s.grid; ++w) {
        tasks[w] = upd_task(w, K, j0, s, qflag);
        if (plant && w == 0) tasks[w].wait[0] = 0;
        if (tasks[w].publish >= 0) {
          if (pub.count(tasks[w].publish)) {
            std::printf("K=%d col=%d: flag %d published twice\n", K, col, tasks[w].publish);
            return 1;
          }
          if (tasks[w].publish >= ld / NB + 5) {
            std::printf("K=%d col=%d: flag %d outside the flag array\n", K, col, tasks[w].publish);
            return 1;
          }
          pub[tasks[w].publish] = w;
        }
      }
      std::set<std::pair<int, int>> tiles, xtiles;
      std::set<int> strips;
      int quarters = 0, look = 0;
      for (int w = 0; w < s.grid; ++w) {
        const UpdTask& t = tasks[w];
        for (int i = 0; i < 4; ++i) {
          if (t.wait[i] < 0) continue;
          auto it = pub.find(t.wait[i]);
          if (it == pub.end()) {
            std::printf("K=%d col=%d: workgroup %d waits on flag %d that nobody publishes\n", K, col, w, t.wait[i]);
            return 1;
          }
          if (it->second >= w) {
            std::printf("K=%d col=%d: workgroup %d waits on workgroup %d (not lower)\n", K, col, w, it->second);
            return 1;
          }
        }
        switch (t.kind) {
          case UPD_QUARTER: ++quarters; break;
          case UPD_LOOKAHEAD: ++look; tiles.insert({0, 0}); break;
          case UPD_TILE: tiles.insert({t.a, t.b}); break;
          case UPD_XINV: xtiles.insert({t.a, t.b}); break;
          case UPD_STRIP: strips.insert(t.a); break;
        }
      }
      // coverage: lower block pairs (bi >= bk) of the trailing matrix, the
      // strips of the next panel, the L^{-1} tiles of block rows k.. by
      // block columns c < k (row k: all c < k; rows i > k: all c < k)
      const int t0 = col + 1, nrb = nb - t0;
      if (quarters != UPD_NQ || look != 1 || (int)tiles.size() != nrb * (nrb + 1) / 2) {
        std::printf("K=%d col=%d: tile cover wrong\n", K, col);
        return 1;
      }
      for (auto& p : tiles)
        if (p.first < p.second || p.first >= nrb) {
          std::printf("K=%d col=%d: bad tile (%d,%d)\n", K, col, p.first, p.second);
          return 1;
        }
      const int rest2 = K - (t0 + 1) * NB;
      if ((int)strips.size() != (rest2 > 0 ? (rest2 + SB - 1) / SB : 0)) {
        std::printf("K=%d col=%d: strip cover wrong\n", K, col);
        return 1;
      }
      for (int i0 : strips)
        if (i0 < (t0 + 1) * NB || i0 >= K || (i0 - (t0 + 1) * NB) % SB) {
          std::printf("K=%d col=%d: bad strip %d\n", K, col, i0);
          return 1;
        }
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
