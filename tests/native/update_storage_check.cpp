// Brute-force check of the multifrontal plan's shared update storage (cells::share_update_storage):
// reads patterns (n, entry count, rows, columns, elimination order) from the files named on the
// command line, builds the supernodes with the lattice plan's front classes, shares the storage
// and tests every pair of supernodes whose update matrices are alive at a common level (written
// at the supernode's level, read at its parent's) for overlap. Prints one line per pattern:
// "n <n> fronts <ns> total <sum of update sizes> peak <shared storage> overlaps <pairs>".
#include "symbolic.h"

#include <cstdio>
#include <fstream>
#include <vector>

static int32_t front_class(int32_t rows) {
  const int64_t tri = (int64_t)rows * (rows + 1) / 2;
  return tri <= 16 ? 0 : tri <= 64 ? 1 : tri <= 128 ? 2 : tri <= 256 ? 3 : tri <= 512 ? 4 : tri <= 1280 ? 5
         : tri <= 5120 ? 6 : 7;
}

static int32_t widen_class(int32_t cls, int32_t level_fronts) {
  if (cls > 5 || level_fronts < 1) return cls;
  static const int32_t lanes[6] = {1, 4, 8, 16, 32, 64};
  int32_t wide = 0;
  for (int c = 5; c >= 0; --c)
    if ((int64_t)level_fronts * lanes[c] <= 256) {
      wide = c;
      break;
    }
  return cls > wide ? cls : wide;
}

int main(int argc, char** argv) {
  for (int a = 1; a < argc; ++a) {
    std::ifstream f(argv[a]);
    int32_t n;
    int64_t m;
    f >> n >> m;
    std::vector<int32_t> rows(m), cols(m), order(n);
    for (auto& x : rows) f >> x;
    for (auto& x : cols) f >> x;
    for (auto& x : order) f >> x;
    if (!f) {
      std::fprintf(stderr, "%s: unreadable pattern\n", argv[a]);
      return 2;
    }
    const cells::Symbolic s = cells::bind(n, m, rows.data(), cols.data(), order.data());
    cells::Supernodes sn = cells::supernodes(s, front_class, widen_class);
    const int64_t total = sn.update_total;
    cells::share_update_storage(sn);
    const int32_t ns = static_cast<int32_t>(sn.first.size());
    auto size = [&](int32_t t) {
      const int64_t S = sn.rows[t] - sn.width[t];
      return sn.parent[t] >= 0 ? S * (S + 1) / 2 : 0;
    };
    int64_t overlaps = 0;
    for (int32_t t = 0; t < ns; ++t) {
      if (size(t) == 0) continue;
      if (sn.update_offset[t] < 0 || sn.update_offset[t] + size(t) > sn.update_total) ++overlaps;
      for (int32_t u = t + 1; u < ns; ++u) {
        if (size(u) == 0) continue;
        const int32_t t0 = sn.level[t], t1 = sn.level[sn.parent[t]];
        const int32_t u0 = sn.level[u], u1 = sn.level[sn.parent[u]];
        if (t1 < u0 || u1 < t0) continue;
        const int64_t x0 = sn.update_offset[t], x1 = x0 + size(t);
        const int64_t y0 = sn.update_offset[u], y1 = y0 + size(u);
        if (x0 < y1 && y0 < x1) ++overlaps;
      }
    }
    std::printf("n %d fronts %d total %lld peak %lld overlaps %lld\n", n, ns, (long long)total,
                (long long)sn.update_total, (long long)overlaps);
  }
  return 0;
}
