// Randomised check of the server's arena allocator (eclcc_arena.h) against a
// map of 256-byte units: no two live blocks overlap, every block lies inside
// the region, alloc fails only when no free run is large enough (so freed
// ranges always coalesce), and after everything is released the whole region
// is one free range again.  Built and run by tests/test_arena.py.
#include <cstdio>
#include <random>
#include <vector>

#include "eclcc_arena.h"

using eclcc::align256;
using eclcc::Arena;

int main() {
  std::mt19937_64 rng(7);
  for (int trial = 0; trial < 200; trial++) {
    Arena a;
    const size_t cap = 256 * (1 + rng() % 4000);
    a.reset(cap);
    std::vector<std::pair<size_t, size_t>> live;  // offset, aligned bytes
    std::vector<char> own(cap / 256, 0);
    for (int step = 0; step < 5000; step++) {
      if (live.empty() || rng() % 2) {
        const size_t want = 1 + rng() % (cap / 8 + 1);
        const size_t b = align256(std::max<size_t>(want, 256));
        size_t off = 0;
        if (a.alloc(want, &off)) {
          if (off % 256 || off + b > cap) return std::printf("block outside the region\n"), 1;
          for (size_t k = off / 256; k < (off + b) / 256; k++) {
            if (own[k]) return std::printf("blocks overlap\n"), 1;
            own[k] = 1;
          }
          live.push_back({off, b});
        } else {
          size_t run = 0, best = 0;
          for (char o : own) {
            run = o ? 0 : run + 1;
            best = std::max(best, run);
          }
          if (best * 256 >= b) return std::printf("free run of %zu bytes missed\n", best * 256), 1;
        }
      } else {
        const size_t i = rng() % live.size();
        const auto [off, b] = live[i];
        a.release(off, b);
        for (size_t k = off / 256; k < (off + b) / 256; k++) own[k] = 0;
        live[i] = live.back();
        live.pop_back();
      }
      size_t used = 0;
      for (const auto& l : live) used += l.second;
      if (used != a.used()) return std::printf("used() disagrees\n"), 1;
    }
    for (const auto& l : live) a.release(l.first, l.second);
    size_t off = 1;
    if (!a.alloc(cap, &off) || off != 0) return std::printf("region not whole after release\n"), 1;
  }
  std::printf("arena ok\n");
  return 0;
}
