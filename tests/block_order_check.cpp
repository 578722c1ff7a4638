// CPU check of the batched walk kernels' block order (eclcc_internal.h):
// every grid size and window setting gives a permutation of the work items,
// a full window's items stay inside that window, and a window's labels fit
// its byte budget.  Built with hipcc (host code only) by tests/test_order.py.
#include <cstdio>
#include <vector>

#include "eclcc_internal.h"

using eclcc::BlockOrder;

static int check(unsigned grid, long long slots, long long wbytes) {
  const BlockOrder o = eclcc::block_order(grid, slots, wbytes);
  std::vector<unsigned char> seen(grid, 0);
  const unsigned full = o.wlog < 32u ? (grid >> o.wlog) << o.wlog : 0u;
  if (o.wlog < 32u && (4ll * slots) << o.wlog > wbytes) {
    std::printf("grid %u slots %lld window %lld: window of 2^%u items exceeds the budget\n", grid, slots, wbytes, o.wlog);
    return 1;
  }
  for (unsigned b = 0; b < grid; b++) {
    const unsigned it = eclcc::order_item(o, b, grid);
    if (it >= grid || seen[it]) {
      std::printf("grid %u slots %lld window %lld: block %u -> item %u %s\n", grid, slots, wbytes, b, it,
                  it >= grid ? "out of range" : "taken twice");
      return 1;
    }
    seen[it] = 1;
    if (b < full && (it >> o.wlog) != (b >> o.wlog)) {
      std::printf("grid %u: block %u -> item %u leaves its window\n", grid, b, it);
      return 1;
    }
  }
  return 0;
}

int main() {
  const long long slots[] = {1024, 2048, 8192, 64};
  const long long windows[] = {0, 32ll << 20, 1ll << 20, 8192, 4096};
  int bad = 0;
  for (unsigned grid = 1; grid <= 3000; grid++)
    for (long long s : slots)
      for (long long w : windows) bad += check(grid, s, w);
  const unsigned big[] = {4096, 4097, 12345, 65536, 65537, 301000, 1u << 20, (1u << 20) + 77};
  for (unsigned grid : big)
    for (long long s : slots)
      for (long long w : windows) bad += check(grid, s, w);
  // the mix's batched mark: 616,327,750 slots in 2048-slot items, 32 MB windows
  bad += check(300942, 2048, 32ll << 20);
  if (bad) return 1;
  std::printf("block order: all permutations\n");
  return 0;
}
