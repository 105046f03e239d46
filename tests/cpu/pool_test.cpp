// CPU unit test of the serving runtime's result pool (curtree_amd/csrc/ct_pool.hpp):
// a random sequence of allocations and releases against a byte-map model.
// Checks: every allocation lies inside the pool and overlaps no live one;
// first fit (the lowest free offset that holds the request); released space is
// merged so that the free list never holds two adjacent blocks; largest() is the
// largest free run; after everything is released the pool is one block again.
#include <stdio.h>
#include <stdlib.h>

#include <random>
#include <vector>

#include "../../curtree_amd/csrc/ct_pool.hpp"

static int fail(const char* what, long a, long b) {
  printf("FAIL %s (%ld, %ld)\n", what, a, b);
  return 1;
}

int main() {
  const int64_t cap = 4096;
  ct::Pool p;
  p.init(cap);
  std::vector<char> used(cap, 0);
  struct Live { int64_t off, n; };
  std::vector<Live> live;
  std::mt19937_64 rng(12345);
  for (int step = 0; step < 200000; step++) {
    const bool do_alloc = live.empty() || (rng() % 100) < 55;
    if (do_alloc) {
      const int64_t n = (int64_t)(rng() % 3 == 0 ? rng() % 600 : rng() % 40) + 1;
      // the model's first fit
      int64_t want = -1;
      for (int64_t o = 0, run = 0; o < cap; o++) {
        run = used[o] ? 0 : run + 1;
        if (run == n) { want = o - n + 1; break; }
      }
      const int64_t got = p.alloc(n);
      if (got != want) return fail("first fit", (long)got, (long)want);
      if (got >= 0) {
        for (int64_t o = got; o < got + n; o++) used[o] = 1;
        live.push_back({got, n});
      }
    } else {
      const size_t k = rng() % live.size();
      p.release(live[k].off, live[k].n);
      for (int64_t o = live[k].off; o < live[k].off + live[k].n; o++) used[o] = 0;
      live[k] = live.back();
      live.pop_back();
    }
    // free list: sorted, disjoint, never adjacent, equal to the model's free runs
    int64_t prev_end = -1, largest = 0;
    for (auto& f : p.free_) {
      if (f.second <= 0) return fail("empty block", (long)f.first, (long)f.second);
      if (f.first <= prev_end) return fail("adjacent or overlapping blocks", (long)f.first, (long)prev_end);
      for (int64_t o = f.first; o < f.first + f.second; o++)
        if (used[o]) return fail("free block over a live one", (long)o, 0);
      if (f.first > 0 && !used[f.first - 1]) return fail("unmerged on the left", (long)f.first, 0);
      if (f.first + f.second < cap && !used[f.first + f.second]) return fail("unmerged on the right", (long)f.first, 0);
      prev_end = f.first + f.second;
      if (f.second > largest) largest = f.second;
    }
    if (p.largest() != largest) return fail("largest", (long)p.largest(), (long)largest);
  }
  for (auto& l : live) p.release(l.off, l.n);
  if (p.free_.size() != 1 || p.free_.begin()->first != 0 || p.free_.begin()->second != cap)
    return fail("not one block after releasing everything", (long)p.free_.size(), 0);
  if (p.alloc(0) != 0 || p.alloc(cap + 1) != -1) return fail("edge sizes", 0, 0);
  printf("pool ok\n");
  return 0;
}
