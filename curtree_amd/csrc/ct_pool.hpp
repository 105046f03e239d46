// ct_pool.hpp -- the serving runtime's result pool: first-fit allocation of
// pair slots in one device buffer (the result budget), with neighbouring
// free blocks merged on release.  Host-only bookkeeping (offsets in pairs); no
// device calls, so it is unit-tested on the CPU.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <iterator>
#include <map>

namespace ct {

// First-fit allocator over the result pool (units: pairs).
struct Pool {
  int64_t cap = 0;
  std::map<int64_t, int64_t> free_;  // offset -> length
  void init(int64_t c) {
    cap = c;
    free_.clear();
    if (c > 0) free_[0] = c;
  }
  int64_t alloc(int64_t n) {
    if (n == 0) return 0;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        int64_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        return off;
      }
    }
    return -1;
  }
  int64_t largest() const {
    int64_t m = 0;
    for (auto& f : free_) m = std::max(m, f.second);
    return m;
  }
  void release(int64_t off, int64_t n) {
    if (n == 0) return;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

}  // namespace ct
