// eclcc_arena.h -- the server's first-fit allocator of offsets in one
// contiguous region (its device arena and its pinned staging arena).
// Host-only C++; no HIP types, so it is also built and tested on its own.
#ifndef ECLCC_ARENA_H
#define ECLCC_ARENA_H

#include <algorithm>
#include <cstddef>
#include <map>
#include <utility>

namespace eclcc {

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// First-fit allocator over one contiguous region (offsets only).  It runs
// on the worker thread, so neither alloc nor release may throw: a block is
// cut from the end of a free range (the range keeps its key), and a freed
// range joins a neighbour by moving that neighbour's node (extract and
// insert of a node allocate nothing); only a range with no free neighbour
// needs a new node, and if even that fails the range is left out of the
// free list rather than throwing.
class Arena {
 public:
  void reset(size_t bytes) {
    free_.clear();
    if (bytes) free_[0] = bytes;
    cap_ = bytes;
    used_ = 0;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = align256(std::max<size_t>(bytes, 256));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        it->second -= bytes;
        *off = it->first + it->second;
        if (it->second == 0) free_.erase(it);
        used_ += bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = align256(std::max<size_t>(bytes, 256));
    used_ -= bytes;
    auto nx = free_.lower_bound(off);  // first free range after the block
    if (nx != free_.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {  // extend the range before it
        pv->second += bytes;
        if (nx != free_.end() && pv->first + pv->second == nx->first) {
          pv->second += nx->second;
          free_.erase(nx);
        }
        return;
      }
    }
    if (nx != free_.end() && off + bytes == nx->first) {  // the range after it moves down
      auto node = free_.extract(nx);
      node.key() = off;
      node.mapped() += bytes;
      free_.insert(std::move(node));
      return;
    }
    try {
      free_.emplace(off, bytes);
    } catch (...) {
      // out of host memory: the range stays unused until reset
    }
  }
  size_t used() const { return used_; }
  size_t capacity() const { return cap_; }

 private:
  std::map<size_t, size_t> free_;
  size_t cap_ = 0, used_ = 0;
};

}  // namespace eclcc

#endif
