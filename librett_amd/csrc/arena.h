// Offset allocator behind the serving engine's device arena: best fit over
// the free ranges (indexed by size, so one allocation is O(log n)), and
// free ranges merged with their neighbours on release. Offsets and sizes
// are bytes in multiples of kGranule. Thread-safe. Host code only (no HIP),
// so it is built and tested on its own as well.
#pragma once

#include <algorithm>
#include <cstddef>
#include <iterator>
#include <map>
#include <mutex>
#include <set>
#include <unordered_map>
#include <utility>

namespace lrt {

class RangeAllocator {
 public:
  static constexpr size_t kGranule = 256;
  static constexpr size_t kNone = ~(size_t)0;

  // Forgets every range and makes [0, capacity rounded up) one free range.
  void reset(size_t capacity) {
    std::lock_guard<std::mutex> lock(mu_);
    cap_ = (capacity + kGranule - 1) / kGranule * kGranule;
    by_off_.clear();
    by_size_.clear();
    used_.clear();
    in_use_ = peak_ = 0;
    if (cap_) add_free(0, cap_);
  }

  // Offset of a new range of at least `bytes` (one granule at least), or
  // kNone when no free range is large enough. Among the smallest free ranges
  // that fit, the lowest offset is taken.
  size_t alloc(size_t bytes) {
    bytes = std::max(kGranule, (bytes + kGranule - 1) / kGranule * kGranule);
    std::lock_guard<std::mutex> lock(mu_);
    auto it = by_size_.lower_bound({bytes, 0});
    if (it == by_size_.end()) return kNone;
    const size_t size = it->first, off = it->second;
    by_size_.erase(it);
    by_off_.erase(off);
    if (size > bytes) add_free(off + bytes, size - bytes);
    used_[off] = bytes;
    in_use_ += bytes;
    peak_ = std::max(peak_, in_use_);
    return off;
  }

  // Returns the range that starts at `off`; false if none does.
  bool release(size_t off) {
    std::lock_guard<std::mutex> lock(mu_);
    auto u = used_.find(off);
    if (u == used_.end()) return false;
    size_t size = u->second;
    used_.erase(u);
    in_use_ -= size;
    size_t start = off;
    auto next = by_off_.find(off + size);
    if (next != by_off_.end()) {
      by_size_.erase({next->second, next->first});
      size += next->second;
      by_off_.erase(next);
    }
    auto after = by_off_.lower_bound(off);
    if (after != by_off_.begin()) {
      auto prev = std::prev(after);
      if (prev->first + prev->second == start) {
        by_size_.erase({prev->second, prev->first});
        start = prev->first;
        size += prev->second;
        by_off_.erase(prev);
      }
    }
    add_free(start, size);
    return true;
  }

  size_t in_use() {
    std::lock_guard<std::mutex> lock(mu_);
    return in_use_;
  }
  size_t peak() {
    std::lock_guard<std::mutex> lock(mu_);
    return peak_;
  }
  // Restarts the peak at the bytes in use now.
  void reset_peak() {
    std::lock_guard<std::mutex> lock(mu_);
    peak_ = in_use_;
  }
  // Free ranges (for tests): their count and their total bytes.
  std::pair<size_t, size_t> free_ranges() {
    std::lock_guard<std::mutex> lock(mu_);
    size_t total = 0;
    for (const auto& r : by_off_) total += r.second;
    return {by_off_.size(), total};
  }
  size_t capacity() const { return cap_; }

 private:
  void add_free(size_t off, size_t size) {
    by_off_[off] = size;
    by_size_.insert({size, off});
  }
  size_t cap_ = 0, in_use_ = 0, peak_ = 0;
  std::map<size_t, size_t> by_off_;              // free ranges: offset -> size
  std::set<std::pair<size_t, size_t>> by_size_;  // the same ranges as (size, offset)
  std::unordered_map<size_t, size_t> used_;      // allocated ranges: offset -> size
  std::mutex mu_;
};

}  // namespace lrt
