// Thread-safe least-recently-used map. The reference keeps one for kernel
// occupancy (LRUCache, src/LRUCache.h:36-104, used at src/kernel.cpp:691);
// here it caches whole plan choices so that re-planning a seen problem skips
// candidate generation, modelling and measurement.
#pragma once

#include <list>
#include <mutex>
#include <unordered_map>
#include <utility>

namespace lrt {

template <typename K, typename V, typename Hash = std::hash<K>>
class LruCache {
 public:
  explicit LruCache(size_t capacity) : cap_(capacity) {}

  bool get(const K& key, V& value) {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = map_.find(key);
    if (it == map_.end()) {
      misses_++;
      return false;
    }
    order_.splice(order_.begin(), order_, it->second);
    value = it->second->second;
    hits_++;
    return true;
  }

  void put(const K& key, const V& value) {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = map_.find(key);
    if (it != map_.end()) {
      it->second->second = value;
      order_.splice(order_.begin(), order_, it->second);
      return;
    }
    order_.emplace_front(key, value);
    map_[key] = order_.begin();
    while (map_.size() > cap_) {
      map_.erase(order_.back().first);
      order_.pop_back();
    }
  }

  void clear() {
    std::lock_guard<std::mutex> lock(mu_);
    map_.clear();
    order_.clear();
  }

  size_t size() {
    std::lock_guard<std::mutex> lock(mu_);
    return map_.size();
  }

  unsigned long long hits() const { return hits_; }
  unsigned long long misses() const { return misses_; }

 private:
  size_t cap_;
  std::mutex mu_;
  std::list<std::pair<K, V>> order_;
  std::unordered_map<K, typename std::list<std::pair<K, V>>::iterator, Hash> map_;
  unsigned long long hits_ = 0, misses_ = 0;
};

}  // namespace lrt
