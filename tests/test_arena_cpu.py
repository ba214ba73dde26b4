"""The serving arena's range allocator (librett_amd/csrc/arena.h), built on
its own with the host compiler and driven through random allocations and
releases from several threads: ranges never overlap or leave the arena, the
bytes in use are the sum of the live ranges, best fit picks the smallest
free range that holds a request, free ranges always merge, and once every
range is returned the whole capacity is one free range again."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include "arena.h"
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

using lrt::RangeAllocator;

static int fail(const char* what) { std::printf("FAIL %s\n", what); return 1; }

int main() {
  // Best fit, by hand: free ranges of 3, 1 and 2 granules (in that order of
  // offset) take a 1-granule request in the 1-granule range and a
  // 2-granule one in the 2-granule range.
  {
    const size_t G = RangeAllocator::kGranule;
    RangeAllocator a;
    a.reset(9 * G);
    size_t r[6];
    const size_t sizes[6] = {3, 1, 1, 1, 2, 1};
    for (int i = 0; i < 6; i++) r[i] = a.alloc(sizes[i] * G);  // 0,3,4,5,6,8 (granules)
    if (a.alloc(1) != RangeAllocator::kNone) return fail("full arena allocates");
    a.release(r[0]);  // free: [0,3)
    a.release(r[2]);  // free: [4,5)
    a.release(r[4]);  // free: [6,8)
    if (a.alloc(G) != 4 * G) return fail("best fit, one granule");
    if (a.alloc(2 * G) != 6 * G) return fail("best fit, two granules");
    if (a.alloc(3 * G) != 0) return fail("best fit, three granules");
    if (a.alloc(1) != RangeAllocator::kNone) return fail("full again");
    if (a.release(12345)) return fail("release of an unknown offset");
  }
  // Random allocations and releases from four threads on one allocator.
  for (int trial = 0; trial < 20; trial++) {
    std::mt19937_64 g0(trial);
    const size_t cap = (size_t)(g0() % 64 + 1) << 20;
    RangeAllocator a;
    a.reset(cap);
    std::vector<std::thread> ts;
    std::vector<int> bad(4, 0);
    for (int k = 0; k < 4; k++)
      ts.emplace_back([&, k] {
        std::mt19937_64 g(1000 * trial + k);
        std::vector<std::pair<size_t, size_t>> live;
        for (int step = 0; step < 4000; step++) {
          if (live.empty() || g() % 3) {
            const size_t n = (size_t)(g() % (cap / 16)) + 1;
            const size_t off = a.alloc(n);
            if (off == RangeAllocator::kNone) continue;
            const size_t len = std::max(RangeAllocator::kGranule, (n + 255) / 256 * 256);
            if (off % RangeAllocator::kGranule || off + len > a.capacity()) bad[k]++;
            live.push_back({off, len});
          } else {
            const size_t i = g() % live.size();
            if (!a.release(live[i].first)) bad[k]++;
            live.erase(live.begin() + (long)i);
          }
        }
        for (auto& r : live)
          if (!a.release(r.first)) bad[k]++;
      });
    for (auto& t : ts) t.join();
    for (int k = 0; k < 4; k++)
      if (bad[k]) return fail("bounds or release");
    if (a.in_use() != 0) return fail("bytes in use after every release");
    const auto fr = a.free_ranges();
    if (fr.first != 1 || fr.second != a.capacity()) return fail("free ranges not merged");
  }
  // One thread, checking overlap and the byte count after every step.
  {
    std::mt19937_64 g(7);
    RangeAllocator a;
    a.reset(16 << 20);
    std::vector<std::pair<size_t, size_t>> live;
    for (int step = 0; step < 20000; step++) {
      if (live.empty() || g() % 2) {
        const size_t n = (size_t)(g() % (1 << 20)) + 1;
        const size_t off = a.alloc(n);
        if (off != RangeAllocator::kNone) live.push_back({off, std::max<size_t>(256, (n + 255) / 256 * 256)});
      } else {
        const size_t i = g() % live.size();
        a.release(live[i].first);
        live.erase(live.begin() + (long)i);
      }
      std::vector<std::pair<size_t, size_t>> v = live;
      std::sort(v.begin(), v.end());
      size_t sum = 0;
      for (size_t i = 0; i < v.size(); i++) {
        sum += v[i].second;
        if (i && v[i - 1].first + v[i - 1].second > v[i].first) return fail("overlap");
      }
      if (sum != a.in_use()) return fail("bytes in use");
      const auto fr = a.free_ranges();
      if (fr.second + sum != a.capacity()) return fail("free + used != capacity");
    }
  }
  std::printf("ok\n");
  return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_range_allocator(tmp_path):
    src = tmp_path / "arena_test.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "arena_test"
    inc = os.path.join(ROOT, "librett_amd", "csrc")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", inc, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
