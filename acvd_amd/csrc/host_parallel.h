// Host loops split across the cores (mesh preparation outside the sweep).
#pragma once

#include <sys/mman.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <new>
#include <thread>
#include <vector>

namespace acvd {

// f(begin, end) over [0, n) split across the host's cores (the per-vertex
// stages are independent; every result is the one a serial loop gives).
template <class F>
inline void parallel_for(int64_t n, F&& f) {
  const unsigned hw = std::thread::hardware_concurrency();
  const int64_t T = std::min<int64_t>(16, std::max<int64_t>(1, hw));
  if (n < (int64_t(1) << 16) || T == 1) {
    f(int64_t(0), n);
    return;
  }
  const int64_t chunk = (n + T - 1) / T;
  std::vector<std::thread> th;
  for (int64_t t = 0; t < T; ++t) {
    const int64_t a = t * chunk, b = std::min(n, a + chunk);
    if (a < b) th.emplace_back([&f, a, b] { f(a, b); });
  }
  for (auto& x : th) x.join();
}

// Allocator for the large arrays the serial region growing walks in random
// order (hundreds of MB): blocks of 2 MiB and more are mapped with
// transparent huge pages requested before the first touch, so that the walk
// does not miss in the TLB on every region.  Smaller blocks come from new.
template <class T>
struct HugeAlloc {
  using value_type = T;
  HugeAlloc() = default;
  template <class U>
  HugeAlloc(const HugeAlloc<U>&) {}
  static constexpr size_t kHuge = size_t(2) << 20;
  T* allocate(size_t n) {
    const size_t bytes = n * sizeof(T);
    if (bytes < kHuge) return static_cast<T*>(::operator new(bytes));
    const size_t len = (bytes + kHuge - 1) / kHuge * kHuge;
    void* p = mmap(nullptr, len, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (p == MAP_FAILED) throw std::bad_alloc();
    (void)madvise(p, len, MADV_HUGEPAGE);
    return static_cast<T*>(p);
  }
  void deallocate(T* p, size_t n) {
    const size_t bytes = n * sizeof(T);
    if (bytes < kHuge) {
      ::operator delete(p);
      return;
    }
    (void)munmap(p, (bytes + kHuge - 1) / kHuge * kHuge);
  }
  template <class U>
  bool operator==(const HugeAlloc<U>&) const { return true; }
  template <class U>
  bool operator!=(const HugeAlloc<U>&) const { return false; }
};

template <class T>
using huge_vector = std::vector<T, HugeAlloc<T>>;

}  // namespace acvd
