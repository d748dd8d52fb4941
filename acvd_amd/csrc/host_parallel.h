// Host loops split across the cores (mesh preparation outside the sweep).
#pragma once

#include <algorithm>
#include <cstdint>
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

}  // namespace acvd
