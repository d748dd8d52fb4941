// Reference initial sampling on the host.
//
// Region growing from shuffled start items until each region holds the mean
// weight (vtkUniformClustering.h:1177-1316).  It depends on the reference's
// topology order: the neighbours of a vertex are visited in the order its
// edges were created when the mesh was built face by face (edges
// (v0,v1), (v1,v2), (v2,v0) of each face, Common/vtkSurfaceBase.cxx:1166-1221,
// 1407-1468; rings in insertion order, :1057-1068), and the item weights are
// the metric's: vertex area = sum over the neighbour faces, in the order
// GetVertexNeighbourFaces lists them (vtkSurfaceBase.cxx:964-984), of the
// face area (Common/vtkSurface.cxx:1280-1283) / 3; times indicator^gradation
// (isotropic: whenever an indicator is set, vtkIsotropicMetricForClustering.h:
// 249-270; QEM: when gradation > 0, vtkQEMetricForClustering.h:324-342);
// clamped to [avg/R, avg*R] with R = 1e5 (isotropic) or 1e4 (QEM).
// The shuffle keeps the reference's std::mt19937 seeded with 0 and its
// `rng() % n` index (vtkUniformClustering.h:1202-1207).
#include "host_sampling.h"

#include "acvd_math.h"
#include "host_parallel.h"

#include <algorithm>
#include <cmath>
#include <atomic>
#include <climits>
#include <random>
#include <thread>
#include <vector>

namespace acvd {

namespace {

double face_area(const double* xyz, const int32_t* t) {
  return tri_area_f64(xyz + 3 * int64_t(t[0]), xyz + 3 * int64_t(t[1]), xyz + 3 * int64_t(t[2]));
}

// vtkMath::Random as VTK computes it: vtkMinimalStandardRandomSequence, the
// Park-Miller "minimal standard" generator (multiplier 16807, modulus
// 2^31 - 1, Schrage's decomposition q = 127773, r = 2836).  SetSeed maps the
// seed into [1, 2^31 - 2] and advances three times; Random() advances, then
// returns seed / (2^31 - 1); Random(a, b) = a + Random() * (b - a).
struct MinimalStandard {
  int32_t seed = 1;
  void set_seed(int32_t v) {
    seed = v;
    if (seed < 1) seed += 2147483646;
    else if (seed == 2147483647) seed = 1;
    next();
    next();
    next();
  }
  void next() {
    const int32_t hi = seed / 127773, lo = seed % 127773;
    seed = 16807 * lo - 2836 * hi;
    if (seed <= 0) seed += 2147483647;
  }
  double random() {
    next();
    return double(seed) / 2147483647.0;
  }
  double random(double a, double b) {
    next();
    return a == b ? a : a + (double(seed) / 2147483647.0) * (b - a);
  }
};

}  // namespace

bool reference_seed_sampling(int64_t nv, int32_t K, int32_t* labels) {
  for (int64_t i = 0; i < nv; ++i) labels[i] = K;
  if (K < 0 || int64_t(K) > nv) return false;
  MinimalStandard rng;
  rng.set_seed(5000);
  (void)rng.random();
  const float size = float(nv);
  for (int32_t c = 0; c < K; ++c)
    for (;;) {
      const float test = float(rng.random(0.0, double(size)));
      const int64_t number = int64_t(std::floor(double(test) + 0.5));
      if (number >= nv) continue;  // the reference reads one past its labels here
      if (labels[number] == K) {
        labels[number] = c;
        break;
      }
    }
  return true;
}

void vertex_topology(int64_t nv, int64_t nf, const int32_t* tri, VertexTopology* out) {
  VertexTopology& T = *out;
  // vertex -> faces, increasing face id (faces with a repeated first edge
  // vertex are inactive in the reference and skipped; a face listing a
  // vertex twice is stored once for it)
  auto slot_of = [](const int32_t* t, int j) { return !(j > 0 && t[j] == t[0]) && !(j > 1 && t[j] == t[1]); };
  T.vfptr.assign(size_t(nv + 1), 0);
  for (int64_t f = 0; f < nf; ++f) {
    const int32_t* t = tri + 3 * f;
    if (t[0] == t[1]) continue;
    for (int j = 0; j < 3; ++j)
      if (slot_of(t, j)) T.vfptr[size_t(t[j] + 1)]++;
  }
  for (int64_t v = 0; v < nv; ++v) T.vfptr[size_t(v + 1)] += T.vfptr[size_t(v)];
  T.vf.assign(size_t(T.vfptr[size_t(nv)]), 0);
  {
    std::vector<int64_t> fill(T.vfptr.begin(), T.vfptr.end() - 1);
    for (int64_t f = 0; f < nf; ++f) {
      const int32_t* t = tri + 3 * f;
      if (t[0] == t[1]) continue;
      for (int j = 0; j < 3; ++j)
        if (slot_of(t, j)) T.vf[size_t(fill[size_t(t[j])]++)] = int32_t(f);
    }
  }
  const std::vector<int64_t>& vfptr = T.vfptr;
  const std::vector<int32_t>& vf = T.vf;
  // rings in edge-creation order: scanning the faces of v in increasing id and
  // each face's edges in order, the first sighting of neighbour u is the
  // creation of edge (v,u), which is when it joins both rings.  Two passes:
  // sizes, then the rings in place.
  auto ring_of = [&](int64_t v, int32_t* dst) {
    int64_t cnt = 0;
    for (int64_t k = vfptr[size_t(v)]; k < vfptr[size_t(v + 1)]; ++k) {
      const int32_t* t = tri + 3 * int64_t(vf[size_t(k)]);
      for (int j = 0; j < 3; ++j) {
        const int32_t a = t[j], b = t[(j + 1) % 3];
        if (a == b) continue;
        int32_t u = -1;
        if (a == v) u = b;
        else if (b == v) u = a;
        if (u < 0) continue;
        bool seen = false;
        for (int64_t q = 0; q < cnt; ++q)
          if (dst[q] == u) {
            seen = true;
            break;
          }
        if (!seen) dst[cnt++] = u;
      }
    }
    return cnt;
  };
  T.rptr.assign(size_t(nv + 1), 0);
  parallel_for(nv, [&](int64_t v0, int64_t v1) {
    std::vector<int32_t> tmp;
    for (int64_t v = v0; v < v1; ++v) {
      tmp.resize(size_t(2 * (vfptr[size_t(v + 1)] - vfptr[size_t(v)])));
      T.rptr[size_t(v + 1)] = ring_of(v, tmp.data());
    }
  });
  for (int64_t v = 0; v < nv; ++v) T.rptr[size_t(v + 1)] += T.rptr[size_t(v)];
  T.ring.assign(size_t(T.rptr[size_t(nv)]), 0);
  parallel_for(nv, [&](int64_t v0, int64_t v1) {
    for (int64_t v = v0; v < v1; ++v) ring_of(v, T.ring.data() + T.rptr[size_t(v)]);
  });
}

void sampling_weights(int64_t nv, int metric, const double* indicator, double gradation, double* w) {
  const bool qem = metric == 1;
  if (indicator && (!qem || gradation > 0))
    parallel_for(nv, [&](int64_t v0, int64_t v1) {
      for (int64_t v = v0; v < v1; ++v) w[v] *= std::pow(indicator[v], gradation);
    });
  double avg = 0.0;
  for (int64_t v = 0; v < nv; ++v) avg += w[v];
  avg /= double(nv);
  const double ratio = qem ? 10000.0 : 100000.0;
  const double mn = avg / ratio, mx = avg * ratio;
  parallel_for(nv, [&](int64_t v0, int64_t v1) {
    for (int64_t v = v0; v < v1; ++v) {
      if (w[v] > mx) w[v] = mx;
      if (w[v] < mn) w[v] = mn;
    }
  });
}

namespace {

constexpr int64_t kParallelGrowthItems = int64_t(1) << 18;

int growth_threads() {
  const unsigned hw = std::thread::hardware_concurrency();
  return int(std::min<unsigned>(16, std::max<unsigned>(1, hw)));
}

// The growth loop of ComputeInitialRandomSampling (vtkUniformClustering.h:
// 1214-1257): regions in turn, each seeded at the first null item of the
// shuffled order and grown breadth first until its weight exceeds the mean.
void grow_regions_serial(int64_t n, const int32_t* items, const int64_t* rbeg, const int32_t* rlen,
                         const int32_t* ring, const double* w, double target, int32_t K, int32_t offset,
                         int32_t* labels, int64_t* rem_items, int32_t* rem_regions) {
  // FIFO of one region as a flat array.  A neighbour already labelled is not
  // queued: it would be skipped when popped (labels never return to null
  // during the growth), so the regions are the reference's.
  std::vector<int32_t> queue;
  queue.reserve(1024);
  int64_t& remaining_items = *rem_items;
  int32_t& remaining_regions = *rem_regions;
  int64_t first = 0;
  double sw = 0.0;
  while (remaining_items > 0 && remaining_regions > 0) {
    int32_t item = -1;
    while (first < n) {
      if (first + 16 < n) __builtin_prefetch(&labels[items[first + 16]]);
      if (labels[items[first]] == K) {
        item = items[first];
        break;
      }
      ++first;
    }
    if (item < 0) break;  // every item labelled
    queue.clear();
    queue.push_back(item);
    size_t head = 0;
    sw = 0.0;
    --remaining_regions;
    while (head < queue.size()) {
      const int32_t it = queue[head++];
      if (labels[it] != K) continue;
      labels[it] = remaining_regions + offset;
      sw += w[it];
      --remaining_items;
      const int32_t* r = ring + rbeg[it];
      for (int32_t k = 0; k < rlen[it]; ++k)
        if (labels[r[k]] == K) queue.push_back(r[k]);
      if (sw > target) break;
    }
  }
}

// Set of item ids as a bitmap over all items (one per thread, cleared
// through the ids it holds).
struct ItemBits {
  std::vector<uint64_t> word;
  explicit ItemBits(int64_t n) : word(static_cast<size_t>(n / 64 + 1), 0) {}
  bool contains(int32_t x) const { return (word[size_t(x) >> 6] >> (x & 63)) & 1u; }
  void insert(int32_t x) { word[size_t(x) >> 6] |= uint64_t(1) << (x & 63); }
  void clear(const std::vector<int32_t>& ids) {
    for (int32_t x : ids) word[size_t(x) >> 6] = 0;
  }
};

// One region grown on the labels as they stand (read only), as the serial
// loop would grow it if no region before it in the batch changed an item it
// read.  The queue holds each item once (a second copy would be skipped when
// popped: the first copy labels it), so no queued item is labelled before it
// is popped.  lab: the items it labels, in order; queue: every item it read
// as null (the seed and every queued item), the only reads an earlier region
// can change.
struct Speculation {
  std::vector<int32_t> lab, queue;
  void grow(int32_t seed, const int64_t* rbeg, const int32_t* rlen, const int32_t* ring, const double* w,
            double target, int32_t K, const int32_t* labels, ItemBits& queued) {
    lab.clear();
    queue.clear();
    queue.push_back(seed);
    queued.insert(seed);
    size_t head = 0;
    double sw = 0.0;
    while (head < queue.size()) {
      const int32_t it = queue[head++];
      lab.push_back(it);
      sw += w[it];
      const int32_t* r = ring + rbeg[it];
      for (int32_t k = 0; k < rlen[it]; ++k) {
        const int32_t u = r[k];
        if (labels[u] == K && !queued.contains(u)) {
          queue.push_back(u);
          queued.insert(u);
        }
      }
      if (sw > target) break;
    }
    queued.clear(queue);
  }
};

// Sense-reversing spin barrier for the growth's fixed team of threads.
struct SpinBarrier {
  explicit SpinBarrier(int n) : n_(n) {}
  void wait() {
    const int gen = gen_.load(std::memory_order_acquire);
    if (count_.fetch_add(1, std::memory_order_acq_rel) + 1 == n_) {
      count_.store(0, std::memory_order_relaxed);
      gen_.store(gen + 1, std::memory_order_release);
      return;
    }
    for (int spins = 0; gen_.load(std::memory_order_acquire) == gen; ++spins)
      if (spins > 1024) std::this_thread::yield();
  }
  const int n_;
  std::atomic<int> count_{0}, gen_{0};
};

// The same growth loop, label for label, with the regions grown
// speculatively in batches.  The next W seeds (the first W null items of the
// shuffled order) are grown in parallel on the labels as they stand.  Region
// i of the batch is the serial loop's region iff its seed and every item it
// read as null are still null once regions 0..i-1 are committed (labels only
// leave null, and only the batch's own regions change them).  In parallel,
// each region marks the items it labels with its batch index (the smallest
// wins) and tests its reads against the marks of the regions before it: the
// regions before the first that meets an earlier one are committed in
// parallel as grown.  From that one on, in order: a region whose reads are
// all still null is committed as grown, any other is grown again on the
// spot; a seed taken by an earlier region ends the batch there (the serial
// loop would seed the next region elsewhere).  The seeds are random, so most
// regions never meet; W follows the length of the conflict-free prefixes.
void grow_regions_parallel(int64_t n, const int32_t* items, const int64_t* rbeg, const int32_t* rlen,
                           const int32_t* ring, const double* w, double target, int32_t K, int32_t offset,
                           int32_t* labels, int64_t* rem_items, int32_t* rem_regions) {
  const int T = growth_threads();
  const int32_t max_batch = 64 * T;
  const size_t nw = static_cast<size_t>(max_batch);
  std::vector<Speculation> spec(nw);
  std::vector<int32_t> cand(nw), queue;
  std::vector<int64_t> cand_pos(nw);
  std::vector<ItemBits> queued(static_cast<size_t>(T), ItemBits(n));
  huge_vector<int32_t> owner(static_cast<size_t>(n), INT32_MAX);
  std::atomic<int32_t> first_conflict{INT32_MAX};
  int64_t first = 0;
  int32_t batch = 8 * T, nc = 0, top_id = 0;
  bool done = false;
  SpinBarrier bar(T);
  // phases run by every thread t on its share of the batch
  auto grow_share = [&](int t) {
    for (int32_t i = t; i < nc; i += T)
      spec[size_t(i)].grow(cand[size_t(i)], rbeg, rlen, ring, w, target, K, labels, queued[size_t(t)]);
  };
  auto mark_share = [&](int t) {
    for (int32_t i = t; i < nc; i += T)
      for (int32_t x : spec[size_t(i)].lab) {
        int32_t o = __atomic_load_n(&owner[size_t(x)], __ATOMIC_RELAXED);
        while (i < o &&
               !__atomic_compare_exchange_n(&owner[size_t(x)], &o, i, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
      }
  };
  auto test_share = [&](int t) {
    for (int32_t i = t; i < nc; i += T)
      for (int32_t x : spec[size_t(i)].queue)
        if (__atomic_load_n(&owner[size_t(x)], __ATOMIC_RELAXED) < i) {
          int32_t c = first_conflict.load(std::memory_order_relaxed);
          while (i < c && !first_conflict.compare_exchange_weak(c, i, std::memory_order_relaxed)) {
          }
          break;
        }
  };
  // commit the conflict-free prefix (region i labels id top_id - i) and
  // clear every mark
  auto commit_share = [&](int t) {
    const int32_t c = std::min(first_conflict.load(std::memory_order_relaxed), nc);
    for (int32_t i = t; i < nc; i += T)
      for (int32_t x : spec[size_t(i)].lab) {
        if (i < c) labels[x] = top_id - i;  // the prefix's regions are disjoint
        __atomic_store_n(&owner[size_t(x)], INT32_MAX, __ATOMIC_RELAXED);
      }
  };
  auto team = [&](int t) {
    for (;;) {
      bar.wait();  // seeds collected (or done)
      if (done) return;
      grow_share(t);
      bar.wait();
      mark_share(t);
      bar.wait();
      test_share(t);
      bar.wait();
      commit_share(t);
      bar.wait();  // prefix committed
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(team, t);
  for (;;) {
    nc = 0;
    int64_t scan = first;
    if (*rem_items > 0 && *rem_regions > 0) {
      const int32_t want = std::min(batch, *rem_regions);
      while (nc < want && scan < n) {
        if (scan + 16 < n) __builtin_prefetch(&labels[items[scan + 16]]);
        if (labels[items[scan]] == K) {
          cand[size_t(nc)] = items[scan];
          cand_pos[size_t(nc)] = scan;
          ++nc;
        }
        ++scan;
      }
    }
    done = nc == 0;
    top_id = *rem_regions - 1 + offset;
    first_conflict.store(INT32_MAX, std::memory_order_relaxed);
    bar.wait();
    if (done) break;
    grow_share(0);
    bar.wait();
    mark_share(0);
    bar.wait();
    test_share(0);
    bar.wait();
    commit_share(0);
    bar.wait();
    const int32_t c = std::min(first_conflict.load(std::memory_order_relaxed), nc);
    for (int32_t i = 0; i < c; ++i) *rem_items -= int64_t(spec[size_t(i)].lab.size());
    *rem_regions -= c;
    // from the first conflict on, in order
    int32_t i = c;
    for (; i < nc; ++i) {
      const Speculation& sp = spec[size_t(i)];
      const int32_t seed = cand[size_t(i)];
      if (labels[seed] != K) break;
      bool valid = true;
      for (int32_t x : sp.queue)
        if (labels[x] != K) {
          valid = false;
          break;
        }
      --*rem_regions;
      const int32_t id = *rem_regions + offset;
      if (valid) {
        for (int32_t x : sp.lab) labels[x] = id;
        *rem_items -= int64_t(sp.lab.size());
        continue;
      }
      queue.clear();
      queue.push_back(seed);
      size_t head = 0;
      double sw = 0.0;
      while (head < queue.size()) {
        const int32_t it = queue[head++];
        if (labels[it] != K) continue;
        labels[it] = id;
        sw += w[it];
        --*rem_items;
        const int32_t* r = ring + rbeg[it];
        for (int32_t k = 0; k < rlen[it]; ++k)
          if (labels[r[k]] == K) queue.push_back(r[k]);
        if (sw > target) break;
      }
    }
    first = i < nc ? cand_pos[size_t(i)] : scan;
    batch = c == nc ? std::min(2 * batch, max_batch) : std::max(T, std::min(max_batch, 2 * c));
  }
  for (auto& th : pool) th.join();
}

}  // namespace

void sampling_order(int64_t n, SamplingOrder* out) {
  huge_vector<int32_t>& items = out->items;
  items.resize(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) items[size_t(i)] = int32_t(i);
  std::mt19937& rng = out->rng;
  rng.seed(0);
  if (n < 2) return;
  // The swap partners do not depend on the items: draw them first (rng() is
  // below 2^32, so rng() % n is the 32-bit remainder, taken by Lemire's
  // multiply-shift: floor(frac(x / n) * n) from a 64-bit reciprocal, exact
  // for 32-bit operands), then swap with the partner a few steps ahead
  // already on its way to the cache.
  const uint64_t d = uint64_t(n);
  const uint64_t M = UINT64_MAX / d + 1;
  huge_vector<uint32_t> partner(static_cast<size_t>(n - 1));
  for (int64_t s = 0; s < n - 1; ++s) {
    const uint64_t x = uint64_t(uint32_t(rng()));
    partner[size_t(s)] = uint32_t((static_cast<unsigned __int128>(M * x) * d) >> 64);
  }
  constexpr int64_t kAhead = 32;
  for (int64_t i = n - 1; i > 0; --i) {
    const int64_t s = n - 1 - i;
    if (s + kAhead < n - 1) __builtin_prefetch(&items[partner[size_t(s + kAhead)]], 1);
    std::swap(items[size_t(i)], items[partner[size_t(s)]]);
  }
}

bool region_growing(int64_t nv, const int64_t* rbeg, const int32_t* rlen, const int32_t* ring, const double* w,
                    int32_t K, int32_t* labels, int64_t n_fixed, const int32_t* fixed, SamplingOrder* order) {
  const int64_t n = nv;
  for (int64_t i = 0; i < n; ++i) labels[i] = K;
  const int32_t offset = int32_t(n_fixed);
  for (int32_t k = 0; k < offset; ++k) labels[fixed[k]] = k;
  SamplingOrder own;
  if (!order) {
    sampling_order(n, &own);
    order = &own;
  }
  huge_vector<int32_t>& items = order->items;
  std::mt19937& rng = order->rng;
  double sw = 0.0;
  for (int64_t i = 0; i < n; ++i) sw += w[i];
  const double target = sw / double(K);
  int64_t remaining_items = n;
  int32_t remaining_regions = K - offset;
  int64_t first = 0;
  if (n >= kParallelGrowthItems && growth_threads() > 1)
    grow_regions_parallel(n, items.data(), rbeg, rlen, ring, w, target, K, offset, labels, &remaining_items,
                          &remaining_regions);
  else
    grow_regions_serial(n, items.data(), rbeg, rlen, ring, w, target, K, offset, labels, &remaining_items,
                        &remaining_regions);
  if (remaining_regions == 0) return true;
  // more regions than the growth produced: single items are moved into the
  // remaining regions (vtkUniformClustering.h:1261-1311)
  std::vector<int32_t> size(size_t(K) + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    items[i] = int32_t(i);
    if (labels[i] != K) size[labels[i]]++;
  }
  for (int64_t i = n - 1; i > 0; --i) std::swap(items[i], items[rng() % uint64_t(n)]);
  first = 0;
  while (remaining_regions) {
    int32_t item = -1;
    while (first < n) {
      item = items[first++];
      const int32_t c = labels[item];
      if (c == K) break;
      if (size[c] == 1) continue;
      labels[item] = remaining_regions + offset;
      size[c]--;
      size[remaining_regions + offset]++;
      break;
    }
    if (item < 0) return false;
    --remaining_regions;
    labels[item] = remaining_regions + offset;
  }
  return true;
}

bool reference_initial_sampling(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, int metric,
                                const double* indicator, double gradation, int32_t K, int32_t* labels,
                                int64_t n_fixed, const int32_t* fixed) {
  VertexTopology topo;
  vertex_topology(nv, nf, tri, &topo);
  const std::vector<int64_t>& vfptr = topo.vfptr;
  const std::vector<int32_t>& vf = topo.vf;
  const std::vector<int64_t>& rptr = topo.rptr;
  const std::vector<int32_t>& ring = topo.ring;
  // metric weights: vertex area = sum, in GetVertexNeighbourFaces order (the
  // faces of each ring neighbour in turn), of the face areas / 3
  std::vector<double> farea(static_cast<size_t>(nf));
  parallel_for(nf, [&](int64_t f0, int64_t f1) {
    for (int64_t f = f0; f < f1; ++f) farea[size_t(f)] = face_area(xyz, tri + 3 * f) / 3.0;
  });
  std::vector<double> w(nv, 0.0);
  parallel_for(nv, [&](int64_t v0, int64_t v1) {
    std::vector<int32_t> nfaces;
    for (int64_t v = v0; v < v1; ++v) {
      nfaces.clear();
      for (int64_t r = rptr[v]; r < rptr[v + 1]; ++r) {
        const int32_t u = ring[size_t(r)];
        for (int64_t k = vfptr[v]; k < vfptr[v + 1]; ++k) {
          const int32_t f = vf[size_t(k)];
          const int32_t* t = tri + 3 * int64_t(f);
          if (t[0] != u && t[1] != u && t[2] != u) continue;
          if (std::find(nfaces.begin(), nfaces.end(), f) == nfaces.end()) nfaces.push_back(f);
        }
      }
      double a = 0.0;
      for (int32_t f : nfaces) a += farea[size_t(f)];
      w[v] = a;
    }
  });
  sampling_weights(nv, metric, indicator, gradation, w.data());
  std::vector<int32_t> rlen(static_cast<size_t>(nv));
  parallel_for(nv, [&](int64_t v0, int64_t v1) {
    for (int64_t v = v0; v < v1; ++v) rlen[size_t(v)] = int32_t(rptr[size_t(v + 1)] - rptr[size_t(v)]);
  });
  return region_growing(nv, rptr.data(), rlen.data(), ring.data(), w.data(), K, labels, n_fixed, fixed);
}

}  // namespace acvd
