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

#include "host_parallel.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

namespace acvd {

namespace {

double face_area(const double* xyz, const int32_t* t) {
  const double* p1 = xyz + 3 * int64_t(t[0]);
  const double* p2 = xyz + 3 * int64_t(t[1]);
  const double* p3 = xyz + 3 * int64_t(t[2]);
  auto d2 = [](const double* a, const double* b) {
    return (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
  };
  const double a = d2(p1, p2), b = d2(p2, p3), c = d2(p3, p1);
  return 0.25 * std::sqrt(std::fabs(4.0 * a * c - (a - b + c) * (a - b + c)));
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
  const bool qem = metric == 1;
  if (indicator && (!qem || gradation > 0))
    parallel_for(nv, [&](int64_t v0, int64_t v1) {
      for (int64_t v = v0; v < v1; ++v) w[v] *= std::pow(indicator[v], gradation);
    });
  {
    double avg = 0.0;
    for (double x : w) avg += x;
    avg /= double(nv);
    const double ratio = qem ? 10000.0 : 100000.0;
    const double mn = avg / ratio, mx = avg * ratio;
    for (double& x : w) {
      if (x > mx) x = mx;
      if (x < mn) x = mn;
    }
  }
  // region growing
  const int64_t n = nv;
  for (int64_t i = 0; i < n; ++i) labels[i] = K;
  const int32_t offset = int32_t(n_fixed);
  for (int32_t k = 0; k < offset; ++k) labels[fixed[k]] = k;
  std::vector<int32_t> items(n);
  for (int64_t i = 0; i < n; ++i) items[i] = int32_t(i);
  std::mt19937 rng;
  rng.seed(0);
  for (int64_t i = n - 1; i > 0; --i) std::swap(items[i], items[rng() % uint64_t(n)]);
  double sw = 0.0;
  for (int64_t i = 0; i < n; ++i) sw += w[i];
  const double target = sw / double(K);
  int64_t remaining_items = n;
  int32_t remaining_regions = K - offset;
  int64_t first = 0;
  // FIFO of one region as a flat array
  std::vector<int32_t> queue;
  queue.reserve(1024);
  while (remaining_items > 0 && remaining_regions > 0) {
    int32_t item = -1;
    while (first < n) {
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
      for (int64_t r = rptr[it]; r < rptr[it + 1]; ++r) queue.push_back(ring[size_t(r)]);
      if (sw > target) break;
    }
  }
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

}  // namespace acvd
