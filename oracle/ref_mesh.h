// TEST INFRASTRUCTURE — CPU oracle, not part of the product.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load or run anything under oracle/, and only as the checker.
//
// VTK-free restatement of the mesh topology the reference's clustering reads:
//   - the edge table built face by face with edges (v_j, v_{j+1 mod 3}), the
//     first-seen direction kept, Poly1/Poly2 = first/second adjacent face and
//     any further faces recorded as non-manifold
//     (Common/vtkSurfaceBase.cxx:1166-1221 AddEdge, :1407-1468
//     CreateFromPolyData);
//   - vertex rings = incident edges in insertion order
//     (vtkSurfaceBase.cxx:1057-1068, vtkSurfaceBase.h:471-477).  Edges are
//     only ever appended while building, so a ring is its incident edge ids in
//     increasing order;
//   - GetVertexNeighbourFaces (vtkSurfaceBase.cxx:964-984), GetVertexNeighbours
//     (:988-1000);
//   - face area by the squared-edge-length formula of Common/vtkSurface.cxx:1280-1283
//     and vertex area = sum of adjacent face areas / 3 (vtkSurface.cxx:1342-1359).
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace oracle {

struct RefMesh {
  int64_t nv = 0, nf = 0;
  std::vector<double> P;         // 3*nv
  std::vector<int32_t> T;        // 3*nf
  // edge table
  std::vector<int32_t> ev1, ev2;      // Vertex1, Vertex2
  std::vector<int32_t> poly1, poly2;  // adjacent faces (-1 = none)
  std::vector<int64_t> nm_ptr;        // CSR of extra (non-manifold) faces per edge
  std::vector<int32_t> nm_faces;
  bool has_nonmanifold = false;
  // rings
  std::vector<int64_t> ring_ptr;  // nv+1
  std::vector<int32_t> ring;      // incident edge ids, increasing

  int64_t num_edges() const { return int64_t(ev1.size()); }

  void point(int64_t v, double* x) const {
    x[0] = P[3 * v];
    x[1] = P[3 * v + 1];
    x[2] = P[3 * v + 2];
  }

  void build(int64_t nverts, const double* xyz, int64_t nfaces, const int32_t* tris) {
    nv = nverts;
    nf = nfaces;
    P.assign(xyz, xyz + 3 * nv);
    T.assign(tris, tris + 3 * nf);
    for (int64_t i = 0; i < 3 * nf; ++i)
      if (T[i] < 0 || T[i] >= nv) throw std::runtime_error("face references a vertex that does not exist");
    // open-addressing table (min,max) -> edge id
    uint64_t cap = 64;
    while (cap < uint64_t(3 * nf) * 2) cap <<= 1;
    std::vector<uint64_t> keys(cap, ~0ull);
    std::vector<int32_t> vals(cap, -1);
    const uint64_t mask = cap - 1;
    std::vector<std::pair<int32_t, int32_t>> extra;  // (edge, face) non-manifold
    ev1.clear(); ev2.clear(); poly1.clear(); poly2.clear();
    ev1.reserve(size_t(nf) * 3 / 2 + 16);
    for (int64_t f = 0; f < nf; ++f) {
      const int32_t* v = &T[3 * f];
      if (v[0] == v[1]) continue;  // inactive face (vtkSurfaceBase.cxx:1443)
      for (int j = 0; j < 3; ++j) {
        int32_t a = v[j], b = v[(j + 1) % 3];
        if (a == b) continue;  // self-loop refused (vtkSurfaceBase.cxx:1168-1172)
        uint64_t lo = uint64_t(a < b ? a : b), hi = uint64_t(a < b ? b : a);
        uint64_t key = (lo << 32) | hi;
        uint64_t h = (key * 0x9E3779B97F4A7C15ull) >> 17;
        for (;;) {
          uint64_t s = h & mask;
          if (keys[s] == ~0ull) {
            int32_t e = int32_t(ev1.size());
            keys[s] = key;
            vals[s] = e;
            ev1.push_back(a);
            ev2.push_back(b);
            poly1.push_back(int32_t(f));
            poly2.push_back(-1);
            break;
          }
          if (keys[s] == key) {
            int32_t e = vals[s];
            if (poly1[e] < 0) poly1[e] = int32_t(f);
            else if (poly2[e] >= 0) extra.emplace_back(e, int32_t(f));
            else poly2[e] = int32_t(f);
            break;
          }
          ++h;
        }
      }
    }
    const int64_t ne = num_edges();
    nm_ptr.assign(ne + 1, 0);
    has_nonmanifold = !extra.empty();
    for (auto& p : extra) nm_ptr[p.first + 1]++;
    for (int64_t e = 0; e < ne; ++e) nm_ptr[e + 1] += nm_ptr[e];
    nm_faces.assign(extra.size(), -1);
    {
      std::vector<int64_t> fill(nm_ptr.begin(), nm_ptr.end() - 1);
      for (auto& p : extra) nm_faces[fill[p.first]++] = p.second;
    }
    ring_ptr.assign(nv + 1, 0);
    for (int64_t e = 0; e < ne; ++e) {
      ring_ptr[ev1[e] + 1]++;
      ring_ptr[ev2[e] + 1]++;
    }
    for (int64_t v = 0; v < nv; ++v) ring_ptr[v + 1] += ring_ptr[v];
    ring.assign(ring_ptr[nv], -1);
    std::vector<int64_t> fill(ring_ptr.begin(), ring_ptr.end() - 1);
    for (int64_t e = 0; e < ne; ++e) {
      ring[fill[ev1[e]]++] = int32_t(e);
      ring[fill[ev2[e]]++] = int32_t(e);
    }
  }

  // vtkSurfaceBase::GetVertexNeighbourFaces: per ring edge, Poly1, Poly2 and
  // the non-manifold faces, each inserted once (InsertUniqueId keeps the
  // first-insertion order).
  void neighbour_faces(int64_t v, std::vector<int32_t>& out) const {
    out.clear();
    auto insert_unique = [&](int32_t f) {
      for (int32_t g : out)
        if (g == f) return;
      out.push_back(f);
    };
    for (int64_t k = ring_ptr[v]; k < ring_ptr[v + 1]; ++k) {
      int32_t e = ring[k];
      if (poly1[e] >= 0) insert_unique(poly1[e]);
      if (poly2[e] >= 0) insert_unique(poly2[e]);
      for (int64_t t = nm_ptr[e]; t < nm_ptr[e + 1]; ++t) insert_unique(nm_faces[t]);
    }
  }

  // vtkSurfaceBase::GetVertexNeighbours: other endpoint of each ring edge.
  void neighbours(int64_t v, std::vector<int32_t>& out) const {
    out.clear();
    for (int64_t k = ring_ptr[v]; k < ring_ptr[v + 1]; ++k) {
      int32_t e = ring[k];
      out.push_back(ev1[e] == v ? ev2[e] : ev1[e]);
    }
  }

  static double dist2(const double* a, const double* b) {
    return (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
  }

  // Common/vtkSurface.cxx:1280-1283
  double face_area(int64_t f) const {
    double p1[3], p2[3], p3[3];
    point(T[3 * f], p1);
    point(T[3 * f + 1], p2);
    point(T[3 * f + 2], p3);
    double a = dist2(p1, p2), b = dist2(p2, p3), c = dist2(p3, p1);
    return 0.25 * std::sqrt(std::fabs(4.0 * a * c - (a - b + c) * (a - b + c)));
  }

  // Common/vtkSurface.cxx:1342-1359, :1361-1378
  void vertex_areas(std::vector<double>& out) const {
    out.assign(nv, 0.0);
    std::vector<int32_t> fl;
    for (int64_t v = 0; v < nv; ++v) {
      neighbour_faces(v, fl);
      double a = 0.0;
      for (int32_t f : fl) a += face_area(f) / 3.0;
      out[v] = a;
    }
  }
};

}  // namespace oracle
