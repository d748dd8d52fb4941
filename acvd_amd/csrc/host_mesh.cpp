// Host-side mesh preparation of Remesh() (vtkDiscreteRemeshing::
// CheckSubsamplingRatio, DiscreteRemeshing/vtkDiscreteRemeshing.h:840-875):
// one level of midpoint subdivision (vtkSurface::Subdivide,
// Common/vtkSurface.cxx:605-677).  Runs once per Remesh on meshes too coarse
// for the requested cluster count, before the device build.
#include "host_mesh.h"

#include <algorithm>
#include <vector>

namespace acvd {

int64_t subdivide_mesh(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, std::vector<double>* out_xyz,
                       std::vector<int32_t>* out_tri, std::vector<int32_t>* parents) {
  // edges in creation order: faces in order (inactive faces, whose first two
  // vertices coincide, create none), edges (v0,v1), (v1,v2), (v2,v0), a new
  // edge at its first sighting (Common/vtkSurfaceBase.cxx:1166-1221, 1443-1451).
  // Each edge is looked up in the list of its smaller end.
  std::vector<int64_t> cap(nv + 1, 0);
  for (int64_t f = 0; f < nf; ++f) {
    const int32_t* t = tri + 3 * f;
    if (t[0] == t[1]) continue;
    for (int j = 0; j < 3; ++j) {
      const int32_t a = t[j], b = t[(j + 1) % 3];
      if (a != b) cap[std::min(a, b) + 1]++;
    }
  }
  for (int64_t v = 0; v < nv; ++v) cap[v + 1] += cap[v];
  std::vector<int32_t> other(size_t(cap[nv])), eid(size_t(cap[nv]));
  std::vector<int64_t> used(nv, 0);
  std::vector<int32_t> face_edge(size_t(3 * nf), -1);
  std::vector<int32_t> ev1, ev2;
  for (int64_t f = 0; f < nf; ++f) {
    const int32_t* t = tri + 3 * f;
    if (t[0] == t[1]) continue;
    for (int j = 0; j < 3; ++j) {
      const int32_t a = t[j], b = t[(j + 1) % 3];
      if (a == b) continue;
      const int32_t lo = std::min(a, b), hi = std::max(a, b);
      const int64_t s0 = cap[lo];
      int32_t id = -1;
      for (int64_t k = s0; k < s0 + used[lo]; ++k)
        if (other[size_t(k)] == hi) {
          id = eid[size_t(k)];
          break;
        }
      if (id < 0) {
        id = int32_t(ev1.size());
        ev1.push_back(a);
        ev2.push_back(b);
        other[size_t(s0 + used[lo])] = hi;
        eid[size_t(s0 + used[lo])] = id;
        used[lo]++;
      }
      face_edge[size_t(3 * f + j)] = id;
    }
  }
  const int64_t ne = int64_t(ev1.size());
  // the input vertices, then one vertex per edge at its midpoint
  out_xyz->assign(xyz, xyz + 3 * nv);
  out_xyz->resize(size_t(3 * (nv + ne)));
  parents->resize(size_t(2 * ne));
  for (int64_t e = 0; e < ne; ++e) {
    const double* p1 = xyz + 3 * int64_t(ev1[size_t(e)]);
    const double* p2 = xyz + 3 * int64_t(ev2[size_t(e)]);
    for (int k = 0; k < 3; ++k) (*out_xyz)[size_t(3 * (nv + e) + k)] = 0.5 * (p1[k] + p2[k]);
    (*parents)[size_t(2 * e)] = ev1[size_t(e)];
    (*parents)[size_t(2 * e + 1)] = ev2[size_t(e)];
  }
  // four faces per active face: (V1,V4,V6), (V4,V2,V5), (V5,V3,V6), (V4,V5,V6)
  // with V4, V5, V6 the children of edges (V1,V2), (V2,V3), (V3,V1)
  out_tri->clear();
  out_tri->reserve(size_t(12 * nf));
  for (int64_t f = 0; f < nf; ++f) {
    const int32_t* t = tri + 3 * f;
    if (t[0] == t[1]) continue;
    const int32_t v1 = t[0], v2 = t[1], v3 = t[2];
    const int32_t v4 = int32_t(nv + face_edge[size_t(3 * f)]);
    const int32_t v5 = int32_t(nv + face_edge[size_t(3 * f + 1)]);
    const int32_t v6 = int32_t(nv + face_edge[size_t(3 * f + 2)]);
    const int32_t nt[12] = {v1, v4, v6, v4, v2, v5, v5, v3, v6, v4, v5, v6};
    out_tri->insert(out_tri->end(), nt, nt + 12);
  }
  return ne;
}

}  // namespace acvd
