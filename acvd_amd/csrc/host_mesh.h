// Host-side mesh preparation of Remesh() (see host_mesh.cpp).
#pragma once

#include <cstdint>
#include <vector>

namespace acvd {

// One level of midpoint subdivision of a triangle mesh (vtkSurface::Subdivide,
// Common/vtkSurface.cxx:605-677): the nv input vertices, then one vertex per
// edge at its midpoint in edge-creation order; four faces per input face.
// parents receives the two ends of each new vertex's edge (2 per edge).
// Returns the number of edges.
int64_t subdivide_mesh(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, std::vector<double>* out_xyz,
                       std::vector<int32_t>* out_tri, std::vector<int32_t>* parents);

}  // namespace acvd
