// Host-side mesh preparation of Remesh() (see host_mesh.cpp).
#pragma once

#include <cstdint>
#include <vector>

#include "host_sampling.h"

namespace acvd {

// Edge table of a triangle mesh in the reference's order (vtkSurfaceBase:
// edges created at their first sighting, faces in order, edges (v0,v1),
// (v1,v2), (v2,v0); Common/vtkSurfaceBase.cxx:1166-1221).  poly1 / poly2: the
// first two faces of each edge (-1: none); vedge: the edges of each vertex in
// creation order (vptr offsets); face_edge: the edge of each face side.
struct EdgeTable {
  std::vector<int32_t> ev1, ev2, poly1, poly2, face_edge, vedge;
  std::vector<int64_t> vptr;
};
void edge_table(int64_t nv, int64_t nf, const int32_t* tri, EdgeTable* out);

// One level of midpoint subdivision of a triangle mesh (vtkSurface::Subdivide,
// Common/vtkSurface.cxx:605-677): the nv input vertices, then one vertex per
// edge at its midpoint in edge-creation order; four faces per input face.
// parents receives the two ends of each new vertex's edge (2 per edge).
// Returns the number of edges.
int64_t subdivide_mesh(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, std::vector<double>* out_xyz,
                       std::vector<int32_t>* out_tri, std::vector<int32_t>* parents);

// IsVertexManifold (Common/vtkSurfaceBase.cxx:259-317) of every vertex of a
// triangle mesh: at least two incident edges, each in exactly two faces
// (IsEdgeManifold, Common/vtkSurfaceBase.h:521-528: a boundary edge is not),
// and one fan of faces around the vertex through all of them.
void vertex_manifold(int64_t nv, int64_t nf, const int32_t* tri, std::vector<uint8_t>* ok);

// One step of Remesh's ForceManifold loop (vtkDiscreteRemeshing.h:937-946):
// DetectNonManifoldOutputVertices (:165-383) on the output triangulation
// (vertices = clusters) with the wire edges BuildDelaunayTriangulation adds
// under ForceManifold (every pair of clusters adjacent in the input without
// an output edge, :1114-1131).  A non-manifold cluster none of whose items is
// non-manifold in the input is an issue: it and its output neighbours are
// unfrozen (all others frozen) and a new cluster takes its first item (its
// only item's first ring neighbour whose cluster holds more than one item
// when it has one item; no new cluster when there is none).  labels
// (0..K-1, other values: none) and K are updated; frozen receives the flags
// of the K clusters after the step.  Returns the number of issues.
int64_t force_manifold_step(int64_t nv, const VertexTopology& topo, const std::vector<uint8_t>& in_ok,
                            int64_t n_out, const int32_t* out_tri, int32_t* K, int32_t* labels,
                            std::vector<int32_t>* frozen, bool verbose);

// FixMeshBoundaries (vtkDiscreteRemeshing.h:425-579, BoundaryFixing on):
// each input boundary edge whose ends lie in two clusters adds an output
// vertex at its midpoint and the triangle (vertex, cluster 1, cluster 2); each
// run of boundary edges inside one cluster adds the triangle (cluster, first
// two cut-edge vertices met); a vertex between two cut edges adds (cluster,
// their vertices) under the reference's GetNumberOfBoundaries test.  labels:
// 0..K-1 (others: none).  Appends to out_xyz / out_tri; true when the output
// changed.
bool fix_mesh_boundaries(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, const int32_t* labels,
                         int32_t K, std::vector<double>* out_xyz, std::vector<int32_t>* out_tri);

}  // namespace acvd
