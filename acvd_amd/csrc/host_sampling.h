// Host-side initial sampling of the reference (sequential by nature).
#pragma once

#include <cstdint>
#include <vector>

namespace acvd {

// Vertex topology in the reference's order: the faces of each vertex in
// increasing face id (GetVertexNeighbourFaces order of the item weights,
// Common/vtkSurfaceBase.cxx:964-984) and the ring of each vertex in
// edge-creation order (GetVertexNeighbours, :1057-1068: edges (v0,v1),
// (v1,v2), (v2,v0) of each face in turn, :1166-1221, 1407-1468).
struct VertexTopology {
  std::vector<int64_t> vfptr, rptr;  // nv + 1 offsets
  std::vector<int32_t> vf, ring;
};
void vertex_topology(int64_t nv, int64_t nf, const int32_t* tri, VertexTopology* out);

// InitSamples with sampling type 1 (vtkUniformClustering.h:1123-1175, default
// type :1409) = ComputeInitialRandomSampling (:1177-1316) over one mesh with
// the metric's FP64 item weights.  labels receives 0..K-1, K = null.  With
// fixed clusters (SetFixedClusters, :1186-1191) item fixed[k] starts in
// cluster k and the grown regions take the ids n_fixed..K-1.
// Returns false on an internal inconsistency (never for valid meshes).
bool reference_initial_sampling(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, int metric,
                                const double* indicator, double gradation, int32_t K, int32_t* labels,
                                int64_t n_fixed = 0, const int32_t* fixed = nullptr);

// InitSamples with sampling type 0 (vtkUniformClustering.h:1123-1159): one
// item per cluster drawn with vtkMath::Random seeded 5000 (an item already
// taken is drawn again), every other item null.  labels receives 0..K-1,
// K = null.  The reference's draw can land one past the last item (it
// rounds a draw in [0, nv]); such a draw is drawn again.  Returns false when
// K exceeds the items.
bool reference_seed_sampling(int64_t nv, int32_t K, int32_t* labels);

}  // namespace acvd
