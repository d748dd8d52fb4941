// Host-side initial sampling of the reference (sequential by nature).
#pragma once

#include <cstdint>
#include <random>
#include <vector>

#include "host_parallel.h"

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

// The reference's sampling weights from the vertex areas in w (FP64, summed
// in GetVertexNeighbourFaces order): times indicator^gradation (isotropic:
// whenever an indicator is set; QEM: when gradation > 0), then clamped to
// [avg/R, avg*R], R = 1e5 (isotropic) or 1e4 (QEM).  In place.
void sampling_weights(int64_t nv, int metric, const double* indicator, double gradation, double* w);

// The shuffled item order of ComputeInitialRandomSampling
// (vtkUniformClustering.h:1202-1207: std::mt19937 seeded 0, then
// swap(Items[i], Items[rng() % n]) for i = n-1 .. 1) and the generator after
// it (the second half of the sampling draws on).  It depends on n alone.
struct SamplingOrder {
  huge_vector<int32_t> items;
  std::mt19937 rng;
};
void sampling_order(int64_t n, SamplingOrder* out);

// ComputeInitialRandomSampling (vtkUniformClustering.h:1177-1316) given the
// rings, ring[rbeg[v] .. rbeg[v] + rlen[v]) in edge-creation order, and the
// weights.  labels receives 0..K-1, K = null.  order: the shuffled order
// for nv items (sampling_order), or null to compute it here; it is used up.
bool region_growing(int64_t nv, const int64_t* rbeg, const int32_t* rlen, const int32_t* ring, const double* w,
                    int32_t K, int32_t* labels, int64_t n_fixed, const int32_t* fixed, SamplingOrder* order = nullptr);

// InitSamples with sampling type 0 (vtkUniformClustering.h:1123-1159): one
// item per cluster drawn with vtkMath::Random seeded 5000 (an item already
// taken is drawn again), every other item null.  labels receives 0..K-1,
// K = null.  The reference's draw can land one past the last item (it
// rounds a draw in [0, nv]); such a draw is drawn again.  Returns false when
// K exceeds the items.
bool reference_seed_sampling(int64_t nv, int32_t K, int32_t* labels);

}  // namespace acvd
