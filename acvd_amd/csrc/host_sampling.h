// Host-side initial sampling of the reference (sequential by nature).
#pragma once

#include <cstdint>

namespace acvd {

// InitSamples with sampling type 1 (vtkUniformClustering.h:1123-1175, default
// type :1409) = ComputeInitialRandomSampling (:1177-1316) over one mesh with
// the metric's FP64 item weights.  labels receives 0..K-1, K = null.  With
// fixed clusters (SetFixedClusters, :1186-1191) item fixed[k] starts in
// cluster k and the grown regions take the ids n_fixed..K-1.
// Returns false on an internal inconsistency (never for valid meshes).
bool reference_initial_sampling(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, int metric,
                                const double* indicator, double gradation, int32_t K, int32_t* labels,
                                int64_t n_fixed = 0, const int32_t* fixed = nullptr);

}  // namespace acvd
