// Device layout of the fixed-assignment energy (k_energy.hip): meshes packed
// back to back, items [item_off, item_off + n), clusters [clus_off, + K),
// faces [face_off, + nf) with global vertex ids.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace acvd {

struct EnergyMesh {
  int64_t item_off;
  int64_t n_items;
  int64_t clus_off;
  int64_t face_off;
  int32_t K;
  int32_t metric;  // 0 isotropic (CVT), 1 QEM
  int32_t level;   // rank truncation of the QEM point
  int32_t pad;
};

struct EnergyArgs {
  int64_t n_items;
  int64_t n_clusters;
  const EnergyMesh* meshes;
  const int32_t* item_mesh;  // null for one mesh
  const int32_t* clus_mesh;  // null for one mesh
  const int32_t* tri;        // 3 global vertex ids per face
  const int32_t* vfptr;      // vertex -> incident faces (mesh-local face ids)
  const int32_t* vfcol;
  // inputs of one call
  const float* xyz;          // 3 per item
  const float* w;            // 1 per item
  const int32_t* labels;     // mesh-local cluster ids, K or out of range = none
  // state kept from the forward for the backward
  double* acc;               // 13 per cluster: W, S[3], Q[9]
  double* pts;               // 6 per cluster: centroid, point
  double* adj;               // 12 per cluster: dE/dQ[9], (dE/dc)/W
};

hipError_t energy_forward(hipStream_t s, const EnergyArgs& a, double* energy, int n_meshes);
// Largest cluster count of energy_forward_cvt_noalloc (its static scratch).
constexpr int32_t kNoAllocMaxClusters = 65536;
// energy_forward for one isotropic mesh with every per-call buffer in the
// library's static device scratch of the current device (no allocation).
hipError_t energy_forward_cvt_noalloc(hipStream_t s, int64_t n, int32_t K, const float* xyz, const float* w,
                                      const int32_t* labels, double* energy);
hipError_t energy_backward(hipStream_t s, const EnergyArgs& a, const double* grad_energy, float* dxyz, float* dw);

}  // namespace acvd
