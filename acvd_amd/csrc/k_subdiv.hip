// Device midpoint subdivision (vtkSurface::Subdivide,
// Common/vtkSurface.cxx:605-677), the step vtkDiscreteRemeshing::
// CheckSubsamplingRatio (DiscreteRemeshing/vtkDiscreteRemeshing.h:840-875)
// repeats while the input has fewer than SubsamplingThreshold * K vertices.
//
// The midpoints are numbered in the reference's edge-creation order: faces in
// order, edges (v0,v1), (v1,v2), (v2,v0) of each face, an edge created at its
// first sighting (Common/vtkSurfaceBase.cxx:1166-1221); faces whose first two
// vertices coincide are inactive and create nothing, a self-loop edge is not
// created.  With half-edge h = 3 f + j, the edge of an undirected vertex pair
// is numbered by its smallest h, so the numbering is a parallel computation:
//   1. key(h) = lo * nv + hi of the pair (a sentinel for no edge);
//   2. a stable radix sort of (key, h) puts each pair's half-edges together,
//      smallest h first;
//   3. the first of each run flags its h; an exclusive scan of the flags in h
//      order numbers the edges in creation order;
//   4. every half-edge takes its run's number; the run's first entry writes
//      the midpoint, oriented as the first sighting (a, b): 0.5 * (x_a + x_b),
//      the indicator likewise (vtkDiscreteRemeshing.h:700-745);
//   5. an exclusive scan over the active faces places their four children
//      (v0,m01,m20), (m01,v1,m12), (m12,v2,m20), (m01,m12,m20).
// Every step is a pass over the half-edges or the faces (HBM-bound, a few
// bytes per half-edge); the sort is hipcub's radix sort over
// ceil(log2(nv^2 + 1)) key bits.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "acvd_ctx.h"

namespace acvd {

namespace {

inline unsigned grid_for(int64_t n) { return unsigned((n + kBlock - 1) / kBlock); }

__global__ void k_sub_keys(int64_t nf, int64_t nv, const int32_t* __restrict__ tri,
                           unsigned long long* __restrict__ keys, int32_t* __restrict__ vals,
                           int32_t* __restrict__ face_edge, int32_t* __restrict__ active) {
  const int64_t h = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (h >= 3 * nf) return;
  const int64_t f = h / 3;
  const int j = int(h - 3 * f);
  const int32_t* t = tri + 3 * f;
  const bool on = t[0] != t[1];
  const int32_t a = t[j], b = t[j == 2 ? 0 : j + 1];
  const unsigned long long sentinel = (unsigned long long)nv * (unsigned long long)nv;
  unsigned long long k = sentinel;
  if (on && a != b) {
    const unsigned long long lo = (unsigned long long)min(a, b), hi = (unsigned long long)max(a, b);
    k = lo * (unsigned long long)nv + hi;
  }
  keys[h] = k;
  vals[h] = int32_t(h);
  face_edge[h] = -1;
  if (j == 0) active[f] = on ? 1 : 0;
}

// flag[h] = 1 for the first half-edge (smallest h) of each vertex pair
__global__ void k_sub_heads(int64_t n3, unsigned long long sentinel, const unsigned long long* __restrict__ keys,
                            const int32_t* __restrict__ vals, int32_t* __restrict__ flag) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  const unsigned long long k = keys[i];
  flag[vals[i]] = (k != sentinel && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
}

__global__ void k_sub_edges(int64_t n3, int64_t nv, unsigned long long sentinel,
                            const unsigned long long* __restrict__ keys, const int32_t* __restrict__ vals,
                            const int32_t* __restrict__ num, const int32_t* __restrict__ tri,
                            const double* __restrict__ xyz, double* __restrict__ xyz_out,
                            const double* __restrict__ ind, double* __restrict__ ind_out,
                            int32_t* __restrict__ face_edge) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  const unsigned long long k = keys[i];
  if (k == sentinel) return;
  int64_t head = i;  // runs are as long as the number of faces on the edge
  while (head > 0 && keys[head - 1] == k) --head;
  const int32_t h0 = vals[head];
  const int32_t e = num[h0];
  face_edge[vals[i]] = e;
  if (head != i) return;
  const int64_t f = h0 / 3;
  const int j = int(h0 - 3 * f);
  const int32_t a = tri[3 * f + j], b = tri[3 * f + (j == 2 ? 0 : j + 1)];
  const int64_t o = nv + e;
  for (int c = 0; c < 3; ++c) xyz_out[3 * o + c] = 0.5 * (xyz[3 * int64_t(a) + c] + xyz[3 * int64_t(b) + c]);
  if (ind) ind_out[o] = 0.5 * (ind[a] + ind[b]);
}

__global__ void k_sub_faces(int64_t nf, int64_t nv, const int32_t* __restrict__ tri,
                            const int32_t* __restrict__ face_edge, const int32_t* __restrict__ fpos,
                            int32_t* __restrict__ tri_out) {
  const int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int32_t* t = tri + 3 * f;
  if (t[0] == t[1]) return;
  const int32_t v1 = t[0], v2 = t[1], v3 = t[2];
  const int32_t v4 = int32_t(nv + face_edge[3 * f]);
  const int32_t v5 = int32_t(nv + face_edge[3 * f + 1]);
  const int32_t v6 = int32_t(nv + face_edge[3 * f + 2]);
  int32_t* o = tri_out + 12 * int64_t(fpos[f]);
  o[0] = v1; o[1] = v4; o[2] = v6;
  o[3] = v4; o[4] = v2; o[5] = v5;
  o[6] = v5; o[7] = v3; o[8] = v6;
  o[9] = v4; o[10] = v5; o[11] = v6;
}

inline int key_bits(int64_t nv) {
  const unsigned long long top = (unsigned long long)nv * (unsigned long long)nv;  // the sentinel
  int b = 1;
  while (b < 64 && (top >> b) != 0) ++b;
  return b;
}

}  // namespace

hipError_t launch_sub_keys(hipStream_t s, int64_t nf, int64_t nv, const int32_t* tri, unsigned long long* keys,
                           int32_t* vals, int32_t* face_edge, int32_t* active) {
  if (nf > 0) k_sub_keys<<<grid_for(3 * nf), kBlock, 0, s>>>(nf, nv, tri, keys, vals, face_edge, active);
  return hipGetLastError();
}

hipError_t sort_sub_keys(hipStream_t s, int64_t n3, int64_t nv, unsigned long long* k0, unsigned long long* k1,
                         int32_t* v0, int32_t* v1, void** tmp, size_t* tmp_bytes) {
  const int bits = key_bits(nv);
  size_t need = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, need, k0, k1, v0, v1, int(n3), 0, bits, s);
  if (e != hipSuccess) return e;
  if (need > *tmp_bytes) {
    if (*tmp) (void)hipFree(*tmp);
    *tmp = nullptr;
    e = hipMalloc(tmp, need);
    if (e != hipSuccess) return e;
    *tmp_bytes = need;
  }
  return hipcub::DeviceRadixSort::SortPairs(*tmp, need, k0, k1, v0, v1, int(n3), 0, bits, s);
}

hipError_t launch_sub_heads(hipStream_t s, int64_t n3, int64_t nv, const unsigned long long* keys,
                            const int32_t* vals, int32_t* flag) {
  const unsigned long long sentinel = (unsigned long long)nv * (unsigned long long)nv;
  if (n3 > 0) k_sub_heads<<<grid_for(n3), kBlock, 0, s>>>(n3, sentinel, keys, vals, flag);
  return hipGetLastError();
}

hipError_t launch_sub_edges(hipStream_t s, int64_t n3, int64_t nv, const unsigned long long* keys,
                            const int32_t* vals, const int32_t* num, const int32_t* tri, const double* xyz,
                            double* xyz_out, const double* ind, double* ind_out, int32_t* face_edge) {
  const unsigned long long sentinel = (unsigned long long)nv * (unsigned long long)nv;
  if (n3 > 0)
    k_sub_edges<<<grid_for(n3), kBlock, 0, s>>>(n3, nv, sentinel, keys, vals, num, tri, xyz, xyz_out, ind, ind_out,
                                                face_edge);
  return hipGetLastError();
}

hipError_t launch_sub_faces(hipStream_t s, int64_t nf, int64_t nv, const int32_t* tri, const int32_t* face_edge,
                            const int32_t* fpos, int32_t* tri_out) {
  if (nf > 0) k_sub_faces<<<grid_for(nf), kBlock, 0, s>>>(nf, nv, tri, face_edge, fpos, tri_out);
  return hipGetLastError();
}

}  // namespace acvd
