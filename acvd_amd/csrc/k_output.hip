// Output mesh of a clustering (vertex clustering branch of
// vtkDiscreteRemeshing::BuildDelaunayTriangulation,
// DiscreteRemeshing/vtkDiscreteRemeshing.h:1002-1152):
//   - one output vertex per cluster at its cluster point; an empty cluster
//     takes the point of the first non-empty cluster (:1012-1035);
//   - each input face whose three vertices lie in three distinct non-null
//     clusters yields the triangle of those clusters, in the face's vertex
//     order (GetDualItemNeighbourClusters, :974-988), unless a triangle on the
//     same three clusters was already added by an earlier face (AddFace /
//     IsFace, :992-1000, Common/vtkSurfaceBase.cxx IsFace ignores
//     orientation).
// Duplicates are removed by a stable sort of the unordered-triple keys: the
// first face of each run is the first occurrence in face order.
// Optional quadric placement (Examples/ACVD.cxx:217-291): each cluster sums
// the full quadrics of the faces around each of its vertices (a face counted
// once per vertex of the cluster, i.e. the sum of the vertex quadrics) and
// moves its output vertex by the rank-truncated step at the given level.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "acvd_ctx.h"

namespace acvd {

namespace {

inline unsigned grid_for(int64_t n) { return unsigned((n + kBlock - 1) / kBlock); }

// key of the unordered triple (c0, c1, c2), each < 2^21; ~0 for no triangle
__global__ void k_dual_keys(int64_t nf, const int32_t* __restrict__ tri, const int32_t* __restrict__ labels,
                            int64_t item_off, int64_t clus_off, unsigned long long* keys, int32_t* face_ids) {
  int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int32_t c[3];
  bool ok = true;
  for (int j = 0; j < 3; ++j) {
    const int32_t l = labels[tri[3 * f + j]];
    ok &= l != kNull;
    c[j] = int32_t(l - clus_off);
  }
  ok = ok && c[0] != c[1] && c[1] != c[2] && c[0] != c[2];
  unsigned long long key = ~0ull;
  if (ok) {
    uint32_t a = uint32_t(c[0]), b = uint32_t(c[1]), d = uint32_t(c[2]), t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > d) { t = b; b = d; d = t; }
    if (a > b) { t = a; a = b; b = t; }
    key = (uint64_t(a) << 42) | (uint64_t(b) << 21) | uint64_t(d);
  }
  keys[f] = key;
  face_ids[f] = int32_t(f);
}

// flag[f] = 1 for the first face (in face order) of each run of equal keys
__global__ void k_dual_heads(int64_t nf, const unsigned long long* __restrict__ skeys,
                             const int32_t* __restrict__ sfaces, int32_t* flag) {
  int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= nf) return;
  const unsigned long long k = skeys[r];
  if (k == ~0ull) return;
  if (r > 0 && skeys[r - 1] == k) return;
  flag[sfaces[r]] = 1;
}

__global__ void k_dual_emit(int64_t nf, const int32_t* __restrict__ tri, const int32_t* __restrict__ labels,
                            int64_t item_off, int64_t clus_off, const int32_t* __restrict__ flag,
                            const int32_t* __restrict__ pos, int32_t* out_tri) {
  int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= nf || !flag[f]) return;
  const int32_t p = pos[f];
  for (int j = 0; j < 3; ++j) out_tri[3 * p + j] = int32_t(labels[tri[3 * f + j]] - clus_off);
}

// first non-empty cluster of the mesh (atomicMin over cluster ids)
__global__ void k_first_nonempty(const int32_t* __restrict__ sizes, int64_t clus_off, int32_t K, int32_t* first) {
  int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= K) return;
  if (sizes[clus_off + r] > 0) atomicMin(first, int32_t(r));
}

__global__ void k_fix_empty_points(const int32_t* __restrict__ sizes, int64_t clus_off, int32_t K,
                                   const int32_t* __restrict__ first, double* pts) {
  int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= K) return;
  if (sizes[clus_off + r] > 0) return;
  const int32_t v = *first;
  if (v >= K) return;
  for (int k = 0; k < 3; ++k) pts[3 * r + k] = pts[3 * int64_t(v) + k];
}

// per-cluster sum of the vertex quadrics (int64 fixed point, order-free)
__global__ void k_cluster_quadrics(int64_t n, int64_t item_off, const int32_t* __restrict__ labels,
                                   int64_t clus_off, const float* __restrict__ q, double sq,
                                   unsigned long long* acc) {
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t i = item_off + t;
  const int32_t l = labels[i];
  if (l == kNull) return;
  const int64_t c = l - clus_off;
  for (int k = 0; k < 10; ++k) atomicAdd(&acc[10 * c + k], (unsigned long long)to_fixed(double(q[10 * i + k]) * sq));
}

__global__ void k_place(int32_t K, const unsigned long long* __restrict__ acc, double isq, int level, double* pts) {
  int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= K) return;
  double Q[9];
  for (int k = 0; k < 9; ++k) Q[k] = double(int64_t(acc[10 * r + k])) * isq;
  double c[3] = {pts[3 * r], pts[3 * r + 1], pts[3 * r + 2]};
  double D[3];
  rep_point_delta(Q, c, level, D);
  for (int k = 0; k < 3; ++k) pts[3 * r + k] = c[k] + D[k];
}

}  // namespace

// Builds the output triangles of one mesh (faces tri[0..nf) with global
// vertex ids, global labels).  out_tri must hold 3*nf entries; *out_nf receives the
// count.  Temporary device memory is allocated here.
hipError_t build_dual_faces(hipStream_t s, int64_t nf, const int32_t* tri, const int32_t* labels, int64_t item_off,
                            int64_t clus_off, int32_t* out_tri, int64_t* out_nf) {
  *out_nf = 0;
  if (nf == 0) return hipSuccess;
  unsigned long long *k0 = nullptr, *k1 = nullptr;
  int32_t *f0 = nullptr, *f1 = nullptr, *flag = nullptr, *pos = nullptr;
  void* tmp = nullptr;
  size_t tb = 0, tb2 = 0;
  hipError_t e = hipSuccess;
  auto cleanup = [&]() {
    (void)hipFree(k0); (void)hipFree(k1); (void)hipFree(f0); (void)hipFree(f1);
    (void)hipFree(flag); (void)hipFree(pos); (void)hipFree(tmp);
  };
#define ACVD_OUT_TRY(x) do { e = (x); if (e != hipSuccess) { cleanup(); return e; } } while (0)
  ACVD_OUT_TRY(hipMalloc(&k0, nf * sizeof(unsigned long long)));
  ACVD_OUT_TRY(hipMalloc(&k1, nf * sizeof(unsigned long long)));
  ACVD_OUT_TRY(hipMalloc(&f0, nf * sizeof(int32_t)));
  ACVD_OUT_TRY(hipMalloc(&f1, nf * sizeof(int32_t)));
  ACVD_OUT_TRY(hipMalloc(&flag, (nf + 1) * sizeof(int32_t)));
  ACVD_OUT_TRY(hipMalloc(&pos, (nf + 1) * sizeof(int32_t)));
  hipLaunchKernelGGL(k_dual_keys, dim3(grid_for(nf)), dim3(kBlock), 0, s, nf, tri, labels, item_off, clus_off, k0, f0);
  ACVD_OUT_TRY(hipGetLastError());
  ACVD_OUT_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, f0, f1, int(nf), 0, 64, s));
  ACVD_OUT_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, flag, pos, int(nf + 1), s));
  ACVD_OUT_TRY(hipMalloc(&tmp, tb > tb2 ? tb : tb2));
  ACVD_OUT_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, f0, f1, int(nf), 0, 64, s));
  ACVD_OUT_TRY(hipMemsetAsync(flag, 0, (nf + 1) * sizeof(int32_t), s));
  hipLaunchKernelGGL(k_dual_heads, dim3(grid_for(nf)), dim3(kBlock), 0, s, nf, k1, f1, flag);
  ACVD_OUT_TRY(hipGetLastError());
  ACVD_OUT_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tb2, flag, pos, int(nf + 1), s));
  hipLaunchKernelGGL(k_dual_emit, dim3(grid_for(nf)), dim3(kBlock), 0, s, nf, tri, labels, item_off, clus_off, flag,
                     pos, out_tri);
  ACVD_OUT_TRY(hipGetLastError());
  int32_t cnt = 0;
  ACVD_OUT_TRY(hipMemcpyAsync(&cnt, pos + nf, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  ACVD_OUT_TRY(hipStreamSynchronize(s));
  *out_nf = cnt;
  cleanup();
#undef ACVD_OUT_TRY
  return hipSuccess;
}

hipError_t fix_empty_cluster_points(hipStream_t s, const int32_t* sizes, int64_t clus_off, int32_t K, int32_t* first,
                                    double* pts) {
  if (K == 0) return hipSuccess;
  hipLaunchKernelGGL(k_first_nonempty, dim3(grid_for(K)), dim3(kBlock), 0, s, sizes, clus_off, K, first);
  hipLaunchKernelGGL(k_fix_empty_points, dim3(grid_for(K)), dim3(kBlock), 0, s, sizes, clus_off, K, first, pts);
  return hipGetLastError();
}

hipError_t quadric_placement(hipStream_t s, int64_t n, int64_t item_off, const int32_t* labels, int64_t clus_off,
                             int32_t K, const float* q, double sq, unsigned long long* acc, int level, double* pts) {
  if (K == 0) return hipSuccess;
  hipError_t e = hipMemsetAsync(acc, 0, size_t(K) * 10 * sizeof(unsigned long long), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cluster_quadrics, dim3(grid_for(n)), dim3(kBlock), 0, s, n, item_off, labels, clus_off, q, sq,
                     acc);
  hipLaunchKernelGGL(k_place, dim3(grid_for(K)), dim3(kBlock), 0, s, K, acc, 1.0 / sq, level, pts);
  return hipGetLastError();
}

}  // namespace acvd
