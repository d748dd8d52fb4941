// Device build of the clustering inputs:
//   (a) the item/edge adjacency as a CSR with rows sorted by neighbour id,
//       plus a vertex -> face CSR (rows sorted by face id), both from the
//       triangle list (the edge graph that vtkSurfaceBase builds in
//       Common/vtkSurfaceBase.cxx:1166-1221, 1407-1468);
//   (b) per-vertex items: FP32 position and clamped weight
//       w = clamp(area * indicator^gradation) with area = sum of adjacent face
//       areas / 3 (Common/vtkSurface.cxx:1280-1283, 1342-1359;
//       vtkIsotropicMetricForClustering.h:249-291 clamp ratio 1e5,
//       vtkQEMetricForClustering.h:324-362 clamp ratio 1e4) and, for QEM
//       meshes, the vertex quadric = sum of the adjacent triangles' plane
//       quadrics (vtkQEMetricForClustering.h:151-167,
//       Common/vtkQuadricTools.cxx:56-79), 10 coefficients;
//   (c) deterministic per-mesh reductions (fixed block partition, fixed tree)
//       and the power-of-two fixed-point scales derived from them.
// All FP64 accumulation happens in a fixed order so that a build is
// reproducible run to run.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "acvd_ctx.h"

namespace acvd {

namespace {

inline unsigned grid_for(int64_t n) { return unsigned((n + kBlock - 1) / kBlock); }

__global__ void k_csr_count(int64_t nf, const int32_t* __restrict__ tri, int32_t* deg2, int32_t* vfdeg) {
  int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  for (int j = 0; j < 3; ++j) {
    int32_t v = tri[3 * f + j];
    atomicAdd(&deg2[v], 2);
    atomicAdd(&vfdeg[v], 1);
  }
}

__global__ void k_csr_fill(int64_t nf, const int32_t* __restrict__ tri, const int32_t* __restrict__ tptr,
                           int32_t* tfill, int32_t* tcol, const int32_t* __restrict__ vfptr, int32_t* vffill,
                           int32_t* vfcol) {
  int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int32_t v[3] = {tri[3 * f], tri[3 * f + 1], tri[3 * f + 2]};
  for (int j = 0; j < 3; ++j) {
    int32_t a = v[j], b = v[(j + 1) % 3], c = v[(j + 2) % 3];
    int32_t slot = tptr[a] + atomicAdd(&tfill[a], 2);
    tcol[slot] = b;
    tcol[slot + 1] = c;
    int32_t s2 = vfptr[a] + atomicAdd(&vffill[a], 1);
    vfcol[s2] = int32_t(f);
  }
}

// insertion sort of a short row held in global memory
__device__ void sort_row(int32_t* r, int32_t n) {
  for (int32_t i = 1; i < n; ++i) {
    int32_t x = r[i];
    int32_t j = i - 1;
    while (j >= 0 && r[j] > x) {
      r[j + 1] = r[j];
      --j;
    }
    r[j + 1] = x;
  }
}

__global__ void k_csr_sort_unique(int64_t n, const int32_t* __restrict__ tptr, int32_t* tcol, int32_t* deg,
                                  const int32_t* __restrict__ vfptr, int32_t* vfcol) {
  int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int32_t* r = tcol + tptr[v];
  int32_t len = tptr[v + 1] - tptr[v];
  sort_row(r, len);
  int32_t u = 0;
  for (int32_t i = 0; i < len; ++i) {
    if (i > 0 && r[i] == r[i - 1]) continue;
    r[u++] = r[i];
  }
  deg[v] = u;
  sort_row(vfcol + vfptr[v], vfptr[v + 1] - vfptr[v]);
}

__global__ void k_csr_compact(int64_t n, const int32_t* __restrict__ tptr, const int32_t* __restrict__ tcol,
                              const int32_t* __restrict__ row_ptr, int32_t* col) {
  int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int32_t len = row_ptr[v + 1] - row_ptr[v];
  const int32_t* src = tcol + tptr[v];
  int32_t* dst = col + row_ptr[v];
  for (int32_t i = 0; i < len; ++i) dst[i] = src[i];
}

// Per-vertex item build (one thread per vertex, faces visited in increasing
// face id).  raw_w receives the unclamped weight; xw.w is filled later.
__global__ void k_item_build(SweepArgs a, const double* __restrict__ pos, const int32_t* __restrict__ tri,
                             const int32_t* __restrict__ vfptr, const int32_t* __restrict__ vfcol,
                             const double* __restrict__ indicator, double* raw_w, float4* xw, float* q) {
  int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= a.n_items) return;
  const int m = a.item_mesh ? a.item_mesh[v] : 0;
  const DevMesh dm = a.meshes[m];
  double area = 0.0;
  double Q[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const bool qem = dm.metric == 1;
  const bool quadrics = q != nullptr;  // vertex quadrics of every mesh (sweep and output placement)
  for (int32_t k = vfptr[v]; k < vfptr[v + 1]; ++k) {
    const int64_t f = vfcol[k];
    const int32_t i0 = tri[3 * f], i1 = tri[3 * f + 1], i2 = tri[3 * f + 2];
    const double p1[3] = {pos[3 * i0], pos[3 * i0 + 1], pos[3 * i0 + 2]};
    const double p2[3] = {pos[3 * i1], pos[3 * i1 + 1], pos[3 * i1 + 2]};
    const double p3[3] = {pos[3 * i2], pos[3 * i2 + 1], pos[3 * i2 + 2]};
    double ea = (p1[0] - p2[0]) * (p1[0] - p2[0]) + (p1[1] - p2[1]) * (p1[1] - p2[1]) + (p1[2] - p2[2]) * (p1[2] - p2[2]);
    double eb = (p2[0] - p3[0]) * (p2[0] - p3[0]) + (p2[1] - p3[1]) * (p2[1] - p3[1]) + (p2[2] - p3[2]) * (p2[2] - p3[2]);
    double ec = (p3[0] - p1[0]) * (p3[0] - p1[0]) + (p3[1] - p1[1]) * (p3[1] - p1[1]) + (p3[2] - p1[2]) * (p3[2] - p1[2]);
    double t = 4.0 * ea * ec - (ea - eb + ec) * (ea - eb + ec);
    area += 0.25 * sqrt(t < 0 ? -t : t) / 3.0;
    if (quadrics) {
      // plane of the triangle: n = (p2-p1) x (p3-p1), d = -n.p1
      double u[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
      double w[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
      double n0 = u[1] * w[2] - u[2] * w[1];
      double n1 = u[2] * w[0] - u[0] * w[2];
      double n2 = u[0] * w[1] - u[1] * w[0];
      double d = -(n0 * p1[0] + n1 * p1[1] + n2 * p1[2]);
      Q[0] += n0 * n0; Q[1] += n0 * n1; Q[2] += n0 * n2; Q[3] += n0 * d;
      Q[4] += n1 * n1; Q[5] += n1 * n2; Q[6] += n1 * d;
      Q[7] += n2 * n2; Q[8] += n2 * d;  Q[9] += d * d;
    }
  }
  double wgt = area;
  if (dm.has_indicator && (!qem || dm.gradation > 0.0)) wgt *= pow(indicator[v], dm.gradation);
  raw_w[v] = wgt;
  xw[v] = make_float4(float(pos[3 * v]), float(pos[3 * v + 1]), float(pos[3 * v + 2]), 0.0f);
  if (quadrics)
    for (int k = 0; k < 10; ++k) q[10 * v + k] = float(Q[k]);
}

// Deterministic segmented block reduction.  Block b covers the item range
// block_tab[b] = (begin, end) of a single mesh.  what:
//   0: sum raw_w                      (average weight before the clamp)
//   1: sum w, sum w * max|x|, sum max|q|, max w   (scales)
__global__ void k_seg_reduce(const int2* __restrict__ block_tab, SweepArgs a, const double* __restrict__ raw_w,
                             int what, double* partial) {
  __shared__ double sh[4][kBlock];
  const int2 r = block_tab[blockIdx.x];
  double v0 = 0, v1 = 0, v2 = 0, v3 = 0;
  // each thread folds a fixed strided subset in increasing order
  for (int64_t i = int64_t(r.x) + threadIdx.x; i < r.y; i += kBlock) {
    if (what == 0) {
      v0 += raw_w[i];
    } else {
      float4 p = a.xw[i];
      double w = p.w;
      double mx = fmax(fabs(double(p.x)), fmax(fabs(double(p.y)), fabs(double(p.z))));
      v0 += w;
      v1 += w * mx;
      if (a.q) {
        double qm = 0;
        for (int k = 0; k < 10; ++k) qm = fmax(qm, fabs(double(a.q[10 * i + k])));
        v2 += qm;
      }
      v3 = fmax(v3, w);
    }
  }
  sh[0][threadIdx.x] = v0;
  sh[1][threadIdx.x] = v1;
  sh[2][threadIdx.x] = v2;
  sh[3][threadIdx.x] = v3;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
      sh[2][threadIdx.x] += sh[2][threadIdx.x + s];
      sh[3][threadIdx.x] = fmax(sh[3][threadIdx.x], sh[3][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) partial[4 * blockIdx.x + k] = sh[k][0];
}

// power of two p with v * p < 2^60 (v > 0)
__device__ double pow2_scale_below(double v) {
  if (!(v > 0.0)) return 1.0;
  int e;
  frexp(v, &e);  // v = f * 2^e, f in [0.5, 1)
  return ldexp(1.0, 60 - e);
}

// stage 0: per-mesh average of raw weights -> clamp -> xw.w
// stage 1: per-mesh scales
__global__ void k_mesh_fold(SweepArgs a, int n_blocks, const int2* __restrict__ block_tab,
                            const double* __restrict__ partial, int stage, DevMesh* meshes_rw, double* mesh_out) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.n_meshes) return;
  DevMesh dm = meshes_rw[m];
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int b = 0; b < n_blocks; ++b) {
    int2 r = block_tab[b];
    if (r.x < dm.item_off || r.x >= dm.item_off + dm.n_items) continue;
    s0 += partial[4 * b];
    s1 += partial[4 * b + 1];
    s2 += partial[4 * b + 2];
    s3 = fmax(s3, partial[4 * b + 3]);
  }
  if (stage == 0) {
    mesh_out[m] = s0 / double(dm.n_items);  // average
  } else {
    Scales sc;
    sc.sw = pow2_scale_below(s0);
    sc.ss = pow2_scale_below(s1);
    sc.sq = pow2_scale_below(s2);
    sc.isw = 1.0 / sc.sw;
    sc.iss = 1.0 / sc.ss;
    sc.isq = 1.0 / sc.sq;
    dm.scales = sc;
    meshes_rw[m] = dm;
  }
}

__global__ void k_clamp(SweepArgs a, const double* __restrict__ raw_w, const double* __restrict__ avg,
                        float4* xw) {
  int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= a.n_items) return;
  const int m = a.item_mesh ? a.item_mesh[v] : 0;
  const double ratio = a.meshes[m].clamp_ratio;
  const double av = avg[m];
  const double mn = av / ratio, mx = av * ratio;
  double w = raw_w[v];
  if (w > mx) w = mx;
  if (w < mn) w = mn;
  xw[v].w = float(w);
}

// Ring adjacency of every item (bits of neighbour pairs sharing an edge).
__global__ void k_ring_adj(int64_t n, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                           unsigned long long* ring_adj) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ring_adj[i] = ring_adjacency(row_ptr, col, int32_t(i));
}

}  // namespace

hipError_t launch_csr_count(hipStream_t s, int64_t nf, const int32_t* tri, int32_t* deg2, int32_t* vfdeg) {
  if (nf == 0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_count, dim3(grid_for(nf)), dim3(kBlock), 0, s, nf, tri, deg2, vfdeg);
  return hipGetLastError();
}

hipError_t launch_csr_fill(hipStream_t s, int64_t nf, const int32_t* tri, const int32_t* tptr, int32_t* tfill,
                           int32_t* tcol, const int32_t* vfptr, int32_t* vffill, int32_t* vfcol) {
  if (nf == 0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_fill, dim3(grid_for(nf)), dim3(kBlock), 0, s, nf, tri, tptr, tfill, tcol, vfptr,
                     vffill, vfcol);
  return hipGetLastError();
}

hipError_t launch_csr_sort_unique(hipStream_t s, int64_t n, const int32_t* tptr, int32_t* tcol, int32_t* deg,
                                  const int32_t* vfptr, int32_t* vfcol) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_sort_unique, dim3(grid_for(n)), dim3(kBlock), 0, s, n, tptr, tcol, deg, vfptr, vfcol);
  return hipGetLastError();
}

hipError_t launch_ring_adj(hipStream_t s, int64_t n, const int32_t* row_ptr, const int32_t* col,
                           unsigned long long* ring_adj) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ring_adj, dim3(grid_for(n)), dim3(kBlock), 0, s, n, row_ptr, col, ring_adj);
  return hipGetLastError();
}

hipError_t launch_csr_compact(hipStream_t s, int64_t n, const int32_t* tptr, const int32_t* tcol,
                              const int32_t* row_ptr, int32_t* col) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_compact, dim3(grid_for(n)), dim3(kBlock), 0, s, n, tptr, tcol, row_ptr, col);
  return hipGetLastError();
}

// out[0..n] = exclusive prefix sum of in[0..n-1] (out has n + 1 entries: the
// input is read with one trailing zero appended by the caller).
hipError_t exclusive_scan_i32(hipStream_t s, const int32_t* in, int32_t* out, int64_t n, void** tmp,
                              size_t* tmp_bytes) {
  size_t need = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, int(n), s);
  if (e != hipSuccess) return e;
  if (need > *tmp_bytes) {
    if (*tmp) (void)hipFree(*tmp);
    *tmp = nullptr;
    e = hipMalloc(tmp, need);
    if (e != hipSuccess) return e;
    *tmp_bytes = need;
  }
  return hipcub::DeviceScan::ExclusiveSum(*tmp, need, in, out, int(n), s);
}

// ------------------------------------------------------ curvature indicator
// The reference's per-vertex curvature indicator (vtkCurvatureMeasure with
// polynomial fitting on vertices, RingSize 3, as
// vtkDiscreteRemeshing::SamplingPreProcessing sets it up,
// DiscreteRemeshing/vtkDiscreteRemeshing.h:644-653): the faces with a vertex
// within two edges of v (Common/vtkNeighbourhoodComputation.cxx:38-110), a
// quadratic height field fitted to their barycentres in the area-weighted
// normal frame, and sqrt(k1^2 + k2^2) of its shape operator
// (Common/vtkCurvatureMeasure.cxx:188-509).  One thread per vertex, FP64.

// True when row r of the (sorted) CSR holds w.
__device__ inline bool row_has(const int32_t* rp, const int32_t* col, int32_t r, int32_t w) {
  int32_t lo = rp[r], hi = rp[r + 1];
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    const int32_t c = col[mid];
    if (c == w) return true;
    if (c < w) lo = mid + 1;
    else hi = mid;
  }
  return false;
}

// True when w lies within two edges of v.
__device__ inline bool within_two(const int32_t* rp, const int32_t* col, int32_t v, int32_t w) {
  if (w == v || row_has(rp, col, v, w)) return true;
  for (int32_t k = rp[v]; k < rp[v + 1]; ++k)
    if (row_has(rp, col, col[k], w)) return true;
  return false;
}

// Calls f(face) once for every face with a vertex within two edges of v: a
// face is taken at its lowest-id vertex among those within two edges, and
// each such vertex u is visited once (v, its neighbours, then the
// neighbours' neighbours at their first occurrence).
template <class F>
__device__ inline void for_each_star2_face(const int32_t* rp, const int32_t* col, const int32_t* vfptr,
                                           const int32_t* vfcol, const int32_t* tri, int32_t v, F&& f) {
  auto visit = [&](int32_t u) {
    for (int32_t t = vfptr[u]; t < vfptr[u + 1]; ++t) {
      const int32_t face = vfcol[t];
      bool first = true;
      for (int c = 0; c < 3; ++c) {
        const int32_t w = tri[3 * int64_t(face) + c];
        if (w < u && within_two(rp, col, v, w)) {
          first = false;
          break;
        }
      }
      if (first) f(face);
    }
  };
  visit(v);
  for (int32_t k = rp[v]; k < rp[v + 1]; ++k) visit(col[k]);
  for (int32_t k = rp[v]; k < rp[v + 1]; ++k) {
    const int32_t p = col[k];
    for (int32_t t = rp[p]; t < rp[p + 1]; ++t) {
      const int32_t u = col[t];
      if (u == v || row_has(rp, col, v, u)) continue;  // already visited
      bool seen = false;
      for (int32_t k2 = rp[v]; k2 < k && !seen; ++k2) seen = row_has(rp, col, col[k2], u);
      if (!seen) visit(u);
    }
  }
}

__global__ void k_curvature(int64_t off, int64_t n, const double* __restrict__ pos, const int32_t* __restrict__ tri,
                            const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                            const int32_t* __restrict__ vfptr, const int32_t* __restrict__ vfcol,
                            double* __restrict__ out) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int32_t v = int32_t(off + idx);
  // face data: area (Common/vtkSurface.cxx:1280-1283), unit normal, barycentre
  auto face = [&](int32_t f, double* area, double* nrm, double* bary) {
    const int32_t i0 = tri[3 * int64_t(f)], i1 = tri[3 * int64_t(f) + 1], i2 = tri[3 * int64_t(f) + 2];
    const double p1[3] = {pos[3 * i0], pos[3 * i0 + 1], pos[3 * i0 + 2]};
    const double p2[3] = {pos[3 * i1], pos[3 * i1 + 1], pos[3 * i1 + 2]};
    const double p3[3] = {pos[3 * i2], pos[3 * i2 + 1], pos[3 * i2 + 2]};
    double ea = 0, eb = 0, ec = 0;
    for (int k = 0; k < 3; ++k) {
      ea += (p1[k] - p2[k]) * (p1[k] - p2[k]);
      eb += (p2[k] - p3[k]) * (p2[k] - p3[k]);
      ec += (p3[k] - p1[k]) * (p3[k] - p1[k]);
    }
    const double t = 4.0 * ea * ec - (ea - eb + ec) * (ea - eb + ec);
    *area = 0.25 * sqrt(t < 0 ? -t : t);
    const double u[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    const double w[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
    nrm[0] = u[1] * w[2] - u[2] * w[1];
    nrm[1] = u[2] * w[0] - u[0] * w[2];
    nrm[2] = u[0] * w[1] - u[1] * w[0];
    const double ln = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    for (int k = 0; k < 3; ++k) {
      nrm[k] = ln > 0.0 ? nrm[k] / ln : 0.0;
      bary[k] = *area > 0.0 ? (p1[k] + p2[k] + p3[k]) / 3.0 : 0.0;  // GetCellMassProperties
    }
  };
  double sa = 0.0, org[3] = {0, 0, 0}, f0[3] = {0, 0, 0};
  int cnt = 0;
  for_each_star2_face(rp, col, vfptr, vfcol, tri, v, [&](int32_t f) {
    double ar, nr[3], bc[3];
    face(f, &ar, nr, bc);
    sa += ar;
    for (int k = 0; k < 3; ++k) {
      org[k] += ar * bc[k];
      f0[k] += ar * nr[k];
    }
    ++cnt;
  });
  double ind = 0.0;
  const double ln = sqrt(f0[0] * f0[0] + f0[1] * f0[1] + f0[2] * f0[2]);
  if (cnt > 6 && sa > 0.0 && ln > 0.0) {
    // frame: normal f2, tangents from the permuted normal
    double f2[3], f1[3], t0[3], t1[3];
    for (int k = 0; k < 3; ++k) {
      org[k] /= sa;
      f2[k] = f0[k] / ln;
    }
    f1[0] = f2[2];
    f1[1] = f2[0];
    f1[2] = f2[1];
    t0[0] = f1[1] * f2[2] - f1[2] * f2[1];
    t0[1] = f1[2] * f2[0] - f1[0] * f2[2];
    t0[2] = f1[0] * f2[1] - f1[1] * f2[0];
    const double lt = sqrt(t0[0] * t0[0] + t0[1] * t0[1] + t0[2] * t0[2]);
    if (lt > 0.0) {
      for (int k = 0; k < 3; ++k) t0[k] /= lt;
      t1[0] = f2[1] * t0[2] - f2[2] * t0[1];
      t1[1] = f2[2] * t0[0] - f2[0] * t0[2];
      t1[2] = f2[0] * t0[1] - f2[1] * t0[0];
      // normal equations of z ~ [1, x, y, x^2, xy, y^2] . q (upper triangle)
      double M[6][6], r[6], hs = 0.0;
      for (int i = 0; i < 6; ++i) {
        r[i] = 0.0;
        for (int j = 0; j < 6; ++j) M[i][j] = 0.0;
      }
      for_each_star2_face(rp, col, vfptr, vfcol, tri, v, [&](int32_t f) {
        double ar, nr[3], bc[3];
        face(f, &ar, nr, bc);
        double d[3] = {bc[0] - org[0], bc[1] - org[1], bc[2] - org[2]};
        const double x = d[0] * t0[0] + d[1] * t0[1] + d[2] * t0[2];
        const double y = d[0] * t1[0] + d[1] * t1[1] + d[2] * t1[2];
        const double z = d[0] * f2[0] + d[1] * f2[1] + d[2] * f2[2];
        hs += sqrt(x * x + y * y);
        const double c[6] = {1.0, x, y, x * x, x * y, y * y};
#pragma unroll
        for (int i = 0; i < 6; ++i) {
          r[i] += c[i] * z;
#pragma unroll
          for (int j = i; j < 6; ++j) M[i][j] += c[i] * c[j];
        }
      });
      // the columns scaled by the mean |(x, y)| (conditioning, :302-315)
      const double h = hs / double(cnt);
      const double sc[6] = {1.0, h, h, h * h, h * h, h * h};
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        r[i] /= sc[i];
#pragma unroll
        for (int j = i; j < 6; ++j) M[i][j] /= sc[i] * sc[j];
      }
      // Cholesky of the scaled normal matrix
      double L[6][6];
      bool ok = h > 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        double dj = M[j][j];
#pragma unroll
        for (int k = 0; k < j; ++k) dj -= L[j][k] * L[j][k];
        ok = ok && dj > 0.0;
        L[j][j] = sqrt(dj > 0.0 ? dj : 1.0);
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
          double s = M[j][i];
#pragma unroll
          for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
          L[i][j] = s / L[j][j];
        }
      }
      if (ok) {
        double yv[6], q[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
          double s = r[i];
#pragma unroll
          for (int k = 0; k < i; ++k) s -= L[i][k] * yv[k];
          yv[i] = s / L[i][i];
        }
#pragma unroll
        for (int i = 5; i >= 0; --i) {
          double s = yv[i];
#pragma unroll
          for (int k = i + 1; k < 6; ++k) s -= L[k][i] * q[k];
          q[i] = s / L[i][i];
        }
        for (int i = 1; i < 6; ++i) q[i] /= sc[i];
        // shape operator -II I^-1 with the reference's f = 2 q4 / s (:331-347)
        const double E = 1.0 + q[1] * q[1], Fm = q[1] * q[2], G = 1.0 + q[2] * q[2];
        const double sq = sqrt(q[1] * q[1] + 1.0 + q[2] * q[2]);
        const double e = 2.0 * q[3] / sq, fm = 2.0 * q[4] / sq, g = 2.0 * q[5] / sq;
        const double det = E * G - Fm * Fm;  // 1 + q1^2 + q2^2 > 0
        const double B00 = G / det, B01 = -Fm / det, B10 = -Fm / det, B11 = E / det;
        const double a00 = -(e * B00 + fm * B10);
        const double a01 = -(fm * B00 + g * B10);
        const double a11 = -(fm * B01 + g * B11);
        // vtkMath::JacobiN reads the upper triangle: k1^2 + k2^2 is the
        // squared Frobenius norm of the upper-symmetrised matrix
        ind = sqrt(a00 * a00 + a11 * a11 + 2.0 * a01 * a01);
      }
    }
  }
  out[v] = ind;
}

namespace {

// Topology and weights of the reference's initial sampling for the vertices
// [off, off + n) of one mesh, one thread per vertex, from the vertex -> face
// rows (faces in increasing id; a face listed twice for a vertex, or whose
// first two vertices coincide, counts once / not at all):
//  - ring: the neighbours in edge-creation order (edges (v0,v1), (v1,v2),
//    (v2,v0) of each face in turn, a neighbour at its first sighting,
//    Common/vtkSurfaceBase.cxx:1057-1068, 1166-1221), written at the row's
//    place in the adjacency (row_ptr: it holds every distinct neighbour, so
//    the ring fits), vertex ids local to the mesh; ring_len its length;
//  - area: the sum of a third of each face's area in the order
//    GetVertexNeighbourFaces lists the faces (the faces of each ring
//    neighbour in turn, a face at its first appearance,
//    vtkSurfaceBase.cxx:964-984).  A face's place is the ring index of its
//    first vertex in the ring (key, stored beside the face in the row), then
//    its id.  FP64 in the host's order of operations.
__global__ void k_sample_topology(int64_t off, int64_t n, const double* __restrict__ pos,
                                  const int32_t* __restrict__ tri, const int32_t* __restrict__ vfptr,
                                  const int32_t* __restrict__ vfcol, const int32_t* __restrict__ row_ptr,
                                  int32_t* __restrict__ ring, int32_t* __restrict__ ring_len, int32_t* __restrict__ key,
                                  double* __restrict__ area) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t v = int32_t(off + i);
  int32_t* rg = ring + row_ptr[v];
  const int32_t k0 = vfptr[v], k1 = vfptr[v + 1];
  int32_t cnt = 0, prev = -1;
  for (int32_t k = k0; k < k1; ++k) {
    const int32_t f = vfcol[k];
    if (f == prev) continue;
    prev = f;
    const int32_t* t = tri + 3 * int64_t(f);
    if (t[0] == t[1]) continue;
    for (int j = 0; j < 3; ++j) {
      const int32_t a = t[j], b = t[j == 2 ? 0 : j + 1];
      if (a == b) continue;
      const int32_t u = a == v ? b : (b == v ? a : -1);
      if (u < 0) continue;
      bool seen = false;
      for (int32_t q = 0; q < cnt; ++q)
        if (rg[q] == u) {
          seen = true;
          break;
        }
      if (!seen) rg[cnt++] = u;
    }
  }
  ring_len[i] = cnt;
  prev = -1;
  for (int32_t k = k0; k < k1; ++k) {
    const int32_t f = vfcol[k];
    const int32_t* t = tri + 3 * int64_t(f);
    int32_t kk = -1;
    if (f != prev && t[0] != t[1])
      for (int32_t r = 0; r < cnt; ++r)
        if (t[0] == rg[r] || t[1] == rg[r] || t[2] == rg[r]) {
          kk = r;
          break;
        }
    prev = f;
    key[k] = kk;
  }
  double sum = 0.0;
  for (int32_t r = 0; r < cnt; ++r)
    for (int32_t k = k0; k < k1; ++k) {
      if (key[k] != r) continue;
      const int32_t* t = tri + 3 * int64_t(vfcol[k]);
      sum += tri_area_f64(pos + 3 * int64_t(t[0]), pos + 3 * int64_t(t[1]), pos + 3 * int64_t(t[2])) / 3.0;
    }
  area[i] = sum;
  for (int32_t r = 0; r < cnt; ++r) rg[r] -= int32_t(off);
}

}  // namespace

hipError_t launch_sample_topology(hipStream_t s, int64_t off, int64_t n, const double* pos, const int32_t* tri,
                                  const int32_t* vfptr, const int32_t* vfcol, const int32_t* row_ptr, int32_t* ring,
                                  int32_t* ring_len, int32_t* key, double* area) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sample_topology, dim3(grid_for(n)), dim3(kBlock), 0, s, off, n, pos, tri, vfptr, vfcol,
                     row_ptr, ring, ring_len, key, area);
  return hipGetLastError();
}

hipError_t launch_curvature(hipStream_t s, int64_t off, int64_t n, const double* pos, const int32_t* tri,
                            const int32_t* row_ptr, const int32_t* col, const int32_t* vfptr, const int32_t* vfcol,
                            double* out) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_curvature, dim3(grid_for(n)), dim3(kBlock), 0, s, off, n, pos, tri, row_ptr, col, vfptr,
                     vfcol, out);
  return hipGetLastError();
}

hipError_t launch_item_build(hipStream_t s, const SweepArgs& a, const double* pos, const int32_t* tri,
                             const int32_t* vfptr, const int32_t* vfcol, const double* indicator, double* raw_w,
                             float4* xw, float* q) {
  if (a.n_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_item_build, dim3(grid_for(a.n_items)), dim3(kBlock), 0, s, a, pos, tri, vfptr, vfcol,
                     indicator, raw_w, xw, q);
  return hipGetLastError();
}

hipError_t launch_seg_reduce(hipStream_t s, int n_blocks, const int2* block_tab, const SweepArgs& a,
                             const double* raw_w, int what, double* partial) {
  if (n_blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_seg_reduce, dim3(n_blocks), dim3(kBlock), 0, s, block_tab, a, raw_w, what, partial);
  return hipGetLastError();
}

// stage 0: fold partial sums into per-mesh averages (into mesh_tmp) and clamp
// the weights; stage 1: fold into scales.
hipError_t launch_clamp_and_scales(hipStream_t s, const SweepArgs& a, int n_blocks, const int2* block_tab,
                                   const double* partial, int stage, const double* raw_w, DevMesh* meshes_rw) {
  // mesh_tmp lives after the partial array (caller sizes it)
  double* mesh_tmp = const_cast<double*>(partial) + 4 * size_t(n_blocks);
  hipLaunchKernelGGL(k_mesh_fold, dim3((a.n_meshes + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a, n_blocks,
                     block_tab, partial, stage, meshes_rw, mesh_tmp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || stage != 0) return e;
  hipLaunchKernelGGL(k_clamp, dim3(grid_for(a.n_items)), dim3(kBlock), 0, s, a, raw_w, mesh_tmp,
                     const_cast<float4*>(a.xw));
  return hipGetLastError();
}


namespace {

// 16 bits -> every third bit of 48
__device__ inline uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0xFFFFu;
  x = (x | (x << 32)) & 0x1F00000000FFFFull;
  x = (x | (x << 16)) & 0x1F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// key = mesh << 48 | Morton code of the position quantised to 16 bits per
// axis inside the mesh's bounding box (bbox: lo[3], scale[3] per mesh)
__global__ void k_morton_keys(SweepArgs a, const float* __restrict__ bbox, unsigned long long* keys, int32_t* vals) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n_items) return;
  const int m = a.item_mesh ? a.item_mesh[i] : 0;
  const float4 p = a.xw[i];
  const float* bb = bbox + 6 * m;
  const float c[3] = {p.x, p.y, p.z};
  uint64_t key = uint64_t(m) << 48;
  for (int k = 0; k < 3; ++k) {
    float u = (c[k] - bb[k]) * bb[3 + k];
    u = fminf(fmaxf(u, 0.0f), 65535.0f);
    key |= spread3(uint32_t(u)) << k;
  }
  keys[i] = key;
  vals[i] = int32_t(i);
}

// sorted position p -> tile slot (tiles of one mesh are contiguous)
__global__ void k_tile_assign(SweepArgs a, const int32_t* __restrict__ sorted, int32_t* tile_items,
                              int32_t* item_tile) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= a.n_items) return;
  const int32_t i = sorted[p];
  const int m = a.item_mesh ? a.item_mesh[i] : 0;
  const DevMesh& dm = a.meshes[m];
  const int64_t local = p - dm.item_off;  // meshes sort in order, so p lies in the mesh's range
  const int64_t t = dm.tile_off + local / kTileItems;
  tile_items[t * kTileItems + local % kTileItems] = i;
  item_tile[i] = int32_t(t);
}

__global__ void k_fill_tiles(int64_t n, int32_t* p) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = -1;
}

}  // namespace

hipError_t launch_tile_order(hipStream_t s, const SweepArgs& a, const float* bbox, int32_t* tile_items,
                             int32_t* item_tile, void** tmp, size_t* tmp_bytes) {
  const int64_t n = a.n_items;
  unsigned long long *k0 = nullptr, *k1 = nullptr;
  int32_t *v0 = nullptr, *v1 = nullptr;
  hipError_t e = hipSuccess;
  auto cleanup = [&]() {
    (void)hipFree(k0); (void)hipFree(k1); (void)hipFree(v0); (void)hipFree(v1);
  };
#define ACVD_TILE_TRY(x) do { e = (x); if (e != hipSuccess) { cleanup(); return e; } } while (0)
  ACVD_TILE_TRY(hipMalloc(&k0, n * sizeof(unsigned long long)));
  ACVD_TILE_TRY(hipMalloc(&k1, n * sizeof(unsigned long long)));
  ACVD_TILE_TRY(hipMalloc(&v0, n * sizeof(int32_t)));
  ACVD_TILE_TRY(hipMalloc(&v1, n * sizeof(int32_t)));
  const int64_t nt = int64_t(a.n_tiles) * kTileItems;
  hipLaunchKernelGGL(k_fill_tiles, dim3(grid_for(nt)), dim3(kBlock), 0, s, nt, tile_items);
  ACVD_TILE_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_morton_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, a, bbox, k0, v0);
  ACVD_TILE_TRY(hipGetLastError());
  size_t need = 0;
  ACVD_TILE_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, need, k0, k1, v0, v1, int(n), 0, 64, s));
  if (need > *tmp_bytes) {
    ACVD_TILE_TRY(hipStreamSynchronize(s));
    if (*tmp) (void)hipFree(*tmp);
    *tmp = nullptr;
    *tmp_bytes = 0;
    ACVD_TILE_TRY(hipMalloc(tmp, need));
    *tmp_bytes = need;
  }
  size_t have = *tmp_bytes;
  ACVD_TILE_TRY(hipcub::DeviceRadixSort::SortPairs(*tmp, have, k0, k1, v0, v1, int(n), 0, 64, s));
  hipLaunchKernelGGL(k_tile_assign, dim3(grid_for(n)), dim3(kBlock), 0, s, a, v1, tile_items, item_tile);
  ACVD_TILE_TRY(hipGetLastError());
  ACVD_TILE_TRY(hipStreamSynchronize(s));
  cleanup();
#undef ACVD_TILE_TRY
  return hipSuccess;
}

}  // namespace acvd
