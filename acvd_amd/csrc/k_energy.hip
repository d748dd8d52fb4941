// Fixed-assignment clustering energy and its adjoint.
//
// For a labelling l (frozen) of the vertices of one or more meshes:
//   isotropic (ACVD):  E = sum_i w_i |x_i - c_l(i)|^2,  c = S/W,
//                      S = sum w x, W = sum w over the cluster
//                      (vtkIsotropicMetricForClustering.h:114-121 unshifted);
//   QEM (ACVDQ):       E = sum_i w_i |x_i - q_l(i)|^2,
//                      q = c + A+_trunc (b - A c), with (A, b) from the cluster
//                      quadric Q_c = sum_{i in c} sum_{f ni i} Q_f(x),
//                      Q_f = [n; d][n; d]^T, n = (x2-x1) x (x3-x1), d = -n.x1
//                      (vtkQEMetricForClustering.h:151-167, 194-208,
//                      Common/vtkQuadricTools.cxx:56-177).
// Everything is computed in FP64 from FP32 positions and weights; the face
// quadrics are recomputed from the positions (so that they are differentiable),
// unlike the sweep, which uses FP32 vertex quadrics.
//
// Forward: cluster sums (W, S, Q) with block-level LDS aggregation and FP64
// atomics, cluster points, per-mesh energy.  Backward (adjoint): per cluster,
// dE/dq = -2 gE (S - W q) is pulled back through q(c, A, b) -- the truncated
// pseudo-inverse's derivative includes the coupling of kept and dropped
// eigen-directions -- into dE/dQ (9 entries) and dE/dc; then per vertex the
// direct terms, the centroid path and the face-quadric path (over the faces
// around the vertex, recomputed, no per-item partials stored) give dE/dx and
// dE/dw.

#include <hip/hip_runtime.h>

#include <mutex>

#include "acvd_energy.h"

namespace acvd {

namespace {

constexpr int kEBlock = 256;
constexpr int kAcc = 13;  // W, Sx, Sy, Sz, Q0..Q8
constexpr double kThreshold = 1e-3;

inline unsigned egrid(int64_t n) { return unsigned((n + kEBlock - 1) / kEBlock); }

__device__ inline int mesh_of(const EnergyArgs& a, int64_t i) { return a.item_mesh ? a.item_mesh[i] : 0; }

// global cluster of item i, or -1
__device__ inline int64_t cluster_of(const EnergyArgs& a, int64_t i, int m) {
  const EnergyMesh& em = a.meshes[m];
  const int32_t l = a.labels[i];
  if (l < 0 || l >= em.K) return -1;
  return em.clus_off + l;
}

// plane quadric entries 0..8 of face f ([n;d][n;d]^T upper part, without d^2)
__device__ inline void face_plane(const EnergyArgs& a, int64_t f, double p[4], double x1[3], double u[3],
                                  double v[3]) {
  const int32_t i0 = a.tri[3 * f], i1 = a.tri[3 * f + 1], i2 = a.tri[3 * f + 2];
  double x2[3], x3[3];
  for (int k = 0; k < 3; ++k) {
    x1[k] = double(a.xyz[3 * int64_t(i0) + k]);
    x2[k] = double(a.xyz[3 * int64_t(i1) + k]);
    x3[k] = double(a.xyz[3 * int64_t(i2) + k]);
    u[k] = x2[k] - x1[k];
    v[k] = x3[k] - x1[k];
  }
  p[0] = u[1] * v[2] - u[2] * v[1];
  p[1] = u[2] * v[0] - u[0] * v[2];
  p[2] = u[0] * v[1] - u[1] * v[0];
  p[3] = -(p[0] * x1[0] + p[1] * x1[1] + p[2] * x1[2]);
}

__device__ inline void plane_to_q9(const double p[4], double q[9]) {
  q[0] = p[0] * p[0];
  q[1] = p[0] * p[1];
  q[2] = p[0] * p[2];
  q[3] = p[0] * p[3];
  q[4] = p[1] * p[1];
  q[5] = p[1] * p[2];
  q[6] = p[1] * p[3];
  q[7] = p[2] * p[2];
  q[8] = p[2] * p[3];
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi to convergence.
// A = (a00, a01, a02, a11, a12, a22); V columns are eigenvectors (V[3*r + k]).
__device__ void eig_sym3(const double* A, double lam[3], double V[9]) {
  double a[3][3] = {{A[0], A[1], A[2]}, {A[1], A[3], A[4]}, {A[2], A[4], A[5]}};
  for (int k = 0; k < 9; ++k) V[k] = (k % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 16; ++sweep) {
    const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    const double dia = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
    if (!(off > dia * 1e-18)) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        const double apq = a[p][q];
        if (apq == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
        const double t = (fabs(theta) > 1e150) ? 0.5 / theta
                                                : (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
        // A <- J^T A J with J the (p,q) rotation
        for (int k = 0; k < 3; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = cs * akp - sn * akq;
          a[k][q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = cs * apk - sn * aqk;
          a[q][k] = sn * apk + cs * aqk;
        }
        a[p][q] = a[q][p] = 0.0;
        for (int r = 0; r < 3; ++r) {
          const double vp = V[3 * r + p], vq = V[3 * r + q];
          V[3 * r + p] = cs * vp - sn * vq;
          V[3 * r + q] = sn * vp + cs * vq;
        }
      }
  }
  lam[0] = a[0][0];
  lam[1] = a[1][1];
  lam[2] = a[2][2];
}

// Kept eigen-directions: by decreasing |lambda| (first index on ties) while
// |lambda| / max|lambda| > 1e-3 and at most `level` of them
// (vtkQuadricTools.cxx:119-154).
__device__ inline void kept_set(const double lam[3], int level, bool keep[3]) {
  double aw[3] = {fabs(lam[0]), fabs(lam[1]), fabs(lam[2])};
  const double mx = fmax(aw[0], fmax(aw[1], aw[2]));
  keep[0] = keep[1] = keep[2] = false;
  int left = level;
  for (int it = 0; it < 3; ++it) {
    int idx = -1;
    double best = -1.0;
    for (int k = 0; k < 3; ++k)
      if (!keep[k] && aw[k] > best && aw[k] >= 0.0) {
        // skip indices already decided as dropped
        best = aw[k];
        idx = k;
      }
    if (idx < 0) break;
    if (mx > 0.0 && best * (1.0 / mx) > kThreshold && left > 0) {
      keep[idx] = true;
      --left;
    } else {
      break;  // every remaining value is smaller
    }
  }
}

__device__ inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// ------------------------------------------------------------------ forward
__global__ __launch_bounds__(kEBlock) void ke_zero(int64_t n, double* p) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Cluster sums with a per-block LDS table keyed by cluster (256 slots for 256
// items: every insertion finds a slot), flushed with one FP64 atomic per
// (cluster, value) per block.
__global__ __launch_bounds__(kEBlock) void ke_accumulate(EnergyArgs a) {
  __shared__ int32_t keys[kEBlock];
  __shared__ double vals[kEBlock * kAcc];
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  keys[threadIdx.x] = -1;
  for (int k = 0; k < kAcc; ++k) vals[threadIdx.x * kAcc + k] = 0.0;
  __syncthreads();
  int64_t c = -1;
  int m = 0;
  if (i < a.n_items) {
    m = mesh_of(a, i);
    c = cluster_of(a, i, m);
  }
  if (c >= 0) {
    const double w = double(a.w[i]);
    double v[kAcc];
    v[0] = w;
    for (int k = 0; k < 3; ++k) v[1 + k] = w * double(a.xyz[3 * i + k]);
    for (int k = 0; k < 9; ++k) v[4 + k] = 0.0;
    if (a.meshes[m].metric == 1) {
      const EnergyMesh& em = a.meshes[m];
      for (int32_t k = a.vfptr[i]; k < a.vfptr[i + 1]; ++k) {
        double p[4], x1[3], u[3], vv[3], q9[9];
        face_plane(a, em.face_off + a.vfcol[k], p, x1, u, vv);
        plane_to_q9(p, q9);
        for (int e = 0; e < 9; ++e) v[4 + e] += q9[e];
      }
    }
    const int32_t key = int32_t(c);
    uint32_t h = (uint32_t(key) * 2654435761u) >> 24;
    int slot = -1;
    for (int probe = 0; probe < kEBlock; ++probe) {
      const int32_t old = atomicCAS(&keys[h], -1, key);
      if (old == -1 || old == key) {
        slot = int(h);
        break;
      }
      h = (h + 1) & (kEBlock - 1);
    }
    for (int k = 0; k < kAcc; ++k) atomicAdd(&vals[slot * kAcc + k], v[k]);
  }
  __syncthreads();
  const int32_t key = keys[threadIdx.x];
  if (key >= 0) {
    double* dst = a.acc + int64_t(key) * kAcc;
    for (int k = 0; k < kAcc; ++k) {
      const double x = vals[threadIdx.x * kAcc + k];
      if (x != 0.0) atomicAdd(&dst[k], x);
    }
  }
}

// Cluster points: centroid c and (QEM) the representative point q.
__global__ __launch_bounds__(kEBlock) void ke_points(EnergyArgs a) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= a.n_clusters) return;
  const int m = a.clus_mesh ? a.clus_mesh[c] : 0;
  const EnergyMesh& em = a.meshes[m];
  const double* s = a.acc + c * kAcc;
  double* out = a.pts + c * 6;  // c[3], q[3]
  if (!(s[0] > 0.0)) {
    for (int k = 0; k < 6; ++k) out[k] = 0.0;
    return;
  }
  double cc[3] = {s[1] / s[0], s[2] / s[0], s[3] / s[0]};
  double q[3] = {cc[0], cc[1], cc[2]};
  if (em.metric == 1) {
    const double* Q = s + 4;
    const double A[6] = {Q[0], Q[1], Q[2], Q[4], Q[5], Q[7]};
    double lam[3], V[9];
    bool keep[3];
    eig_sym3(A, lam, V);
    kept_set(lam, em.level, keep);
    const double r[3] = {-Q[3] - (Q[0] * cc[0] + Q[1] * cc[1] + Q[2] * cc[2]),
                         -Q[6] - (Q[1] * cc[0] + Q[4] * cc[1] + Q[5] * cc[2]),
                         -Q[8] - (Q[2] * cc[0] + Q[5] * cc[1] + Q[7] * cc[2])};
    for (int k = 0; k < 3; ++k) {
      if (!keep[k]) continue;
      const double vk[3] = {V[k], V[3 + k], V[6 + k]};
      const double f = dot3(vk, r) / lam[k];
      for (int e = 0; e < 3; ++e) q[e] += f * vk[e];
    }
  }
  for (int k = 0; k < 3; ++k) {
    out[k] = cc[k];
    out[3 + k] = q[k];
  }
}

// Per-mesh energy: block sum when the block lies in one mesh, else per item.
__global__ __launch_bounds__(kEBlock) void ke_energy(EnergyArgs a, double* energy) {
  __shared__ double red[kEBlock];
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double e = 0.0;
  int m = 0;
  if (i < a.n_items) {
    m = mesh_of(a, i);
    const int64_t c = cluster_of(a, i, m);
    if (c >= 0) {
      const double* q = a.pts + c * 6 + 3;
      double d2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double d = double(a.xyz[3 * i + k]) - q[k];
        d2 += d * d;
      }
      e = double(a.w[i]) * d2;
    }
  }
  const int64_t first = int64_t(blockIdx.x) * blockDim.x;
  const int64_t last = min(first + kEBlock, a.n_items) - 1;
  const bool one_mesh = mesh_of(a, first) == mesh_of(a, last);
  if (!one_mesh) {
    if (e != 0.0) atomicAdd(&energy[m], e);
    return;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = kEBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&energy[mesh_of(a, first)], red[0]);
}

// ----------------------------------------------------------------- backward
// Per cluster: adj[0..8] = dE/dQ (entries 0..8), adj[9..11] = (dE/dc) / W.
__global__ __launch_bounds__(kEBlock) void ke_cluster_adjoint(EnergyArgs a, const double* __restrict__ gE) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= a.n_clusters) return;
  const int m = a.clus_mesh ? a.clus_mesh[c] : 0;
  const EnergyMesh& em = a.meshes[m];
  double* adj = a.adj + c * 12;
  for (int k = 0; k < 12; ++k) adj[k] = 0.0;
  const double* s = a.acc + c * kAcc;
  if (em.metric != 1 || !(s[0] > 0.0)) return;  // isotropic: dE/dc = 0 at the centroid
  const double g = gE[m];
  const double W = s[0];
  const double* cc = a.pts + c * 6;
  const double* q = cc + 3;
  // dE/dq = -2 g (S - W q)
  const double gq[3] = {-2.0 * g * (s[1] - W * q[0]), -2.0 * g * (s[2] - W * q[1]), -2.0 * g * (s[3] - W * q[2])};
  const double* Q = s + 4;
  const double A[6] = {Q[0], Q[1], Q[2], Q[4], Q[5], Q[7]};
  double lam[3], V[9];
  bool keep[3];
  eig_sym3(A, lam, V);
  kept_set(lam, em.level, keep);
  const double r[3] = {-Q[3] - (Q[0] * cc[0] + Q[1] * cc[1] + Q[2] * cc[2]),
                       -Q[6] - (Q[1] * cc[0] + Q[4] * cc[1] + Q[5] * cc[2]),
                       -Q[8] - (Q[2] * cc[0] + Q[5] * cc[1] + Q[7] * cc[2])};
  double vg[3], vr[3];
  for (int k = 0; k < 3; ++k) {
    const double vk[3] = {V[k], V[3 + k], V[6 + k]};
    vg[k] = dot3(vk, gq);
    vr[k] = dot3(vk, r);
  }
  // y = A+ r = q - c; h = A+ gq; P gq = gq - sum_kept v_k (v_k . gq)
  double h[3] = {0, 0, 0}, pg[3] = {gq[0], gq[1], gq[2]};
  for (int k = 0; k < 3; ++k) {
    if (!keep[k]) continue;
    for (int e = 0; e < 3; ++e) {
      h[e] += V[3 * e + k] * vg[k] / lam[k];
      pg[e] -= V[3 * e + k] * vg[k];
    }
  }
  // G = -h (y + c)^T + sum_{k kept, j dropped} coef_kj v_j v_k^T
  const double yc[3] = {q[0], q[1], q[2]};  // y + c = q
  double G[3][3];
  for (int r0 = 0; r0 < 3; ++r0)
    for (int c0 = 0; c0 < 3; ++c0) G[r0][c0] = -h[r0] * yc[c0];
  for (int k = 0; k < 3; ++k) {
    if (!keep[k]) continue;
    for (int j = 0; j < 3; ++j) {
      if (keep[j]) continue;
      const double gap = lam[k] - lam[j];
      if (gap == 0.0) continue;
      const double coef = (vg[j] * vr[k] + vg[k] * vr[j]) / (lam[k] * gap);
      for (int r0 = 0; r0 < 3; ++r0)
        for (int c0 = 0; c0 < 3; ++c0) G[r0][c0] += coef * V[3 * r0 + j] * V[3 * c0 + k];
    }
  }
  // symmetric part onto the quadric entries; b = -(Q3, Q6, Q8) gets A+ gq
  adj[0] = G[0][0];
  adj[1] = G[0][1] + G[1][0];
  adj[2] = G[0][2] + G[2][0];
  adj[4] = G[1][1];
  adj[5] = G[1][2] + G[2][1];
  adj[7] = G[2][2];
  adj[3] = -h[0];
  adj[6] = -h[1];
  adj[8] = -h[2];
  for (int e = 0; e < 3; ++e) adj[9 + e] = pg[e] / W;
}

// Per vertex: dE/dx (3) and dE/dw.
__global__ __launch_bounds__(kEBlock) void ke_item_grad(EnergyArgs a, const double* __restrict__ gE, float* dxyz,
                                                        float* dw) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n_items) return;
  const int m = mesh_of(a, i);
  const EnergyMesh& em = a.meshes[m];
  const double g = gE[m];
  double gx[3] = {0, 0, 0}, gw = 0.0;
  const double x[3] = {double(a.xyz[3 * i]), double(a.xyz[3 * i + 1]), double(a.xyz[3 * i + 2])};
  const int64_t c = cluster_of(a, i, m);
  if (c >= 0) {
    const double w = double(a.w[i]);
    const double* cc = a.pts + c * 6;
    const double* q = cc + 3;
    const double* adj = a.adj + c * 12;
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double d = x[k] - q[k];
      gx[k] = 2.0 * g * w * d + w * adj[9 + k];
      d2 += d * d;
    }
    gw = g * d2 + adj[9] * (x[0] - cc[0]) + adj[10] * (x[1] - cc[1]) + adj[11] * (x[2] - cc[2]);
  }
  if (em.metric == 1) {
    // faces around i: their quadrics feed the clusters of their three vertices
    for (int32_t k = a.vfptr[i]; k < a.vfptr[i + 1]; ++k) {
      const int64_t f = em.face_off + a.vfcol[k];
      double gq9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      int role = 0;
      for (int j = 0; j < 3; ++j) {
        const int32_t vj = a.tri[3 * f + j];
        if (vj == i) role = j;
        const int64_t cj = cluster_of(a, vj, m);
        if (cj < 0) continue;
        const double* aj = a.adj + cj * 12;
        for (int e = 0; e < 9; ++e) gq9[e] += aj[e];
      }
      double p[4], x1[3], u[3], v[3];
      face_plane(a, f, p, x1, u, v);
      const double dp[4] = {2.0 * gq9[0] * p[0] + gq9[1] * p[1] + gq9[2] * p[2] + gq9[3] * p[3],
                            gq9[1] * p[0] + 2.0 * gq9[4] * p[1] + gq9[5] * p[2] + gq9[6] * p[3],
                            gq9[2] * p[0] + gq9[5] * p[1] + 2.0 * gq9[7] * p[2] + gq9[8] * p[3],
                            gq9[3] * p[0] + gq9[6] * p[1] + gq9[8] * p[2]};
      // d = -n.x1: n receives -dp3 x1, x1 receives -dp3 n
      const double gn[3] = {dp[0] - dp[3] * x1[0], dp[1] - dp[3] * x1[1], dp[2] - dp[3] * x1[2]};
      // n = u x v: gu = v x gn, gv = gn x u
      const double gu[3] = {v[1] * gn[2] - v[2] * gn[1], v[2] * gn[0] - v[0] * gn[2], v[0] * gn[1] - v[1] * gn[0]};
      const double gv[3] = {gn[1] * u[2] - gn[2] * u[1], gn[2] * u[0] - gn[0] * u[2], gn[0] * u[1] - gn[1] * u[0]};
      for (int e = 0; e < 3; ++e) {
        if (role == 0) gx[e] += -gu[e] - gv[e] - dp[3] * p[e];
        else if (role == 1) gx[e] += gu[e];
        else gx[e] += gv[e];
      }
    }
  }
  for (int k = 0; k < 3; ++k) dxyz[3 * i + k] = float(gx[k]);
  dw[i] = float(gw);
}

}  // namespace

hipError_t energy_forward(hipStream_t s, const EnergyArgs& a, double* energy, int n_meshes) {
  const int64_t nacc = a.n_clusters * kAcc;
  hipLaunchKernelGGL(ke_zero, dim3(egrid(nacc)), dim3(kEBlock), 0, s, nacc, a.acc);
  hipLaunchKernelGGL(ke_zero, dim3(egrid(n_meshes)), dim3(kEBlock), 0, s, int64_t(n_meshes), energy);
  if (a.n_items > 0) hipLaunchKernelGGL(ke_accumulate, dim3(egrid(a.n_items)), dim3(kEBlock), 0, s, a);
  if (a.n_clusters > 0) hipLaunchKernelGGL(ke_points, dim3(egrid(a.n_clusters)), dim3(kEBlock), 0, s, a);
  if (a.n_items > 0) hipLaunchKernelGGL(ke_energy, dim3(egrid(a.n_items)), dim3(kEBlock), 0, s, a, energy);
  return hipGetLastError();
}

// ------------------------------------------------- forward without allocation
namespace {

// Scratch of energy_forward_cvt_noalloc: the mesh descriptor, then the
// per-cluster sums (kAcc doubles) and points (6 doubles) of at most
// kNoAllocMaxClusters clusters; one copy per device, held until the process
// ends.  Host code never passes `scratch` itself to a kernel (in host code the
// name is only a stand-in): it uses the device address found once per device
// by hipGetSymbolAddress.
constexpr size_t kNoAllocMeshBytes = 256;
constexpr size_t kNoAllocBytes =
    kNoAllocMeshBytes + size_t(kNoAllocMaxClusters) * size_t(kAcc + 6) * sizeof(double);
static_assert(sizeof(EnergyMesh) <= kNoAllocMeshBytes, "mesh descriptor slot");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;

// Device address of `scratch` on the current device (looked up once per device).
hipError_t noalloc_scratch(unsigned char** out) {
  static std::mutex mu;
  static unsigned char* ptr[kNoAllocMaxDevices] = {};
  int dev = 0;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess) return err;
  if (dev < 0 || dev >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess) return err;
    ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = ptr[dev];
  return hipSuccess;
}

__global__ void ke_set_mesh(EnergyMesh m, EnergyMesh* dst) { *dst = m; }

}  // namespace

hipError_t energy_forward_cvt_noalloc(hipStream_t s, int64_t n, int32_t K, const float* xyz, const float* w,
                                      const int32_t* labels, double* energy) {
  unsigned char* base = nullptr;
  hipError_t err = noalloc_scratch(&base);
  if (err != hipSuccess) return err;
  EnergyMesh* mesh = reinterpret_cast<EnergyMesh*>(base);
  double* acc = reinterpret_cast<double*>(base + kNoAllocMeshBytes);
  EnergyArgs a{};
  a.n_items = n;
  a.n_clusters = K;
  a.meshes = mesh;
  a.xyz = xyz;
  a.w = w;
  a.labels = labels;
  a.acc = acc;
  a.pts = acc + size_t(kNoAllocMaxClusters) * kAcc;
  // the descriptor is written by a kernel (no host copy: the call may be captured)
  hipLaunchKernelGGL(ke_set_mesh, dim3(1), dim3(1), 0, s, EnergyMesh{0, n, 0, 0, K, 0, 0, 0}, mesh);
  return energy_forward(s, a, energy, 1);
}

hipError_t energy_backward(hipStream_t s, const EnergyArgs& a, const double* grad_energy, float* dxyz, float* dw) {
  if (a.n_clusters > 0)
    hipLaunchKernelGGL(ke_cluster_adjoint, dim3(egrid(a.n_clusters)), dim3(kEBlock), 0, s, a, grad_energy);
  if (a.n_items > 0)
    hipLaunchKernelGGL(ke_item_grad, dim3(egrid(a.n_items)), dim3(kEBlock), 0, s, a, grad_energy, dxyz, dw);
  return hipGetLastError();
}

}  // namespace acvd
