// Per-item / per-cluster arithmetic of the boundary sweep, shared verbatim by
// the HIP kernels and the host-side replay so that both produce bit-identical
// labels.  Rules that keep host and device in lock-step:
//   - every translation unit that includes this header is compiled with
//     -ffp-contract=off and without fast-math (no silent FMA formation);
//   - only +, -, *, / and sqrt are used (IEEE correctly rounded on both sides),
//     no transcendental functions;
//   - conversions double -> int64 round to nearest even (llrint on the host,
//     __double2ll_rn on the device), int64 -> double is exact-rounding on both;
//   - integer priorities and integer tie-breaks only.
//
// Energy model (both metrics, per cluster c with weighted sum S, weight W,
// centroid c = S/W and quadric-optimal point q):
//   isotropic (vtkIsotropicMetricForClustering.h:114-121):  E_c = -|S|^2/W
//   QEM       (vtkQEMetricForClustering.h:194-208):        E_c = W|q|^2 - 2 q.S
//                                                           = -|S|^2/W + W|q-c|^2
// Moving item i (position x, weight w) from cluster a to cluster b changes the
// isotropic part by the Hartigan form
//   dE_iso = w [ W_b/(W_b+w) |x-c_b|^2 - W_a/(W_a-w) |x-c_a|^2 ]
// which equals the reference's Try2 - Try1 (vtkUniformClustering.h:925-940)
// without its cancellation, and the QEM part by
//   dT = T_a' + T_b' - T_a - T_b,  T = W |q - c|^2 = W |Delta|^2,
// with Delta the rank-truncated pseudo-inverse step of
// vtkQuadricTools::ComputeDisplacement (Common/vtkQuadricTools.cxx:83-163).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ACVD_HD __host__ __device__ inline
#define ACVD_UNROLL _Pragma("unroll")
#else
#include <cmath>
#define ACVD_HD inline
#define ACVD_UNROLL _Pragma("GCC unroll 16")
#endif

namespace acvd {

// Null label used inside the library (the reference uses K, the cluster count).
constexpr int32_t kNull = -1;

// Decision threshold: a move is taken only if dE < -kAcceptEps * M where M is
// the sum of magnitudes of the terms of dE.  kAcceptEps is well above the FP32
// rounding of the terms, so an accepted move lowers the exact energy and no
// sequence of moves can cycle.
constexpr float kAcceptEps = 1.0f / 262144.0f;  // 2^-18

// Singular-value truncation threshold (vtkQuadricTools.h:36).
constexpr double kSvThreshold = 1e-3;

ACVD_HD double dsqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_sqrt(x);
#else
  return std::sqrt(x);
#endif
}

// Triangle area in FP64 from the squared edge lengths a, b, c:
// sqrt(|4ac - (a - b + c)^2|) / 4 (Common/vtkSurface.cxx:1280-1283); the
// same operations on the host and the device.
ACVD_HD double tri_area_f64(const double* p1, const double* p2, const double* p3) {
  const double a = (p1[0] - p2[0]) * (p1[0] - p2[0]) + (p1[1] - p2[1]) * (p1[1] - p2[1]) +
                   (p1[2] - p2[2]) * (p1[2] - p2[2]);
  const double b = (p2[0] - p3[0]) * (p2[0] - p3[0]) + (p2[1] - p3[1]) * (p2[1] - p3[1]) +
                   (p2[2] - p3[2]) * (p2[2] - p3[2]);
  const double c = (p3[0] - p1[0]) * (p3[0] - p1[0]) + (p3[1] - p1[1]) * (p3[1] - p1[1]) +
                   (p3[2] - p1[2]) * (p3[2] - p1[2]);
  const double d = 4.0 * a * c - (a - b + c) * (a - b + c);
  return 0.25 * dsqrt(__builtin_fabs(d));
}

ACVD_HD int64_t to_fixed(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __double2ll_rn(v);
#else
  return std::llrint(v);
#endif
}

// Power-of-two fixed-point scales of one mesh: value * s -> int64.
struct Scales {
  double sw, ss, sq;     // weight, weighted position, quadric
  double isw, iss, isq;  // inverses (exact, powers of two)
};

// Fixed-point images of one item (weights / weighted positions / quadric).
ACVD_HD int64_t fix_w(float w, const Scales& s) { return to_fixed(double(w) * s.sw); }
ACVD_HD int64_t fix_s(float w, float x, const Scales& s) { return to_fixed(double(w) * double(x) * s.ss); }
ACVD_HD int64_t fix_q(float q, const Scales& s) { return to_fixed(double(q) * s.sq); }

// ---------------------------------------------------------------------------
// Rank-truncated representative point (FP64).
// Q: 9 (or more) quadric coefficients in the compact layout
//   [0]=A00 [1]=A01 [2]=A02 [3]=A0d [4]=A11 [5]=A12 [6]=A1d [7]=A22 [8]=A2d
// c: start point (the centroid).  Writes Delta = A+_trunc (b - A c) with
// b = -(Q3,Q6,Q8); returns the rank deficiency (0..3).
// Eigen-decomposition: cyclic Jacobi, at most 8 sweeps, stopping as soon as
// the off-diagonal part is negligible (same decisions on host and device since
// the inputs are bit-identical).
// V0 (optional, 3x3 row-major, orthonormal columns): a warm start, e.g. the
// eigenvectors of a cluster whose quadric this one perturbs by one item; the
// solve then runs on V0^T A V0, nearly diagonal, in fewer sweeps.  Vout
// (optional) receives the eigenvectors (columns, row-major).
ACVD_HD int rep_point_delta(const double* Q, const double* c, int level, double* D, const double* V0 = nullptr,
                            double* Vout = nullptr) {
  double a00 = Q[0], a01 = Q[1], a02 = Q[2], a11 = Q[4], a12 = Q[5], a22 = Q[7];
  double r0 = -Q[3] - (Q[0] * c[0] + Q[1] * c[1] + Q[2] * c[2]);
  double r1 = -Q[6] - (Q[1] * c[0] + Q[4] * c[1] + Q[5] * c[2]);
  double r2 = -Q[8] - (Q[2] * c[0] + Q[5] * c[1] + Q[7] * c[2]);
  if (V0) {
    // A' = V0^T A V0, r' = V0^T r
    const double A[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
    double M[3][3];
    for (int i = 0; i < 3; ++i)
      for (int q = 0; q < 3; ++q) M[i][q] = A[i][0] * V0[q] + A[i][1] * V0[3 + q] + A[i][2] * V0[6 + q];
    double B[3][3];
    for (int p = 0; p < 3; ++p)
      for (int q = p; q < 3; ++q) B[p][q] = V0[p] * M[0][q] + V0[3 + p] * M[1][q] + V0[6 + p] * M[2][q];
    a00 = B[0][0];
    a01 = B[0][1];
    a02 = B[0][2];
    a11 = B[1][1];
    a12 = B[1][2];
    a22 = B[2][2];
    const double s0 = V0[0] * r0 + V0[3] * r1 + V0[6] * r2;
    const double s1 = V0[1] * r0 + V0[4] * r1 + V0[7] * r2;
    const double s2 = V0[2] * r0 + V0[5] * r1 + V0[8] * r2;
    r0 = s0;
    r1 = s1;
    r2 = s2;
  }
  // V = identity; columns are eigenvectors
  double v00 = 1, v01 = 0, v02 = 0, v10 = 0, v11 = 1, v12 = 0, v20 = 0, v21 = 0, v22 = 1;
  for (int sweep = 0; sweep < 8; ++sweep) {
    double off = (a01 < 0 ? -a01 : a01) + (a02 < 0 ? -a02 : a02) + (a12 < 0 ? -a12 : a12);
    double dia = (a00 < 0 ? -a00 : a00) + (a11 < 0 ? -a11 : a11) + (a22 < 0 ? -a22 : a22);
    if (!(off > dia * 1e-17)) break;
    // three rotations (p,q) = (0,1), (0,2), (1,2)
#define ACVD_JACOBI_ROT(app, aqq, apq, akp, akq, vp0, vq0, vp1, vq1, vp2, vq2)              \
    if (apq != 0.0) {                                                                     \
      double theta = (aqq - app) / (2.0 * apq);                                           \
      double at = theta < 0 ? -theta : theta;                                             \
      double t = (at > 1e150) ? 0.5 / theta                                               \
                              : (theta >= 0 ? 1.0 : -1.0) / (at + dsqrt(theta * theta + 1.0)); \
      double cs = 1.0 / dsqrt(t * t + 1.0);                                               \
      double sn = t * cs;                                                                 \
      double tau = t * apq;                                                               \
      app = app - tau;                                                                    \
      aqq = aqq + tau;                                                                    \
      apq = 0.0;                                                                          \
      double kp = akp, kq = akq;                                                          \
      akp = cs * kp - sn * kq;                                                            \
      akq = sn * kp + cs * kq;                                                            \
      double x0 = vp0, y0 = vq0;                                                          \
      vp0 = cs * x0 - sn * y0;                                                            \
      vq0 = sn * x0 + cs * y0;                                                            \
      double x1 = vp1, y1 = vq1;                                                          \
      vp1 = cs * x1 - sn * y1;                                                            \
      vq1 = sn * x1 + cs * y1;                                                            \
      double x2 = vp2, y2 = vq2;                                                          \
      vp2 = cs * x2 - sn * y2;                                                            \
      vq2 = sn * x2 + cs * y2;                                                            \
    }
    // (0,1): third index 2 -> couples a02 (k=2,p=0) and a12 (k=2,q=1)
    ACVD_JACOBI_ROT(a00, a11, a01, a02, a12, v00, v01, v10, v11, v20, v21)
    // (0,2): third index 1 -> a01 (p=0) and a12 (q=2)
    ACVD_JACOBI_ROT(a00, a22, a02, a01, a12, v00, v02, v10, v12, v20, v22)
    // (1,2): third index 0 -> a01 (p=1) and a02 (q=2)
    ACVD_JACOBI_ROT(a11, a22, a12, a01, a02, v01, v02, v11, v12, v21, v22)
#undef ACVD_JACOBI_ROT
  }
  double lam[3] = {a00, a11, a22};
  double aw[3] = {lam[0] < 0 ? -lam[0] : lam[0], lam[1] < 0 ? -lam[1] : lam[1], lam[2] < 0 ? -lam[2] : lam[2]};
  double maxw = aw[0];
  if (aw[1] > maxw) maxw = aw[1];
  if (aw[2] > maxw) maxw = aw[2];
  // visit eigenvalues by decreasing |lambda|, first index on ties
  int keep0 = 0, keep1 = 0, keep2 = 0;
  int rank_def = 0;
  int left = level;
  // |lambda| * (1 / max|lambda|) > threshold, the reference's form
  // (Common/vtkQuadricTools.cxx:117, 134), which can differ from
  // |lambda| / max|lambda| in the last bit at the threshold
  const double inv_maxw = maxw > 0.0 ? 1.0 / maxw : 0.0;
  double rem0 = aw[0], rem1 = aw[1], rem2 = aw[2];
  for (int it = 0; it < 3; ++it) {
    int idx = 0;
    double best = rem0;
    if (best < rem1) { best = rem1; idx = 1; }
    if (best < rem2) { best = rem2; idx = 2; }
    bool ok = (maxw > 0.0) && (best * inv_maxw > kSvThreshold) && (left > 0);
    if (ok) {
      if (idx == 0) keep0 = 1;
      else if (idx == 1) keep1 = 1;
      else keep2 = 1;
    } else {
      rank_def++;
    }
    if (idx == 0) rem0 = -2.0;
    else if (idx == 1) rem1 = -2.0;
    else rem2 = -2.0;
    left--;
  }
  double d0 = 0, d1 = 0, d2 = 0;
  if (keep0) {
    double p = (v00 * r0 + v10 * r1 + v20 * r2) / lam[0];
    d0 += v00 * p; d1 += v10 * p; d2 += v20 * p;
  }
  if (keep1) {
    double p = (v01 * r0 + v11 * r1 + v21 * r2) / lam[1];
    d0 += v01 * p; d1 += v11 * p; d2 += v21 * p;
  }
  if (keep2) {
    double p = (v02 * r0 + v12 * r1 + v22 * r2) / lam[2];
    d0 += v02 * p; d1 += v12 * p; d2 += v22 * p;
  }
  if (V0) {
    D[0] = V0[0] * d0 + V0[1] * d1 + V0[2] * d2;
    D[1] = V0[3] * d0 + V0[4] * d1 + V0[5] * d2;
    D[2] = V0[6] * d0 + V0[7] * d1 + V0[8] * d2;
  } else {
    D[0] = d0;
    D[1] = d1;
    D[2] = d2;
  }
  if (Vout) {
    const double W[9] = {v00, v01, v02, v10, v11, v12, v20, v21, v22};
    if (V0) {
      for (int i = 0; i < 3; ++i)
        for (int q = 0; q < 3; ++q) Vout[3 * i + q] = V0[3 * i] * W[q] + V0[3 * i + 1] * W[3 + q] + V0[3 * i + 2] * W[6 + q];
    } else {
      for (int k = 0; k < 9; ++k) Vout[k] = W[k];
    }
  }
  return rank_def;
}

// ---------------------------------------------------------------------------
// Cluster summaries from fixed-point accumulators.
//   ws[0] = W, ws[1..3] = S  (int64 fixed)
//   q[0..9]                  (int64 fixed, QEM only)
struct ClusterGeom {
  double cx, cy, cz, W;
};

ACVD_HD ClusterGeom geom_from_fixed(int64_t W, int64_t Sx, int64_t Sy, int64_t Sz, const Scales& s) {
  ClusterGeom g;
  g.W = double(W) * s.isw;
  double sx = double(Sx) * s.iss, sy = double(Sy) * s.iss, sz = double(Sz) * s.iss;
  if (g.W > 0.0) {
    g.cx = sx / g.W;
    g.cy = sy / g.W;
    g.cz = sz / g.W;
  } else {
    g.cx = g.cy = g.cz = 0.0;
  }
  return g;
}

// T = W |q - c|^2 of a cluster given its fixed accumulators (QEM metric with
// active constraints; 0 otherwise).  Optionally returns q.  Vout (when given)
// always receives orthonormal vectors: the identity when there is no solve.
ACVD_HD double qem_T_from_fixed(const int64_t* q10, const ClusterGeom& g, const Scales& s, int level,
                                double* q_out, const double* V0 = nullptr, double* Vout = nullptr) {
  if (!(g.W > 0.0) || level <= 0) {
    if (q_out) { q_out[0] = g.cx; q_out[1] = g.cy; q_out[2] = g.cz; }
    if (Vout)
      for (int k = 0; k < 9; ++k) Vout[k] = (k % 4 == 0) ? 1.0 : 0.0;
    return 0.0;
  }
  double Q[9];
  for (int k = 0; k < 9; ++k) Q[k] = double(q10[k]) * s.isq;
  double c[3] = {g.cx, g.cy, g.cz};
  double D[3];
  rep_point_delta(Q, c, level, D, V0, Vout);
  if (q_out) { q_out[0] = c[0] + D[0]; q_out[1] = c[1] + D[1]; q_out[2] = c[2] + D[2]; }
  return g.W * (D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
}

// ---------------------------------------------------------------------------
// Hartigan terms (FP32 with the position difference taken in FP64 first, so
// that |x - c| keeps FP32 relative precision for clusters much smaller than
// the coordinates).
//   leave term:  W_a/(W_a-w) |x-c_a|^2   (valid only for W_a > w)
//   join term:   W_b/(W_b+w) |x-c_b|^2
ACVD_HD float dist2_f(float x, float y, float z, const ClusterGeom& g) {
  float dx = float(double(x) - g.cx);
  float dy = float(double(y) - g.cy);
  float dz = float(double(z) - g.cz);
  return dx * dx + dy * dy + dz * dz;
}

ACVD_HD float leave_term(float x, float y, float z, float w, const ClusterGeom& a) {
  float f = float(a.W / (a.W - double(w)));
  return f * dist2_f(x, y, z, a);
}

ACVD_HD float join_term(float x, float y, float z, float w, const ClusterGeom& b) {
  float f = float(b.W / (b.W + double(w)));
  return f * dist2_f(x, y, z, b);
}

// True when the leave term is usable (the item does not carry the whole
// cluster weight up to FP64 resolution).
ACVD_HD bool can_leave(float w, const ClusterGeom& a) { return a.W - double(w) > a.W * 9.5367431640625e-7; }

// ---------------------------------------------------------------------------
// Seeded priorities for the conflict-free accept step.
ACVD_HD uint32_t prio_hash(uint32_t seed, uint32_t pass, uint32_t item) {
  uint32_t h = item * 0x9E3779B1u ^ (pass * 0x85EBCA77u + seed * 0xC2B2AE3Du);
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

ACVD_HD uint64_t prio_key(uint32_t seed, uint32_t pass, uint32_t item) {
  return (uint64_t(prio_hash(seed, pass, item)) << 32) | uint64_t(item);
}

// Accept step.  A pass runs kMatchRounds rounds over its proposals.  Round 0
// takes the proposals of the pass's evaluation (state frozen at the pass
// start); round r > 0 re-evaluates the proposals not applied yet against the
// state left by round r-1.  In every round each live proposal registers its
// key at both of its clusters (atomicMin) and a proposal holding the minimum
// at both is applied.  The moves of one round touch disjoint clusters, so
// every applied dE is exact with respect to the state at the start of its
// round and the energy decreases monotonically.  A round that applies no move
// ends the pass: nothing is stale after it and the item hashes of the keys do
// not depend on the round, so every later round would register the same
// minima and apply nothing either.  On a graded QEM torus (131,072 vertices,
// K = 1,560) 12 rounds need 170 passes where 6 need 240, with 42 % fewer
// proposals, at a lower final energy.
constexpr int kMatchRounds = 12;

// Accept key:
//   [63:40] 0xFFFFFF - (loop * kMatchRounds + round): later rounds give
//           smaller keys, so a slot never needs resetting
//   [39:27] seeded hash of (loop, item)   random priority
//   [26:0]  mesh-local item index         unique tie-break
ACVD_HD uint64_t accept_key(uint32_t seed, int32_t loop, int round, uint32_t local_item) {
  const uint32_t pr = uint32_t(loop) * uint32_t(kMatchRounds) + uint32_t(round);
  uint64_t tag = uint64_t(0xFFFFFFu - (pr & 0xFFFFFFu));
  uint64_t h = uint64_t(prio_hash(seed, uint32_t(loop), local_item) & 0x1FFFu);
  return (tag << 40) | (h << 27) | uint64_t(local_item & 0x7FFFFFFu);
}
constexpr int64_t kMaxItemsPerMesh = int64_t(1) << 27;
constexpr int32_t kMaxLoops = int32_t(0xFFFFFF / kMatchRounds) - 2;

// Isotropic acceptance: dE = w (tb - ta) < -eps * w (ta + tb).
ACVD_HD bool iso_accept(float ta, float tb) { return (tb - ta) < -kAcceptEps * (ta + tb); }

// QEM move energy and its magnitude (for the acceptance threshold).
ACVD_HD double qem_delta(float w, float ta, float tb, double Ta_new, double Ta, double Tb_new, double Tb) {
  return double(w * (tb - ta)) + ((Ta_new - Ta) + (Tb_new - Tb));
}
ACVD_HD double qem_mag(float w, float ta, float tb, double Ta_new, double Ta, double Tb_new, double Tb) {
  return double(w * (ta + tb)) + (Ta_new + Ta + Tb_new + Tb);
}
ACVD_HD bool qem_accept(double delta, double mag) { return delta < -double(kAcceptEps) * mag; }

// Fixed clusters (SetFixedClusters with the QEM metric's AnchorItem,
// Common/vtkUniformClustering.h:134, DiscreteRemeshing/vtkQEMetricForClustering.h:
// 196-199, 259, 270-274): the point of an anchored cluster is its anchor
// item's position in every phase, constrained or not, so its energy is
// -|S|^2/W + T with T = W |a - c|^2 and no solve.  The anchor item itself
// never leaves (the reference prices that at 1e100).
// Anchor position of a cluster, passed by value (valid = false: the cluster
// has none) so that no private array holds it.
struct AnchorPos {
  double x, y, z;
  bool valid;
};

ACVD_HD double anchored_T(const ClusterGeom& g, const AnchorPos& a) {
  if (!(g.W > 0.0)) return 0.0;
  const double dx = a.x - g.cx, dy = a.y - g.cy, dz = a.z - g.cz;
  return g.W * (dx * dx + dy * dy + dz * dz);
}

// Energy of one cluster as the energy log sums it (GetClusterEnergy:
// -|S|^2/W isotropic, vtkIsotropicMetricForClustering.h:114-121; W|p|^2 - 2 p.S
// QEM, vtkQEMetricForClustering.h:194-208, with p the anchor, the solved point
// when constrained, else the centroid), written -W|c|^2 + T with
// T = W|p - c|^2 (0 for isotropic clusters).  An empty cluster counts 0.
ACVD_HD double log_cluster_energy(const ClusterGeom& g, double T) {
  if (!(g.W > 0.0)) return 0.0;
  return T - g.W * (g.cx * g.cx + g.cy * g.cy + g.cz * g.cz);
}

// Per-mesh sum of the energy log in a fixed order: kLogThreads partial sums
// (partial t adds clusters t, t + kLogThreads, ... in turn, from 0), then a
// halving tree (partial t += partial t + h for h = kLogThreads / 2, ..., 1).
// The device reduction and the host replay take exactly these steps.
constexpr int kLogThreads = 256;

// T of a trial cluster = accumulators (ws: W,Sx,Sy,Sz; q10) + sign * item:
// with the anchor position an when the cluster is anchored, else (solve)
// the rank-truncated solve warm-started from the cluster's own eigenvectors
// V0 when given, else 0 (the cluster's point is its centroid).
ACVD_HD double qem_T_trial(const int64_t* ws, const int64_t* q10, int64_t iw, const int64_t* is,
                           const int64_t* iq, int sign, const Scales& s, int level, const double* V0 = nullptr,
                           AnchorPos an = AnchorPos{0.0, 0.0, 0.0, false}, bool solve = true) {
  int64_t W = ws[0] + sign * iw;
  int64_t Sx = ws[1] + sign * is[0], Sy = ws[2] + sign * is[1], Sz = ws[3] + sign * is[2];
  ClusterGeom g = geom_from_fixed(W, Sx, Sy, Sz, s);
  if (an.valid) return anchored_T(g, an);
  if (!solve) return 0.0;
  int64_t Q[10];
  for (int k = 0; k < 10; ++k) Q[k] = q10[k] + sign * iq[k];
  return qem_T_from_fixed(Q, g, s, level, nullptr, V0);
}

// Ring check of the vertex path (ConnexityConstraintProblemLocal,
// DiscreteRemeshing/vtkVerticesProcessing.h:168-237): would removing item i
// from cluster a split the ring neighbours of i that belong to a, when those
// neighbours are connected only through mesh edges between themselves?
// rp/col is the CSR adjacency with each row sorted by neighbour id.  Rings
// with more than 64 neighbours are refused (treated as a problem).
ACVD_HD bool ring_problem_rows(const int32_t* rp, const int32_t* col, const int32_t* lab, int32_t i, int32_t a);

// Rings of at most kRingCache neighbours whose neighbour rows are also that
// short (the usual case) are checked with all row entries loaded at once into
// registers and an all-pairs intersection (the rows hold distinct ids, so it
// finds exactly what the merge-walk of sorted rows finds), which replaces the
// chain of dependent loads of the walk; longer rows take the walk.  The same
// bound selects the register-resident QEM rule (valence <= 8 covers the
// usual meshes and keeps the evaluate kernels at 3 waves per SIMD).
constexpr int kRingCache = 8;
ACVD_HD bool ring_problem(const int32_t* rp, const int32_t* col, const int32_t* lab, int32_t i, int32_t a) {
  const int32_t r0 = rp[i], r1 = rp[i + 1];
  const int32_t deg = r1 - r0;
  if (deg > kRingCache) return ring_problem_rows(rp, col, lab, i, a);
  int32_t ci[kRingCache];
ACVD_UNROLL
  for (int p = 0; p < kRingCache; ++p) ci[p] = p < deg ? col[r0 + p] : -1;
  uint64_t M = 0;
ACVD_UNROLL
  for (int p = 0; p < kRingCache; ++p)
    if (p < deg && lab[ci[p]] == a) M |= uint64_t(1) << p;
  if (M == 0) return false;
  uint64_t reach = M & (~M + 1);
  uint64_t frontier = reach;
  while (frontier) {
    const int t = __builtin_ctzll(frontier);
    frontier &= frontier - 1;
    const int32_t j = col[r0 + t];
    const int32_t s0 = rp[j], len = rp[j + 1] - s0;
    if (len > kRingCache) return ring_problem_rows(rp, col, lab, i, a);
    int32_t cj[kRingCache];
ACVD_UNROLL
    for (int u = 0; u < kRingCache; ++u) cj[u] = u < len ? col[s0 + u] : -2;
    uint64_t hit = 0;
ACVD_UNROLL
    for (int p = 0; p < kRingCache; ++p) {
      bool h = false;
ACVD_UNROLL
      for (int u = 0; u < kRingCache; ++u) h |= (ci[p] == cj[u]);
      if (p < deg && h) hit |= uint64_t(1) << p;
    }
    const uint64_t add = hit & M & ~reach;
    reach |= add;
    frontier |= add;
  }
  return reach != M;
}

// The same check with merge-walks of the sorted rows (any length up to 64).
ACVD_HD bool ring_problem_rows(const int32_t* rp, const int32_t* col, const int32_t* lab, int32_t i, int32_t a) {
  const int32_t r0 = rp[i], r1 = rp[i + 1];
  const int32_t deg = r1 - r0;
  if (deg > 64) return true;
  uint64_t M = 0;
  for (int32_t k = 0; k < deg; ++k)
    if (lab[col[r0 + k]] == a) M |= uint64_t(1) << k;
  if (M == 0) return false;
  uint64_t reach = M & (~M + 1);
  uint64_t frontier = reach;
  while (frontier) {
    int t = __builtin_ctzll(frontier);
    frontier &= frontier - 1;
    const int32_t j = col[r0 + t];
    const int32_t s0 = rp[j], s1 = rp[j + 1];
    // merge-walk the two sorted rows
    int32_t p = r0, q = s0;
    while (p < r1 && q < s1) {
      const int32_t u = col[p], v = col[q];
      if (u < v) {
        ++p;
      } else if (v < u) {
        ++q;
      } else {
        const uint64_t bit = uint64_t(1) << (p - r0);
        if ((M & bit) && !(reach & bit)) {
          reach |= bit;
          frontier |= bit;
        }
        ++p;
        ++q;
      }
    }
  }
  return reach != M;
}

// Ring adjacency of item i (rows of at most kRingAdjMax neighbours): bit
// 8 * p + q is set when ring neighbours p and q (positions in the CSR row of
// i) share a mesh edge.  A property of the mesh alone, built once; with it
// the ring check needs only the labels of the ring neighbours.
constexpr int kRingAdjMax = 8;
ACVD_HD uint64_t ring_adjacency(const int32_t* rp, const int32_t* col, int32_t i) {
  const int32_t r0 = rp[i], deg = rp[i + 1] - r0;
  if (deg > kRingAdjMax) return 0;
  uint64_t adj = 0;
  for (int p = 0; p < deg; ++p) {
    const int32_t j = col[r0 + p];
    for (int32_t k = rp[j]; k < rp[j + 1]; ++k) {
      const int32_t u = col[k];
      for (int q = 0; q < deg; ++q)
        if (col[r0 + q] == u) adj |= uint64_t(1) << (8 * p + q);
    }
  }
  return adj;
}

// The ring check from the ring adjacency and the ring labels nb[0..deg)
// (deg <= kRingAdjMax): do the neighbours labelled a form one connected set
// through edges between themselves?  Same answer as ring_problem.
ACVD_HD bool ring_problem_mask(uint64_t adj, const int32_t* nb, int deg, int32_t a) {
  uint32_t M = 0;
  for (int p = 0; p < kRingAdjMax; ++p)
    if (p < deg && nb[p] == a) M |= 1u << p;
  if (M == 0) return false;
  uint32_t reach = M & (~M + 1u);
  uint32_t frontier = reach;
  while (frontier) {
    const int t = __builtin_ctz(frontier);
    frontier &= frontier - 1;
    const uint32_t add = uint32_t(adj >> (8 * t)) & 0xffu & M & ~reach;
    reach |= add;
    frontier |= add;
  }
  return reach != M;
}

// ---------------------------------------------------------------------------
// Read-only view of the sweep state used by the propose rule.  On the device
// the pointers are global memory; in the replay they are host arrays.
struct SweepView {
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* labels;
  const float* xw;                   // 4 floats per item: x, y, z, w
  const float* q;                    // 10 floats per item (QEM)
  const unsigned long long* acc_ws;  // 4 per cluster (int64 fixed bits)
  const unsigned long long* acc_q;   // 10 per cluster
  const double* evec;                // 9 per cluster: eigenvectors of its quadric (QEM, constrained), or null
  const ClusterGeom* geom;
  const double* T;
  const int32_t* sizes;
  const int32_t* stamp;
  const int32_t* cand_pass;
  const unsigned long long* ring_adj;  // ring adjacency per item (rows <= kRingAdjMax), or null
  const int32_t* anchor;             // per cluster: its anchor item (fixed clusters) or -1; null if none
  const int32_t* frozen;             // per cluster: nonzero when frozen; null if none is
};

// Frozen clusters (IsClusterFreezed, vtkUniformClustering.h:910-919, set by
// the ForceManifold restart, DiscreteRemeshing/vtkDiscreteRemeshing.h:
// 165-250): an item moves only between two clusters that are not frozen; a
// null item still joins any neighbouring cluster (:881-907 come first).
ACVD_HD bool is_frozen(const SweepView& v, int32_t c) { return v.frozen && v.frozen[c]; }

// Anchor position of cluster c (FP32 item position, as the sweep's items), if any.
ACVD_HD AnchorPos anchor_pos(const SweepView& v, int32_t c) {
  AnchorPos p{0.0, 0.0, 0.0, false};
  if (!v.anchor) return p;
  const int32_t it = v.anchor[c];
  if (it < 0) return p;
  p.x = double(v.xw[4 * int64_t(it)]);
  p.y = double(v.xw[4 * int64_t(it) + 1]);
  p.z = double(v.xw[4 * int64_t(it) + 2]);
  p.valid = true;
  return p;
}

// Item i is the anchor of its cluster lab: it never moves.
ACVD_HD bool is_anchor(const SweepView& v, int64_t i, int32_t lab) {
  return v.anchor && lab != kNull && v.anchor[lab] == int32_t(i);
}

// Ring check of item i leaving cluster a: with the ring adjacency for short
// rows, else by the row walk.
ACVD_HD bool ring_check(const SweepView& v, int64_t i, int32_t a) {
  const int32_t r0 = v.row_ptr[i], deg = v.row_ptr[i + 1] - r0;
  if (v.ring_adj && deg <= kRingAdjMax) {
    int32_t nb[kRingAdjMax];
    ACVD_UNROLL
    for (int p = 0; p < kRingAdjMax; ++p) nb[p] = p < deg ? v.labels[v.col[r0 + p]] : kNull;
    return ring_problem_mask(v.ring_adj[i], nb, deg, a);
  }
  return ring_problem(v.row_ptr, v.col, v.labels, int32_t(i), a);
}

// Item i must be evaluated in the pass of loop L: it is on a cluster
// boundary, and its own cluster or a neighbouring one changed in this or the
// previous loop (the reference's skip test, vtkUniformClustering.h:909-920),
// or its proposal of the previous loop lost the accept step.
ACVD_HD bool needs_eval(const SweepView& v, int64_t i, int32_t lab, int32_t L) {
  bool dirty = (v.cand_pass[i] == L - 1);
  if (lab != kNull) dirty |= (v.stamp[lab] + 1 >= L);
  bool boundary = false;
  const int32_t r1 = v.row_ptr[i + 1];
  for (int32_t k = v.row_ptr[i]; k < r1; ++k) {
    const int32_t b = v.labels[v.col[k]];
    if (b != lab) {
      boundary = true;
      if (b != kNull && v.stamp[b] + 1 >= L) dirty = true;
    }
  }
  return boundary && dirty;
}

ACVD_HD void load_acc(const SweepView& v, int32_t c, int64_t* ws, int64_t* q10) {
  for (int k = 0; k < 4; ++k) ws[k] = int64_t(v.acc_ws[4 * int64_t(c) + k]);
  for (int k = 0; k < 10; ++k) q10[k] = int64_t(v.acc_q[10 * int64_t(c) + k]);
}

// ---------------------------------------------------------------------------
// QEM evaluation steps: screen (guards and exact prune), FP64 solves, choice.
// QemItem holds what the choice needs besides the solves; everything is read
// from the state frozen for the current round.
struct QemItem {
  float px, py, pz, pw;
  float ta;      // leave term of the item's cluster
  double Ta;     // T of the item's cluster (0 when unconstrained)
  int32_t r0, r1;
};

ACVD_HD void qem_item_terms(const SweepView& v, int64_t i, int32_t lab, int constrained, QemItem* s) {
  s->px = v.xw[4 * i];
  s->py = v.xw[4 * i + 1];
  s->pz = v.xw[4 * i + 2];
  s->pw = v.xw[4 * i + 3];
  s->r0 = v.row_ptr[i];
  s->r1 = v.row_ptr[i + 1];
  s->ta = leave_term(s->px, s->py, s->pz, s->pw, v.geom[lab]);
  s->Ta = constrained ? v.T[lab] : 0.0;
}

// Exact prune before any FP64 solve: T' >= 0, so with monotone rounding the
// computed dE of a move to b is at least LB_b = double(w (tb - ta)) + (-(Ta + Tb));
// a target with LB_b >= 0 can neither pass the acceptance test nor be the
// best of a set that contains an acceptable move, so it is skipped.
ACVD_HD bool qem_open(const SweepView& v, const QemItem& s, int32_t b, int constrained) {
  if (is_frozen(v, b)) return false;
  const float tb = join_term(s.px, s.py, s.pz, s.pw, v.geom[b]);
  const double Tb = constrained ? v.T[b] : 0.0;
  return double(s.pw * (tb - s.ta)) + (-(s.Ta + Tb)) < 0.0;
}

// CSR position of the next open target after position k (each neighbouring
// cluster once, at its first position), or s.r1.
ACVD_HD int32_t qem_next_open(const SweepView& v, int32_t lab, const QemItem& s, int constrained, int32_t k) {
  for (++k; k < s.r1; ++k) {
    const int32_t b = v.labels[v.col[k]];
    if (b == kNull || b == lab) continue;
    bool seen = false;
    for (int32_t t = s.r0; t < k; ++t)
      if (v.labels[v.col[t]] == b) {
        seen = true;
        break;
      }
    if (seen) continue;
    if (qem_open(v, s, b, constrained)) return k;
  }
  return s.r1;
}

// Guards (size, ring check, leave condition) and prune of a labelled item.
// Returns true when at least one target is open (solves and choice follow).
ACVD_HD bool qem_screen(const SweepView& v, int64_t i, int32_t lab, int connexity, int constrained, QemItem* s) {
  if (v.sizes[lab] <= 1) return false;
  if (is_anchor(v, i, lab) || is_frozen(v, lab)) return false;
  if (connexity && ring_check(v, i, lab)) return false;
  if (!can_leave(v.xw[4 * i + 3], v.geom[lab])) return false;
  qem_item_terms(v, i, lab, constrained, s);
  for (int32_t k = s->r0; k < s->r1; ++k) {
    const int32_t b = v.labels[v.col[k]];
    if (b == kNull || b == lab) continue;
    if (qem_open(v, *s, b, constrained)) return true;
  }
  return false;
}

// T of the item's cluster without the item, and of cluster b with it.
ACVD_HD void item_fixed(const SweepView& v, int64_t i, const Scales& sc, int64_t* iw, int64_t* is, int64_t* iq) {
  const float px = v.xw[4 * i], py = v.xw[4 * i + 1], pz = v.xw[4 * i + 2], pw = v.xw[4 * i + 3];
  *iw = fix_w(pw, sc);
  is[0] = fix_s(pw, px, sc);
  is[1] = fix_s(pw, py, sc);
  is[2] = fix_s(pw, pz, sc);
  for (int k = 0; k < 10; ++k) iq[k] = fix_q(v.q[10 * i + k], sc);
}
// mode (qem_mode): 1 every cluster's point is solved, 2 only anchored
// clusters carry a T (the others keep their centroid), 0 no T at all.
ACVD_HD double qem_leave_T(const SweepView& v, int64_t i, int32_t lab, const Scales& sc, int level, int mode) {
  const AnchorPos an = anchor_pos(v, lab);
  if (!an.valid && mode != 1) return 0.0;
  int64_t iw, is[3], iq[10], ws[4], q10[10];
  item_fixed(v, i, sc, &iw, is, iq);
  load_acc(v, lab, ws, q10);
  return qem_T_trial(ws, q10, iw, is, iq, -1, sc, level, v.evec ? v.evec + 9 * int64_t(lab) : nullptr, an);
}
ACVD_HD double qem_join_T(const SweepView& v, int64_t i, int32_t b, const Scales& sc, int level, int mode) {
  const AnchorPos an = anchor_pos(v, b);
  if (!an.valid && mode != 1) return 0.0;
  int64_t iw, is[3], iq[10], ws[4], q10[10];
  item_fixed(v, i, sc, &iw, is, iq);
  load_acc(v, b, ws, q10);
  return qem_T_trial(ws, q10, iw, is, iq, +1, sc, level, v.evec ? v.evec + 9 * int64_t(b) : nullptr, an);
}

// Choice among the open targets (CSR order; ties: lowest cluster id) and the
// acceptance test.  tb_new(k, b) returns T of cluster b (first seen at CSR
// position k) with the item.
template <class TbNew>
ACVD_HD bool qem_choose(const SweepView& v, int32_t lab, const QemItem& s, int constrained, double Ta_new,
                        TbNew tb_new, int32_t* target_out) {
  int32_t target = kNull;
  double best_d = 0.0, best_m = 0.0;
  for (int32_t k = qem_next_open(v, lab, s, constrained, s.r0 - 1); k < s.r1;
       k = qem_next_open(v, lab, s, constrained, k)) {
    const int32_t b = v.labels[v.col[k]];
    const float tb = join_term(s.px, s.py, s.pz, s.pw, v.geom[b]);
    const double Tb = constrained ? v.T[b] : 0.0;
    const double Tb_new = tb_new(k, b);
    const double d = qem_delta(s.pw, s.ta, tb, Ta_new, s.Ta, Tb_new, Tb);
    if (target == kNull || d < best_d || (d == best_d && b < target)) {
      best_d = d;
      best_m = qem_mag(s.pw, s.ta, tb, Ta_new, s.Ta, Tb_new, Tb);
      target = b;
    }
  }
  *target_out = target;
  return target != kNull && qem_accept(best_d, best_m);
}

// nb[p] for a run-time p, selected in registers (nb stays a register array).
ACVD_HD int32_t ring_label(const int32_t* nb, int p) {
  int32_t b = kNull;
  ACVD_UNROLL
  for (int q = 0; q < kRingCache; ++q)
    if (q == p) b = nb[q];
  return b;
}

// Screen of the QEM rule for rows of at most kRingCache neighbours: the
// neighbour labels are read once into nb (independent loads, no chains
// through the row), then the guards of qem_screen (size, ring check, leave
// condition) and the exact prune of each neighbouring cluster at its first
// CSR position.  Bit p of *open marks an open target nb[p]; returns *open != 0.
ACVD_HD bool qem_screen_short_row(const SweepView& v, int64_t i, int32_t lab, int connexity, int constrained,
                                  int32_t* nb, QemItem* s, uint32_t* open_out) {
  const int32_t r0 = v.row_ptr[i], deg = v.row_ptr[i + 1] - r0;
  // the words of the item and of its cluster, read with the row bounds (they
  // do not depend on the row), so that the guards below wait on one round trip
  const int32_t size_a = v.sizes[lab];
  const unsigned long long adj = v.ring_adj ? v.ring_adj[i] : 0ull;
  const float w_i = v.xw[4 * i + 3];
  const ClusterGeom ga = v.geom[lab];
  ACVD_UNROLL
  for (int p = 0; p < kRingCache; ++p) nb[p] = p < deg ? v.labels[v.col[r0 + p]] : kNull;
  *open_out = 0;
  if (size_a <= 1) return false;
  if (is_anchor(v, i, lab) || is_frozen(v, lab)) return false;
  if (connexity) {
    const bool split = (v.ring_adj && deg <= kRingAdjMax)
                           ? ring_problem_mask(adj, nb, deg, lab)
                           : ring_problem(v.row_ptr, v.col, v.labels, int32_t(i), lab);
    if (split) return false;
  }
  if (!can_leave(w_i, ga)) return false;
  qem_item_terms(v, i, lab, constrained, s);
  uint32_t open = 0;
  ACVD_UNROLL
  for (int p = 0; p < kRingCache; ++p) {
    bool first = nb[p] != kNull && nb[p] != lab;
    ACVD_UNROLL
    for (int t = 0; t < p; ++t) first = first && nb[t] != nb[p];
    if (first && qem_open(v, *s, nb[p], constrained)) open |= 1u << p;
  }
  *open_out = open;
  return open != 0;
}

// The QEM rule for rows of at most kRingCache neighbours: the same tests in
// the same CSR order as qem_screen / qem_choose, so it returns the same move.
ACVD_HD bool qem_propose_short_row(const SweepView& v, int64_t i, int32_t lab, int connexity, int constrained,
                                   int level, const Scales& sc, int32_t* target_out) {
  int32_t nb[kRingCache];
  QemItem s;
  uint32_t open;
  if (!qem_screen_short_row(v, i, lab, connexity, constrained, nb, &s, &open)) return false;
  const double Ta_new = qem_leave_T(v, i, lab, sc, level, constrained);
  int32_t target = kNull;
  double best_d = 0.0, best_m = 0.0;
  while (open) {
    const int p = __builtin_ctz(open);
    open &= open - 1;
    const int32_t b = ring_label(nb, p);
    const float tb = join_term(s.px, s.py, s.pz, s.pw, v.geom[b]);
    const double Tb = constrained ? v.T[b] : 0.0;
    const double Tb_new = qem_join_T(v, i, b, sc, level, constrained);
    const double d = qem_delta(s.pw, s.ta, tb, Ta_new, s.Ta, Tb_new, Tb);
    if (target == kNull || d < best_d || (d == best_d && b < target)) {
      best_d = d;
      best_m = qem_mag(s.pw, s.ta, tb, Ta_new, s.Ta, Tb_new, Tb);
      target = b;
    }
  }
  *target_out = target;
  return target != kNull && qem_accept(best_d, best_m);
}

// Best move of an evaluated item (label lab, loop L).  Returns true and the
// target cluster if the move is proposed.
//  - null item: joins the neighbouring cluster with the smallest join term
//    (the reference absorbs it into the first neighbour met,
//    vtkUniformClustering.h:881-907);
//  - otherwise: no move out of a size-1 cluster or against the ring check
//    (vtkUniformClustering.h:929, 945), then the neighbouring cluster of
//    lowest dE (ties: lowest cluster id) if dE passes the acceptance test.
// kShortRows = false forces the generic QEM path for every row (the oracle
// checks both paths give the same moves).
template <int METRIC, bool kShortRows = true>
ACVD_HD bool propose_item(const SweepView& v, int64_t i, int32_t lab, int connexity, int constrained, int level,
                          const Scales& sc, int32_t* target_out) {
  const float px = v.xw[4 * i], py = v.xw[4 * i + 1], pz = v.xw[4 * i + 2], pw = v.xw[4 * i + 3];
  const int32_t r0 = v.row_ptr[i], r1 = v.row_ptr[i + 1];
  int32_t target = kNull;
  // the item's cluster words, read with the item (independent of the row)
  const int32_t size_a = lab != kNull ? v.sizes[lab] : 0;
  const ClusterGeom ga = lab != kNull ? v.geom[lab] : ClusterGeom{0.0, 0.0, 0.0, 0.0};
  if (lab == kNull) {
    float best = 3.0e38f;
    for (int32_t k = r0; k < r1; ++k) {
      const int32_t b = v.labels[v.col[k]];
      if (b == kNull) continue;
      const float tb = join_term(px, py, pz, pw, v.geom[b]);
      if (tb < best || (tb == best && b < target)) {
        best = tb;
        target = b;
      }
    }
    *target_out = target;
    return target != kNull;
  }
  // QEM with active constraints (constrained = qem_mode: 1, or 2 when only
  // anchored clusters carry a T): the T terms take part.  Without them (the
  // unconstrained initialisation of a mesh without fixed clusters,
  // vtkUniformClustering.h:683-688) the QEM move energy is w (tb - ta), the
  // isotropic one, and the isotropic rule below decides; METRIC 2 is that
  // case alone (QEM meshes swept while none is constrained), compiled without
  // the FP64 solves.
  if (METRIC == 1 && constrained) {
    if (kShortRows && r1 - r0 <= kRingCache)
      return qem_propose_short_row(v, i, lab, connexity, constrained, level, sc, target_out);
    // QEM: screen, solves, choice
    QemItem s;
    if (!qem_screen(v, i, lab, connexity, constrained, &s)) return false;
    const double Ta_new = qem_leave_T(v, i, lab, sc, level, constrained);
    return qem_choose(
        v, lab, s, constrained, Ta_new, [&](int32_t, int32_t b) { return qem_join_T(v, i, b, sc, level, constrained); },
        target_out);
  }
  if (size_a <= 1 || is_frozen(v, lab)) return false;
  if (connexity && ring_check(v, i, lab)) return false;
  if (!can_leave(pw, ga)) return false;
  const float ta = leave_term(px, py, pz, pw, ga);
  float best = 3.0e38f;
  for (int32_t k = r0; k < r1; ++k) {
    const int32_t b = v.labels[v.col[k]];
    if (b == kNull || b == lab || is_frozen(v, b)) continue;
    const float tb = join_term(px, py, pz, pw, v.geom[b]);
    if (tb < best || (tb == best && b < target)) {
      best = tb;
      target = b;
    }
  }
  *target_out = target;
  return target != kNull && iso_accept(ta, best);
}

// Re-check, in a later accept round of the pass, of a stale proposal that the
// isotropic rule decides (the isotropic metric, or a QEM mesh while no T term
// is active), item i of cluster lab proposing target b: the move is kept
// while it is still one the evaluate could propose (b still a neighbouring
// cluster, the same guards as propose_item) and still improving; no better
// target is searched (the next pass's evaluate does that).  Null items are
// proposed again in full.
ACVD_HD bool iso_recheck(const SweepView& v, int64_t i, int32_t lab, int32_t b, int connexity) {
  if (v.sizes[lab] <= 1 || is_frozen(v, lab) || is_frozen(v, b)) return false;
  const int32_t r0 = v.row_ptr[i], r1 = v.row_ptr[i + 1];
  bool adjacent = false;
  for (int32_t k = r0; k < r1 && !adjacent; ++k) adjacent = v.labels[v.col[k]] == b;
  if (!adjacent) return false;
  if (connexity && ring_check(v, i, lab)) return false;
  const float px = v.xw[4 * i], py = v.xw[4 * i + 1], pz = v.xw[4 * i + 2], pw = v.xw[4 * i + 3];
  const ClusterGeom ga = v.geom[lab];
  if (!can_leave(pw, ga)) return false;
  return iso_accept(leave_term(px, py, pz, pw, ga), join_term(px, py, pz, pw, v.geom[b]));
}

// Per-mesh convergence state, advanced once per pass by the controller with
// the reference's phase logic (vtkUniformClustering.h:724-830).
enum MeshPhase : int32_t {
  kPhaseSweep = 0,     // boundary passes running
  kPhaseNeedClean = 1, // a convergence was reached; clean + fill pending
  kPhaseDone = 2,
  kPhaseIdle = 3       // not part of the current run
};

struct MeshState {
  int32_t phase;
  int32_t loops;            // NumberOfLoops
  int32_t convergences;     // NumberOfConvergences
  int32_t connexity;        // ConnexityConstraint
  int32_t constrained;      // QEM ActiveConstraints
  int32_t unconstrained_init;
  int32_t max_loops;
  int32_t max_convergences;
  int64_t early_stop_items; // items whose cluster is not frozen, at MinimizeEnergy start
  int32_t last_mods;        // modifications of the last pass
  int32_t pass_mods;        // accumulated by the apply step of the current pass
  int64_t total_mods;
  int64_t evaluated;        // boundary items evaluated (statistics)
  int64_t proposals;        // candidates proposed (statistics)
  int32_t last_disconnected;
  int32_t pass_props;       // proposals (improving moves found) of the current pass
  int32_t first_rounds;     // accept rounds per pass until the first convergence (1..kMatchRounds)
  int32_t anchored;         // QEM mesh with fixed clusters (anchored cluster points in every phase)
};

// T terms of a mesh's QEM rule: 1 every cluster point solved (constrained),
// 2 only anchored clusters carry one, 0 none (the isotropic rule decides).
ACVD_HD int qem_mode(const MeshState& s) { return s.constrained ? 1 : (s.anchored ? 2 : 0); }

// Accept rounds of the mesh's current pass.  Until the first convergence a
// pass runs first_rounds rounds, afterwards kMatchRounds.  The phase before
// the first convergence decides which local minimum the run reaches and most
// of its moves.  With one round per pass every cluster moves at most one item
// per pass and the clusters relax evenly; with twelve the moves of a pass
// follow the proposals of its start and far fewer passes are needed.
// Measured on the replay against the serial reference (same initial labels):
//   icosphere 40,962 vertices, K = 2,000, isotropic, seed-vertex start, eight
//   accept seeds: gap median +0.02 % (range -0.14 .. +0.24) with 12 rounds,
//   -0.10 % (-0.19 .. +0.04) with 1; 44 against 112 passes;
//   cube 960,002 vertices, K = 20,000, QEM: +0.01 .. +0.03 % with 12,
//   -0.001 .. +0.002 % with 1;
//   graded QEM torus 524,288 vertices, K = 6,250: -0.04 / +0.01 % with 12
//   rounds (173 passes), +0.02 / -0.03 % with 1 (~700 passes);
//   on the GPU the 8,388,608-vertex torus (K = 100,000) takes 384 passes and
//   239 ms with 12 rounds, 1,556 passes and 403 ms with 1.
//   star-shaped icospheres, K = V/15, isotropic, two meshes x four accept
//   seeds per size (gap median, max; mean passes) with 1 / 2 / 12 rounds:
//   10,242 vertices  -0.30 % +0.13 54 / -0.08 % +0.39 33 / +0.01 % +0.40 24;
//   40,962           -0.04 % +0.04 60 / -0.02 % +0.07 39 / +0.10 % +0.18 29;
//   163,842          -0.04 % -0.01 62 / +0.02 % +0.06 43 / +0.03 % +0.09 34;
//   655,362          -0.07 % -0.05 68 / -0.03 % +0.00 49 / +0.02 % +0.02 39.
// So the default is one round for meshes of at most 2^17 items, where the
// distinct local minima differ most and a pass costs little, two up to 2^19
// (within 0.1 % at 30 % fewer passes) and twelve above, where the gap is
// within noise and the passes dominate.
constexpr int32_t kFirstRoundsOneItems = int32_t(1) << 17;
constexpr int32_t kFirstRoundsTwoItems = int32_t(1) << 19;
ACVD_HD int32_t auto_first_rounds(int64_t n_items) {
  return n_items <= kFirstRoundsOneItems ? 1 : (n_items <= kFirstRoundsTwoItems ? 2 : kMatchRounds);
}

ACVD_HD int32_t pass_rounds(const MeshState& s) { return s.convergences == 0 ? s.first_rounds : kMatchRounds; }

// Early stop of the first two phases.  The reference stops a phase when a
// loop applies at most N/1000 moves (vtkUniformClustering.h:773-776), a loop
// being a sweep of the whole edge queue with immediate updates.  A GPU pass
// applies only matchings of the improving moves it finds, so the analogue
// here counts the improving moves found (proposals) in a pass.  The
// threshold N/3000 was chosen by measuring final energies against the serial
// reference: on a graded QEM torus (524,288 vertices, K = 6,250) N/10000,
// N/3000 and N/1000 end within 0.06 % of each other while N/3000 saves 15 %
// of the passes of N/10000.
constexpr int64_t kEarlyStopDivisor = 3000;

// Controller transition after one pass (the decision part of MinimizeEnergy).
// Returns true if the mesh reached a convergence (clean + fill must follow).
ACVD_HD bool controller_step(MeshState& s) {
  const int mods = s.pass_mods;
  const int props = s.pass_props;
  s.last_mods = mods;
  s.total_mods += mods;
  s.proposals += props;
  s.pass_mods = 0;
  s.pass_props = 0;
  s.loops++;
  bool conv = (mods == 0) || (s.loops > s.max_loops) ||
              (int64_t(props) <= s.early_stop_items / kEarlyStopDivisor && s.convergences <= 1);
  if (!conv) return false;
  if (s.unconstrained_init && s.convergences == 0) s.constrained = 1;
  if (s.convergences >= 1) s.connexity = 1;
  s.convergences++;
  s.phase = kPhaseNeedClean;
  return true;
}

// Decision after clean + fill: true if the mesh is finished.
// Deliberate difference from the reference (vtkUniformClustering.h:800-801):
// when the first convergence also switched the metric from unconstrained to
// constrained (ACVDQ, :778-782), a zero-move pass under the old metric does
// not end the run, so the QEM phase always runs.
ACVD_HD bool finish_after_clean(MeshState& s, int disconnected) {
  s.last_disconnected = disconnected;
  const bool metric_just_switched = s.unconstrained_init && s.convergences == 1;
  if (disconnected == 0 && s.last_mods == 0 && !metric_just_switched) return true;
  if (s.loops >= s.max_loops) return true;
  if (s.convergences >= s.max_convergences) return true;
  return false;
}

}  // namespace acvd
