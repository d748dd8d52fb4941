// TEST INFRASTRUCTURE — CPU restatement of the reference's per-vertex
// curvature indicator (vtkCurvatureMeasure with ComputationMethod 1,
// ElementsType 1, RingSize 3, as vtkDiscreteRemeshing::SamplingPreProcessing
// sets it up, DiscreteRemeshing/vtkDiscreteRemeshing.h:644-653).  Only tests/
// use it, as the checker of the device kernel.
//
//   face list   vtkNeighbourhoodComputation::ComputeNRingCells
//               (Common/vtkNeighbourhoodComputation.cxx:38-110): the faces
//               around the vertex, then three rounds over the ring edges of
//               the not yet visited vertices of the last round, adding the
//               faces of each edge once (= every face incident to a vertex
//               within graph distance 2);
//   fit         vtkSinglePolynomialMeasure::ComputeFitting
//               (Common/vtkCurvatureMeasure.cxx:188-509): area-weighted
//               origin and mean normal, the frame built from the permuted
//               normal, z = a0 + a1 x + a2 y + a3 x^2 + a4 xy + a5 y^2 fitted
//               to the face barycentres (columns scaled by the mean |(x,y)|),
//               least squares by the normal equations and their inverse
//               (SolveLeastSquares, :511-620), the shape operator
//               -II I^-1 with f = 2 a4 / sqrt(1 + a1^2 + a2^2) as written at
//               :340-347, and the two eigenvalues of its upper-symmetrised
//               2x2 (vtkMath::JacobiN reads the upper triangle); indicator =
//               sqrt(k1^2 + k2^2), 0 for fewer than 7 faces or a singular
//               system.
// The order of the face list follows the ring-edge order of RefMesh (the
// reference's edge-insertion order); it changes only the rounding of the sums.
// Parity: pinned by closed forms (a plane gives 0, a sphere of radius R tends
// to sqrt(2)/R, a cylinder of radius r to 1/r) and by an independent numpy
// least-squares fit of the same face data (tests/test_curvature_oracle.py).
// Unpinned: the rounding of vtkMath::InvertMatrix / JacobiN and of
// vtkTriangle::ComputeNormal (VTK, not available here).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "ref_mesh.h"

namespace oracle {

// Inverse of an n x n matrix by LU with partial pivoting (the role of
// vtkMath::InvertMatrix); false when a pivot is zero.
inline bool invert_matrix(const std::vector<double>& A, std::vector<double>& Ainv, int n) {
  std::vector<double> M(A);
  Ainv.assign(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) Ainv[size_t(i) * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(M[size_t(r) * n + c]) > std::fabs(M[size_t(p) * n + c])) p = r;
    if (M[size_t(p) * n + c] == 0.0) return false;
    if (p != c)
      for (int k = 0; k < n; ++k) {
        std::swap(M[size_t(p) * n + k], M[size_t(c) * n + k]);
        std::swap(Ainv[size_t(p) * n + k], Ainv[size_t(c) * n + k]);
      }
    const double d = M[size_t(c) * n + c];
    for (int k = 0; k < n; ++k) {
      M[size_t(c) * n + k] /= d;
      Ainv[size_t(c) * n + k] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = M[size_t(r) * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        M[size_t(r) * n + k] -= f * M[size_t(c) * n + k];
        Ainv[size_t(r) * n + k] -= f * Ainv[size_t(c) * n + k];
      }
    }
  }
  return true;
}

// Face list of vertex v (Common/vtkNeighbourhoodComputation.cxx:38-110, vertex
// branch, RingSize 3).
inline void curvature_face_list(const RefMesh& m, int64_t v, std::vector<int32_t>& flist) {
  std::vector<char> face_seen(size_t(m.nf), 0), vert_seen(size_t(m.nv), 0);
  m.neighbour_faces(v, flist);
  for (int32_t f : flist) face_seen[f] = 1;
  std::vector<int32_t> vlist{int32_t(v)}, vlist2;
  for (int round = 0; round < 3; ++round) {
    vlist2.clear();
    for (int32_t v1 : vlist) {
      if (vert_seen[v1]) continue;
      vert_seen[v1] = 1;
      for (int64_t k = m.ring_ptr[v1]; k < m.ring_ptr[v1 + 1]; ++k) {
        const int32_t e = m.ring[k];
        vlist2.push_back(m.ev1[e] == v1 ? m.ev2[e] : m.ev1[e]);
        for (int32_t f : {m.poly1[e], m.poly2[e]}) {
          if (f < 0) break;
          if (!face_seen[f]) {
            face_seen[f] = 1;
            flist.push_back(f);
          }
        }
      }
    }
    vlist.swap(vlist2);
  }
}

// vtkSinglePolynomialMeasure::ComputeFitting (Common/vtkCurvatureMeasure.cxx:188-509).
inline double curvature_fit(const RefMesh& m, const std::vector<int32_t>& flist) {
  const int n = int(flist.size());
  double s_area = 0.0, origin[3] = {0, 0, 0}, frame[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  std::vector<double> area(n), bary(3 * size_t(n));
  for (int j = 0; j < n; ++j) {
    const int32_t f = flist[j];
    double p1[3], p2[3], p3[3];
    m.point(m.T[3 * f], p1);
    m.point(m.T[3 * f + 1], p2);
    m.point(m.T[3 * f + 2], p3);
    // vtkTriangle::ComputeNormal: the unit normal of (p1, p2, p3)
    const double a[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    const double b[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
    double nrm[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double ln = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    if (ln != 0.0)
      for (double& x : nrm) x /= ln;
    // vtkSurface::GetCellMassProperties (Common/vtkSurface.cxx:1380-1419)
    const double ar = m.face_area(f);
    double* bc = &bary[3 * size_t(j)];
    for (int k = 0; k < 3; ++k) bc[k] = ar > 0.0 ? ar * (p1[k] + p2[k] + p3[k]) / 3.0 / ar : 0.0;
    area[j] = ar;
    for (int k = 0; k < 3; ++k) {
      origin[k] += ar * bc[k];
      frame[0][k] += ar * nrm[k];
    }
    s_area += ar;
  }
  for (double& x : origin) x /= s_area;
  auto normalize = [](double* x) {
    const double l = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    if (l != 0.0)
      for (int k = 0; k < 3; ++k) x[k] /= l;
  };
  auto cross = [](const double* x, const double* y, double* z) {
    z[0] = x[1] * y[2] - x[2] * y[1];
    z[1] = x[2] * y[0] - x[0] * y[2];
    z[2] = x[0] * y[1] - x[1] * y[0];
  };
  normalize(frame[0]);
  for (int k = 0; k < 3; ++k) frame[2][k] = frame[0][k];
  frame[1][1] = frame[2][0];
  frame[1][2] = frame[2][1];
  frame[1][0] = frame[2][2];
  cross(frame[1], frame[2], frame[0]);
  normalize(frame[0]);
  cross(frame[2], frame[0], frame[1]);
  if (n <= 6) return 0.0;
  std::vector<double> X(size_t(n) * 6), Z(n);
  double h = 0.0;
  for (int j = 0; j < n; ++j) {
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = bary[3 * size_t(j) + k] - origin[k];
    const double x = d[0] * frame[0][0] + d[1] * frame[0][1] + d[2] * frame[0][2];
    const double y = d[0] * frame[1][0] + d[1] * frame[1][1] + d[2] * frame[1][2];
    const double z = d[0] * frame[2][0] + d[1] * frame[2][1] + d[2] * frame[2][2];
    double* r = &X[size_t(j) * 6];
    r[0] = 1.0;
    r[1] = x;
    r[2] = y;
    r[3] = x * x;
    r[4] = x * y;
    r[5] = y * y;
    Z[j] = z;
    h += std::sqrt(x * x + y * y);
  }
  h = h / double(n);
  for (int j = 0; j < n; ++j) {
    double* r = &X[size_t(j) * 6];
    r[1] /= h;
    r[2] /= h;
    r[3] /= h * h;
    r[4] /= h * h;
    r[5] /= h * h;
  }
  // SolveLeastSquares (:511-620): XXt (upper half, mirrored), XYt, inverse
  std::vector<double> XX(36, 0.0), XY(6, 0.0), XXi;
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < 6; ++i) {
      for (int j = i; j < 6; ++j) XX[i * 6 + j] += X[size_t(k) * 6 + i] * X[size_t(k) * 6 + j];
      XY[i] += X[size_t(k) * 6 + i] * Z[k];
    }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < i; ++j) XX[i * 6 + j] = XX[j * 6 + i];
  if (!invert_matrix(XX, XXi, 6)) return 0.0;
  double q[6];
  for (int i = 0; i < 6; ++i) {
    q[i] = 0.0;
    for (int k = 0; k < 6; ++k) q[i] += XXi[i * 6 + k] * XY[k];
  }
  q[1] /= h;
  q[2] /= h;
  q[3] /= h * h;
  q[4] /= h * h;
  q[5] /= h * h;
  const double E = 1.0 + q[1] * q[1], F = q[1] * q[2], G = 1.0 + q[2] * q[2];
  const double s = std::sqrt(q[1] * q[1] + 1.0 + q[2] * q[2]);
  const double e = 2.0 * q[3] / s, f = 2.0 * q[4] / s, g = 2.0 * q[5] / s;
  std::vector<double> A2 = {E, F, F, G}, B2;
  if (!invert_matrix(A2, B2, 2)) return 0.0;
  const double B[2][2] = {{B2[0], B2[1]}, {B2[2], B2[3]}};
  const double a00 = -(e * B[0][0] + f * B[1][0]);
  const double a01 = -(f * B[0][0] + g * B[1][0]);
  const double a11 = -(f * B[0][1] + g * B[1][1]);
  // vtkMath::JacobiN on the 2x2 (upper triangle): one rotation zeroes a01
  double d0 = a00, d1 = a11;
  if (a01 != 0.0) {
    const double theta = 0.5 * (d1 - d0) / a01;
    double t = 1.0 / (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
    if (theta < 0.0) t = -t;
    d0 -= t * a01;
    d1 += t * a01;
  }
  return std::sqrt(d0 * d0 + d1 * d1);
}

// Indicator of every vertex of the mesh.
inline void curvature_indicator(const RefMesh& m, std::vector<double>& out) {
  out.assign(size_t(m.nv), 0.0);
  std::vector<int32_t> flist;
  for (int64_t v = 0; v < m.nv; ++v) {
    curvature_face_list(m, v, flist);
    out[v] = curvature_fit(m, flist);
  }
}

}  // namespace oracle
