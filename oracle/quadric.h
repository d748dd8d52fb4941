// TEST INFRASTRUCTURE — CPU oracle, not part of the product.
//
// Restatement of Common/vtkQuadricTools.cxx for the clustering path:
//   - AddTriangleQuadric (:56-79): the plane quadric of a triangle, in the
//     compact 9/10-coefficient layout
//       [0]=a*a [1]=a*b [2]=a*c [3]=a*d [4]=b*b [5]=b*c [6]=b*d [7]=c*c [8]=c*d [9]=d*d
//     with (a,b,c) = x1 x x2 + x2 x x3 + x3 x x1 and d = -det[x1;x2;x3]
//     (Lindstrom 2000, the construction Examples/ACVD.cxx:218-219 cites; the
//     reference gets it from VTK's vtkTriangle::ComputeQuadric, which is not in
//     the reference tree);
//   - ComputeDisplacement / ComputeRepresentativePoint (:83-177): the
//     rank-truncated pseudo-inverse step q = c + A+_trunc (b - A c), where A
//     is the symmetric 3x3 block and b = -(Q[3],Q[6],Q[8]).  Singular values
//     are visited in decreasing |sigma| order (first index wins ties, :119-133)
//     and kept while |sigma|/sigma_max > threshold (DEFAULT_SV_THRESHOLD =
//     1e-3, vtkQuadricTools.h:36) and at most `level` of them are used.
//
// The reference computes the SVD with VTK's SingularValueDecomposition3x3.
// For a symmetric A the truncated pseudo-inverse V diag(1/w) U^T equals
// sum_k v_k v_k^T / lambda_k over the kept eigenpairs, whatever sign/axis
// convention the SVD uses, so this file uses its own cyclic Jacobi
// eigen-decomposition in FP64, iterated to convergence.
#pragma once

#include <cmath>

namespace oracle {

inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

inline double det3(const double* c1, const double* c2, const double* c3) {
  return c1[0] * c2[1] * c3[2] + c2[0] * c3[1] * c1[2] + c3[0] * c1[1] * c2[2] - c1[0] * c3[1] * c2[2] -
         c2[0] * c1[1] * c3[2] - c3[0] * c2[1] * c1[2];
}

// Adds the plane quadric of triangle (x1,x2,x3) to Q (9 or 10 coefficients).
inline void add_triangle_quadric(double* Q, const double* x1, const double* x2, const double* x3, bool full) {
  double c12[3], c23[3], c31[3];
  cross3(x1, x2, c12);
  cross3(x2, x3, c23);
  cross3(x3, x1, c31);
  double n[4];
  for (int k = 0; k < 3; ++k) n[k] = c12[k] + c23[k] + c31[k];
  n[3] = -det3(x1, x2, x3);
  Q[0] += n[0] * n[0];
  Q[1] += n[0] * n[1];
  Q[2] += n[0] * n[2];
  Q[3] += n[0] * n[3];
  Q[4] += n[1] * n[1];
  Q[5] += n[1] * n[2];
  Q[6] += n[1] * n[3];
  Q[7] += n[2] * n[2];
  Q[8] += n[2] * n[3];
  if (full) Q[9] += n[3] * n[3];
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (FP64),
// iterated until the off-diagonal part vanishes relative to the diagonal.
// On return lam[k] is the eigenvalue for eigenvector column V[.][k].
inline void sym_eigen3(const double A_in[3][3], double lam[3], double V[3][3]) {
  double A[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      A[i][j] = A_in[i][j];
      V[i][j] = (i == j) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = std::fabs(A[0][1]) + std::fabs(A[0][2]) + std::fabs(A[1][2]);
    double dia = std::fabs(A[0][0]) + std::fabs(A[1][1]) + std::fabs(A[2][2]);
    if (off == 0.0 || off <= 1e-300 || off < dia * 1e-18) break;
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        double apq = A[p][q];
        if (apq == 0.0) continue;
        double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
        double t;
        if (std::fabs(theta) > 1e150) t = 0.5 / theta;
        else t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0);
        double s = t * c;
        // A <- J^T A J with J the rotation in plane (p,q)
        for (int k = 0; k < 3; ++k) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        A[p][q] = A[q][p] = 0.0;
        for (int k = 0; k < 3; ++k) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
  for (int k = 0; k < 3; ++k) lam[k] = A[k][k];
}

// vtkQuadricTools::ComputeDisplacement (Common/vtkQuadricTools.cxx:83-163).
// Returns the rank deficiency.
inline int compute_displacement(const double* Q, const double* P, double* D, int max_sv, double thr) {
  double A[3][3] = {{Q[0], Q[1], Q[2]}, {Q[1], Q[4], Q[5]}, {Q[2], Q[5], Q[7]}};
  double b[3] = {-Q[3], -Q[6], -Q[8]};
  double lam[3], V[3][3];
  sym_eigen3(A, lam, V);
  double absw[3], maxw = -1.0;
  for (int j = 0; j < 3; ++j) {
    absw[j] = std::fabs(lam[j]);
    if (absw[j] > maxw) maxw = absw[j];
  }
  // r = b - A P
  double r[3];
  for (int i = 0; i < 3; ++i) r[i] = b[i] - (A[i][0] * P[0] + A[i][1] * P[1] + A[i][2] * P[2]);
  int rank_def = 0;
  double keep[3] = {0, 0, 0};
  for (int i = 0; i < 3; ++i) {
    double local = -1.0;
    int idx = -1;
    for (int j = 0; j < 3; ++j)
      if (local < absw[j]) {
        local = absw[j];
        idx = j;
      }
    // |sigma| * invmaxW > thr as the reference computes it (:117, :134);
    // maxw == 0 gives 0 * inf = NaN in the reference: the test fails and the
    // component is dropped; handled explicitly here.
    const double invmaxw = 1.0 / maxw;
    bool ok = (maxw > 0.0) && (absw[idx] * invmaxw > thr) && (max_sv > 0);
    if (ok) keep[idx] = 1.0;
    else rank_def++;
    absw[idx] = -2.0;
    max_sv--;
  }
  for (int i = 0; i < 3; ++i) D[i] = 0.0;
  for (int k = 0; k < 3; ++k) {
    if (keep[k] == 0.0) continue;
    double proj = (V[0][k] * r[0] + V[1][k] * r[1] + V[2][k] * r[2]) / lam[k];
    for (int i = 0; i < 3; ++i) D[i] += V[i][k] * proj;
  }
  return rank_def;
}

// vtkQuadricTools::ComputeRepresentativePoint (vtkQuadricTools.cxx:168-177)
inline int compute_representative_point(const double* Q, double* P, int max_sv, double thr = 1e-3) {
  double D[3];
  int rd = compute_displacement(Q, P, D, max_sv, thr);
  for (int i = 0; i < 3; ++i) P[i] += D[i];
  return rd;
}

// vtkQuadricTools::Evaluate (vtkQuadricTools.cxx:179-193)
inline double evaluate_quadric(const double* Q, const double* p, bool full) {
  double e = p[0] * p[0] * Q[0] + p[1] * p[1] * Q[4] + p[2] * p[2] * Q[7] + 2.0 * p[0] * p[1] * Q[1] +
             2.0 * p[0] * p[2] * Q[2] + 2.0 * p[1] * p[2] * Q[5] + 2.0 * p[0] * Q[3] + 2.0 * p[1] * Q[6] +
             2.0 * p[2] * Q[8];
  if (full) e += Q[9];
  return e;
}

}  // namespace oracle
