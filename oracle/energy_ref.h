// TEST INFRASTRUCTURE — CPU oracle, not part of the product.
//
// FP64 reference of the fixed-assignment clustering energy and of its
// derivatives, independent of the device adjoint:
//   isotropic: E = sum_i w_i |x_i - c|^2, c = S/W per cluster
//              (vtkIsotropicMetricForClustering.h:114-121, unshifted);
//   QEM:       E = sum_i w_i |x_i - q|^2, q = c + A+_trunc (b - A c) from the
//              cluster quadric Q = sum_{i in c} sum_{f ni i} [n;d][n;d]^T,
//              n = (x2-x1) x (x3-x1), d = -n.x1
//              (vtkQEMetricForClustering.h:151-167, 194-208;
//              Common/vtkQuadricTools.cxx:56-177, threshold 1e-3, `level`
//              singular values at most).
// Derivatives come from forward-mode dual numbers pushed through the same
// arithmetic (including the Jacobi eigen-solve), one input coordinate at a
// time and only over the clusters that coordinate reaches, so they share no
// formula with the hand-derived adjoint.  Central finite differences use the
// same cluster-local energy in plain FP64.
#pragma once

#include <cmath>
#include <cstdint>
#include <algorithm>
#include <set>
#include <type_traits>
#include <vector>

namespace oracle {

struct Dual {
  double v = 0.0, d = 0.0;
  Dual() = default;
  Dual(double x) : v(x), d(0.0) {}  // NOLINT
  Dual(double x, double y) : v(x), d(y) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual dsqrt_(Dual a) {
  const double s = std::sqrt(a.v);
  return {s, s > 0 ? a.d / (2.0 * s) : 0.0};
}
inline double val(double x) { return x; }
inline double val(Dual x) { return x.v; }
inline double dsqrt_(double a) { return std::sqrt(a); }

// Symmetric 3x3 eigen-decomposition (cyclic Jacobi).  For dual numbers a
// rotation whose off-diagonal value vanishes but whose derivative does not
// takes the first-order angle t = a_pq / (a_qq - a_pp), which annihilates the
// derivative of a_pq; the eigen-pairs' derivatives then converge with the
// values.
template <class S>
void eig3(S a[3][3], S V[3][3]) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) V[r][c] = S(r == c ? 1.0 : 0.0);
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0, dia = 0;
    for (int p = 0; p < 3; ++p) {
      dia += std::fabs(val(a[p][p]));
      for (int q = p + 1; q < 3; ++q) off += std::fabs(val(a[p][q]));
    }
    bool done = off <= dia * 1e-22;
    if constexpr (!std::is_same<S, double>::value) {
      double offd = 0;
      for (int p = 0; p < 3; ++p)
        for (int q = p + 1; q < 3; ++q) offd += std::fabs(a[p][q].d);
      double diad = 0;
      for (int p = 0; p < 3; ++p) diad += std::fabs(a[p][p].d);
      done = done && offd <= (diad + dia) * 1e-22;
    }
    if (done) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        S apq = a[p][q];
        S t;
        if (std::fabs(val(apq)) <= 1e-300 + 1e-30 * std::fabs(val(a[q][q]) - val(a[p][p]))) {
          if constexpr (std::is_same<S, double>::value) {
            continue;
          } else {
            if (apq.d == 0.0 || val(a[q][q]) == val(a[p][p])) continue;
            t = apq / (a[q][q] - a[p][p]);
          }
        } else {
          S theta = (a[q][q] - a[p][p]) / (S(2.0) * apq);
          const double th = val(theta);
          S at = th < 0 ? -theta : theta;
          t = S(th >= 0 ? 1.0 : -1.0) / (at + dsqrt_(theta * theta + S(1.0)));
        }
        S cs = S(1.0) / dsqrt_(t * t + S(1.0));
        S sn = t * cs;
        for (int k = 0; k < 3; ++k) {
          S akp = a[k][p], akq = a[k][q];
          a[k][p] = cs * akp - sn * akq;
          a[k][q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < 3; ++k) {
          S apk = a[p][k], aqk = a[q][k];
          a[p][k] = cs * apk - sn * aqk;
          a[q][k] = sn * apk + cs * aqk;
        }
        for (int r = 0; r < 3; ++r) {
          S vp = V[r][p], vq = V[r][q];
          V[r][p] = cs * vp - sn * vq;
          V[r][q] = sn * vp + cs * vq;
        }
      }
  }
}

struct EnergyRef {
  int64_t nv = 0, nf = 0;
  int K = 0, metric = 0, level = 3;
  std::vector<int32_t> tri, labels;
  std::vector<std::vector<int32_t>> vf;       // vertex -> faces
  std::vector<std::vector<int32_t>> members;  // cluster -> vertices

  EnergyRef(int64_t nv_, int64_t nf_, const int32_t* tri_, const int32_t* lab, int K_, int metric_, int level_)
      : nv(nv_), nf(nf_), K(K_), metric(metric_), level(level_), tri(tri_, tri_ + 3 * nf_), labels(lab, lab + nv_),
        vf(nv_), members(K_) {
    for (int64_t f = 0; f < nf; ++f)
      for (int j = 0; j < 3; ++j) vf[tri[3 * f + j]].push_back(int32_t(f));
    for (int64_t i = 0; i < nv; ++i)
      if (labels[i] >= 0 && labels[i] < K) members[labels[i]].push_back(int32_t(i));
  }

  // energy of cluster c; coordinate `coord` (3*i+k: x_i[k], 3*nv+i: w_i) is
  // seeded with derivative 1 when S is Dual, and shifted by h when h != 0
  template <class S>
  S cluster_energy(int c, const double* x, const double* w, int64_t coord, double h) const {
    auto X = [&](int64_t i, int k) -> S {
      S s(x[3 * i + k]);
      if (coord == 3 * i + k) {
        s = s + S(h);
        if constexpr (!std::is_same<S, double>::value) s.d = 1.0;
      }
      return s;
    };
    auto Wt = [&](int64_t i) -> S {
      S s(w[i]);
      if (coord == 3 * nv + i) {
        s = s + S(h);
        if constexpr (!std::is_same<S, double>::value) s.d = 1.0;
      }
      return s;
    };
    const auto& mem = members[c];
    if (mem.empty()) return S(0.0);
    S W(0.0), Sx[3] = {S(0.0), S(0.0), S(0.0)};
    for (int32_t i : mem) {
      S wi = Wt(i);
      W = W + wi;
      for (int k = 0; k < 3; ++k) Sx[k] = Sx[k] + wi * X(i, k);
    }
    S cc[3] = {Sx[0] / W, Sx[1] / W, Sx[2] / W};
    S q[3] = {cc[0], cc[1], cc[2]};
    if (metric == 1) {
      S Q[9];
      for (int k = 0; k < 9; ++k) Q[k] = S(0.0);
      for (int32_t i : mem)
        for (int32_t f : vf[i]) {
          const int32_t i0 = tri[3 * f], i1 = tri[3 * f + 1], i2 = tri[3 * f + 2];
          S x1[3], u[3], v[3];
          for (int k = 0; k < 3; ++k) {
            x1[k] = X(i0, k);
            u[k] = X(i1, k) - x1[k];
            v[k] = X(i2, k) - x1[k];
          }
          S n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
          S d = -(n[0] * x1[0] + n[1] * x1[1] + n[2] * x1[2]);
          S p[4] = {n[0], n[1], n[2], d};
          Q[0] = Q[0] + p[0] * p[0];
          Q[1] = Q[1] + p[0] * p[1];
          Q[2] = Q[2] + p[0] * p[2];
          Q[3] = Q[3] + p[0] * p[3];
          Q[4] = Q[4] + p[1] * p[1];
          Q[5] = Q[5] + p[1] * p[2];
          Q[6] = Q[6] + p[1] * p[3];
          Q[7] = Q[7] + p[2] * p[2];
          Q[8] = Q[8] + p[2] * p[3];
        }
      S a[3][3] = {{Q[0], Q[1], Q[2]}, {Q[1], Q[4], Q[5]}, {Q[2], Q[5], Q[7]}};
      S r[3] = {-Q[3] - (Q[0] * cc[0] + Q[1] * cc[1] + Q[2] * cc[2]),
                -Q[6] - (Q[1] * cc[0] + Q[4] * cc[1] + Q[5] * cc[2]),
                -Q[8] - (Q[2] * cc[0] + Q[5] * cc[1] + Q[7] * cc[2])};
      S V[3][3];
      eig3(a, V);
      double aw[3] = {std::fabs(val(a[0][0])), std::fabs(val(a[1][1])), std::fabs(val(a[2][2]))};
      double mx = std::max(aw[0], std::max(aw[1], aw[2]));
      bool used[3] = {false, false, false};
      int left = level;
      for (int it = 0; it < 3; ++it) {
        int idx = -1;
        double best = -1;
        for (int k = 0; k < 3; ++k)
          if (!used[k] && aw[k] > best) {
            best = aw[k];
            idx = k;
          }
        if (!(mx > 0 && best * (1.0 / mx) > 1e-3 && left > 0)) break;
        used[idx] = true;
        --left;
        S vr = V[0][idx] * r[0] + V[1][idx] * r[1] + V[2][idx] * r[2];
        S f = vr / a[idx][idx];
        for (int e = 0; e < 3; ++e) q[e] = q[e] + f * V[e][idx];
      }
    }
    S E(0.0);
    for (int32_t i : mem) {
      S d2(0.0);
      for (int k = 0; k < 3; ++k) {
        S dk = X(i, k) - q[k];
        d2 = d2 + dk * dk;
      }
      E = E + Wt(i) * d2;
    }
    return E;
  }

  // clusters whose energy depends on coordinate `coord`
  std::vector<int> affected(int64_t coord) const {
    std::set<int> s;
    const int64_t i = coord < 3 * nv ? coord / 3 : coord - 3 * nv;
    if (labels[i] >= 0 && labels[i] < K) s.insert(labels[i]);
    if (metric == 1 && coord < 3 * nv)
      for (int32_t f : vf[i])
        for (int j = 0; j < 3; ++j) {
          const int32_t l = labels[tri[3 * f + j]];
          if (l >= 0 && l < K) s.insert(l);
        }
    return std::vector<int>(s.begin(), s.end());
  }

  double total(const double* x, const double* w) const {
    double e = 0;
    for (int c = 0; c < K; ++c) e += cluster_energy<double>(c, x, w, -1, 0.0);
    return e;
  }

  double derivative(const double* x, const double* w, int64_t coord) const {
    double d = 0;
    for (int c : affected(coord)) d += cluster_energy<Dual>(c, x, w, coord, 0.0).d;
    return d;
  }

  double central_difference(const double* x, const double* w, int64_t coord, double h) const {
    double ep = 0, em = 0;
    for (int c : affected(coord)) {
      ep += cluster_energy<double>(c, x, w, coord, h);
      em += cluster_energy<double>(c, x, w, coord, -h);
    }
    return (ep - em) / (2.0 * h);
  }
};

}  // namespace oracle
