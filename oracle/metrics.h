// TEST INFRASTRUCTURE — CPU oracle, not part of the product.
//
// Restatement of the two clustering metrics on the vertex path (items =
// vertices), all in FP64:
//   - IsoMetric  = DiscreteRemeshing/vtkIsotropicMetricForClustering.h
//       Item {S = w x, W = w} (:83-87), Cluster {S[3], W, E} (:101-106),
//       E = -|S|^2 / W (:114-121), add/sub (:146-174), centroid S/W (:178-183),
//       BuildMetric vertex branch (:249-270), ClampWeights ratio 1e5 (:273-291)
//   - QEMMetric  = DiscreteRemeshing/vtkQEMetricForClustering.h
//       Item {Q[9], S = w x, W} (:95-105), Cluster (:107-129),
//       ComputeVertexQuadric (:151-167), E = W|q|^2 - 2 q.S (:194-208),
//       add/sub (:237-261), ComputeClusterCentroid (:268-285),
//       BuildMetric vertex branch (:324-342), ClampWeights ratio 1e4 (:347-362)
// The anchor (fixed vertex) machinery of the QEM metric is restated too
// (AnchorItem >= 0 pins the cluster point, -2 marks a broken constraint).
#pragma once

#include <cmath>
#include <vector>

#include "quadric.h"
#include "ref_mesh.h"

namespace oracle {

// Both metrics clamp weights to [avg/ratio, avg*ratio].
inline void clamp_weights(std::vector<double>& w, double ratio) {
  double avg = 0.0;
  for (double x : w) avg += x;
  avg /= double(w.size());
  double mn = avg / ratio, mx = avg * ratio;
  for (double& x : w) {
    if (x > mx) x = mx;
    if (x < mn) x = mn;
  }
}

// Per-vertex weights before the clamp: area (* indicator^gradation).
inline void vertex_weights(const RefMesh& m, const double* indicator, double gradation, bool qem,
                           std::vector<double>& w) {
  m.vertex_areas(w);
  if (indicator) {
    // iso: weight = area * pow(ind, gradation) whenever custom weights exist;
    // qem: only when gradation > 0.
    if (!qem || gradation > 0)
      for (int64_t i = 0; i < m.nv; ++i) w[i] *= std::pow(indicator[i], gradation);
  }
}

struct IsoMetric {
  static constexpr bool kQEM = false;
  struct Item {
    double S[3];
    double W;
  };
  struct Cluster {
    double S[3];
    double W;
    double E;
  };
  std::vector<Item> items;
  const RefMesh* mesh = nullptr;

  void set_constrained(int) {}
  void build(const RefMesh& m, const double* indicator, double gradation) {
    mesh = &m;
    std::vector<double> w;
    vertex_weights(m, indicator, gradation, false, w);
    clamp_weights(w, 100000.0);
    items.resize(m.nv);
    for (int64_t i = 0; i < m.nv; ++i) {
      for (int k = 0; k < 3; ++k) items[i].S[k] = m.P[3 * i + k] * w[i];
      items[i].W = w[i];
    }
  }
  double item_weight(int64_t i) const { return items[i].W; }
  static void reset(Cluster* c) {
    c->S[0] = c->S[1] = c->S[2] = 0;
    c->W = 0;
    c->E = 0;
  }
  static void copy(const Cluster* s, Cluster* d) { *d = *s; }
  void add_item(int64_t i, Cluster* c) const {
    for (int k = 0; k < 3; ++k) c->S[k] += items[i].S[k];
    c->W += items[i].W;
  }
  void sub_item(int64_t i, Cluster* c) const {
    for (int k = 0; k < 3; ++k) c->S[k] -= items[i].S[k];
    c->W -= items[i].W;
  }
  void compute_centroid(Cluster*) const {}
  static void compute_energy(Cluster* c) {
    c->E = (-c->S[0] * c->S[0] - c->S[1] * c->S[1] - c->S[2] * c->S[2]) / c->W;
  }
  static double energy(const Cluster* c) { return c->E; }
  static void centroid(const Cluster* c, double* p) {
    for (int k = 0; k < 3; ++k) p[k] = c->S[k] / c->W;
  }
};

struct QEMMetric {
  static constexpr bool kQEM = true;
  struct Item {
    double Q[9];
    double S[3];
    double W;
  };
  struct Cluster {
    double E;
    double Q[9];
    double C[3];  // representative point
    double S[3];
    double W;
    int rank_def = 0;
    int64_t anchor = -1;
  };
  std::vector<Item> items;
  const RefMesh* mesh = nullptr;
  int active_constraints = 1;
  int level = 3;

  void set_constrained(int c) { active_constraints = c; }
  void build(const RefMesh& m, const double* indicator, double gradation) {
    mesh = &m;
    std::vector<double> w;
    vertex_weights(m, indicator, gradation, true, w);
    clamp_weights(w, 10000.0);
    items.resize(m.nv);
    std::vector<int32_t> fl;
    for (int64_t i = 0; i < m.nv; ++i) {
      Item& it = items[i];
      for (double& q : it.Q) q = 0.0;
      m.neighbour_faces(i, fl);
      for (int32_t f : fl) {
        double x1[3], x2[3], x3[3];
        m.point(m.T[3 * f], x1);
        m.point(m.T[3 * f + 1], x2);
        m.point(m.T[3 * f + 2], x3);
        add_triangle_quadric(it.Q, x1, x2, x3, false);
      }
      for (int k = 0; k < 3; ++k) it.S[k] = m.P[3 * i + k] * w[i];
      it.W = w[i];
    }
  }
  double item_weight(int64_t i) const { return items[i].W; }
  static void reset(Cluster* c) {
    for (double& q : c->Q) q = 0;
    for (int k = 0; k < 3; ++k) c->S[k] = c->C[k] = 0;
    c->W = 0;
    c->E = 0;
  }
  static void copy(const Cluster* s, Cluster* d) { *d = *s; }
  void add_item(int64_t i, Cluster* c) const {
    const Item& it = items[i];
    for (int k = 0; k < 3; ++k) c->S[k] += it.S[k];
    for (int k = 0; k < 9; ++k) c->Q[k] += it.Q[k];
    c->W += it.W;
  }
  void sub_item(int64_t i, Cluster* c) const {
    const Item& it = items[i];
    for (int k = 0; k < 3; ++k) c->S[k] -= it.S[k];
    for (int k = 0; k < 9; ++k) c->Q[k] -= it.Q[k];
    c->W -= it.W;
    if (c->anchor == i) c->anchor = -2;
  }
  void compute_centroid(Cluster* c) const {
    if (c->anchor >= 0) {
      mesh->point(c->anchor, c->C);
      return;
    }
    for (int k = 0; k < 3; ++k) c->C[k] = c->S[k] / c->W;
    if (!active_constraints || !level) return;
    c->rank_def = compute_representative_point(c->Q, c->C, level);
  }
  static void compute_energy(Cluster* c) {
    if (c->anchor == -2) {
      c->E = 1e100;
      return;
    }
    c->E = (c->C[0] * c->C[0] + c->C[1] * c->C[1] + c->C[2] * c->C[2]) * c->W -
           2.0 * (c->C[0] * c->S[0] + c->C[1] * c->S[1] + c->C[2] * c->S[2]);
  }
  static double energy(const Cluster* c) { return c->E; }
  static void centroid(const Cluster* c, double* p) {
    for (int k = 0; k < 3; ++k) p[k] = c->C[k];
  }
};

}  // namespace oracle
