// TEST INFRASTRUCTURE — CPU oracle, not part of the product.
//
// Host restatement of the device build of the sweep inputs (the kernels of
// acvd_amd/csrc/k_build.hip): the CSR adjacency with rows sorted by
// neighbour id, per-vertex FP32 items (position, clamped weight, 10-float
// quadric) and the power-of-two fixed-point scales, with the same formulas
// and the same summation order (faces of a vertex by increasing id; the
// per-mesh sums by blocks of 16384 items, 256 strided lanes and a tree).
// It lets the replay (replay.h) run without a device, and checks the device
// build.  The reference counterparts are cited in k_build.hip.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace oracle {

struct DevInputs {
  std::vector<int32_t> row_ptr, col;
  std::vector<float> xw;  // 4 per item
  std::vector<float> q;   // 10 per item (QEM)
  double scales[3] = {1, 1, 1};
};

inline double pow2_scale_below(double v) {
  if (!(v > 0.0)) return 1.0;
  int e;
  std::frexp(v, &e);
  return std::ldexp(1.0, 60 - e);
}

// per-mesh deterministic sum in the device's order
template <class F>
inline double device_order_sum(int64_t n, F value) {
  const int64_t chunk = 16384;
  double total = 0.0;
  for (int64_t s = 0; s < n; s += chunk) {
    const int64_t e = std::min(n, s + chunk);
    double sh[256];
    for (int t = 0; t < 256; ++t) {
      double v = 0.0;
      for (int64_t i = s + t; i < e; i += 256) v += value(i);
      sh[t] = v;
    }
    for (int st = 128; st > 0; st >>= 1)
      for (int t = 0; t < st; ++t) sh[t] += sh[t + st];
    total += sh[0];
  }
  return total;
}

inline void build_device_inputs(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, int metric,
                                const double* indicator, double gradation, DevInputs& out) {
  // adjacency (sorted, unique) and vertex -> faces (sorted)
  std::vector<std::vector<int32_t>> adj(nv), vf(nv);
  for (int64_t f = 0; f < nf; ++f) {
    const int32_t v[3] = {tri[3 * f], tri[3 * f + 1], tri[3 * f + 2]};
    for (int j = 0; j < 3; ++j) {
      adj[v[j]].push_back(v[(j + 1) % 3]);
      adj[v[j]].push_back(v[(j + 2) % 3]);
      vf[v[j]].push_back(int32_t(f));
    }
  }
  out.row_ptr.assign(nv + 1, 0);
  out.col.clear();
  for (int64_t i = 0; i < nv; ++i) {
    auto& r = adj[i];
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    out.row_ptr[i + 1] = out.row_ptr[i] + int32_t(r.size());
    out.col.insert(out.col.end(), r.begin(), r.end());
    std::sort(vf[i].begin(), vf[i].end());
  }
  const bool qem = metric == 1;
  std::vector<double> raw(nv);
  out.xw.assign(4 * nv, 0.0f);
  out.q.assign(10 * nv, 0.0f);  // vertex quadrics of every mesh
  for (int64_t v = 0; v < nv; ++v) {
    double area = 0.0;
    double Q[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int32_t f : vf[v]) {
      const double* p1 = &xyz[3 * int64_t(tri[3 * int64_t(f)])];
      const double* p2 = &xyz[3 * int64_t(tri[3 * int64_t(f) + 1])];
      const double* p3 = &xyz[3 * int64_t(tri[3 * int64_t(f) + 2])];
      double ea = (p1[0] - p2[0]) * (p1[0] - p2[0]) + (p1[1] - p2[1]) * (p1[1] - p2[1]) + (p1[2] - p2[2]) * (p1[2] - p2[2]);
      double eb = (p2[0] - p3[0]) * (p2[0] - p3[0]) + (p2[1] - p3[1]) * (p2[1] - p3[1]) + (p2[2] - p3[2]) * (p2[2] - p3[2]);
      double ec = (p3[0] - p1[0]) * (p3[0] - p1[0]) + (p3[1] - p1[1]) * (p3[1] - p1[1]) + (p3[2] - p1[2]) * (p3[2] - p1[2]);
      double t = 4.0 * ea * ec - (ea - eb + ec) * (ea - eb + ec);
      area += 0.25 * std::sqrt(t < 0 ? -t : t) / 3.0;
      {
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
    if (indicator && (!qem || gradation > 0.0)) wgt *= std::pow(indicator[v], gradation);
    raw[v] = wgt;
    for (int k = 0; k < 3; ++k) out.xw[4 * v + k] = float(xyz[3 * v + k]);
    for (int k = 0; k < 10; ++k) out.q[10 * v + k] = float(Q[k]);
  }
  const double avg = device_order_sum(nv, [&](int64_t i) { return raw[i]; }) / double(nv);
  const double ratio = qem ? 10000.0 : 100000.0;
  const double mn = avg / ratio, mx = avg * ratio;
  for (int64_t v = 0; v < nv; ++v) {
    double w = raw[v];
    if (w > mx) w = mx;
    if (w < mn) w = mn;
    out.xw[4 * v + 3] = float(w);
  }
  const double sw = device_order_sum(nv, [&](int64_t i) { return double(out.xw[4 * i + 3]); });
  const double ss = device_order_sum(nv, [&](int64_t i) {
    double mxc = std::max(std::fabs(double(out.xw[4 * i])),
                          std::max(std::fabs(double(out.xw[4 * i + 1])), std::fabs(double(out.xw[4 * i + 2]))));
    return double(out.xw[4 * i + 3]) * mxc;
  });
  double sq = device_order_sum(nv, [&](int64_t i) {
    double m = 0;
    for (int k = 0; k < 10; ++k) m = std::max(m, std::fabs(double(out.q[10 * i + k])));
    return m;
  });
  out.scales[0] = pow2_scale_below(sw);
  out.scales[1] = pow2_scale_below(ss);
  out.scales[2] = pow2_scale_below(sq);
}

}  // namespace oracle
