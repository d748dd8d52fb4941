// Drop-in check of include/acvd/remeshing.hpp: the calls of
// DiscreteRemeshing/Examples/ACVD.cxx:208-215 and ACVDQ.cxx:319-338 on an
// icosphere built here.  Prints one line "ok nv nf labels_null" on success,
// "fail <reason>" and exit status 1 on failure.
#include <cmath>
#include <cstdio>
#include <map>
#include <utility>

#include "acvd/remeshing.hpp"

static acvd::Mesh icosphere(int level) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  acvd::Mesh m;
  const double v[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                           {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  const int f[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                        {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                        {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  for (auto& p : v) {
    double n = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    for (int k = 0; k < 3; ++k) m.points.push_back(p[k] / n);
  }
  for (auto& t3 : f)
    for (int k = 0; k < 3; ++k) m.triangles.push_back(t3[k]);
  for (int l = 0; l < level; ++l) {
    std::map<std::pair<int, int>, int> mid;
    std::vector<int32_t> nt;
    auto midpoint = [&](int a, int b) {
      auto key = std::make_pair(a < b ? a : b, a < b ? b : a);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      double p[3], n = 0;
      for (int k = 0; k < 3; ++k) {
        p[k] = 0.5 * (m.points[3 * a + k] + m.points[3 * b + k]);
        n += p[k] * p[k];
      }
      n = std::sqrt(n);
      int id = int(m.points.size() / 3);
      for (int k = 0; k < 3; ++k) m.points.push_back(p[k] / n);
      mid[key] = id;
      return id;
    };
    for (size_t i = 0; i < m.triangles.size(); i += 3) {
      int a = m.triangles[i], b = m.triangles[i + 1], c = m.triangles[i + 2];
      int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      int32_t add[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      nt.insert(nt.end(), add, add + 12);
    }
    m.triangles.swap(nt);
  }
  return m;
}

int main(int argc, char** argv) {
  const bool qem = argc > 1 && argv[1][0] == 'q';
  acvd::Mesh mesh = icosphere(4);
  const int K = 200;
  const acvd::Mesh* out = nullptr;
  const acvd::Labels* labels = nullptr;
  std::string err;
  if (qem) {
    auto* r = acvd::QIsotropicDiscreteRemeshing::New();
    r->SetInput(&mesh);
    r->SetNumberOfClusters(K);
    r->SetConsoleOutput(0);
    r->SetUnconstrainedInitialization(1);
    r->GetMetric()->SetQuadricsOptimizationLevel(3);
    if (r->Remesh()) {
      out = r->GetOutput();
      labels = r->GetClustering();
      acvd::Mesh copy = *out;
      int64_t nnull = 0;
      for (int32_t l : *labels) nnull += l == K;
      std::printf("ok %lld %lld %lld\n", (long long)copy.GetNumberOfPoints(), (long long)copy.GetNumberOfCells(),
                  (long long)nnull);
      r->Delete();
      return 0;
    }
    err = r->GetLastError();
    r->Delete();
  } else {
    auto* r = acvd::IsotropicDiscreteRemeshing::New();
    r->SetInput(&mesh);
    r->SetNumberOfClusters(K);
    r->SetConsoleOutput(0);
    r->GetMetric()->SetGradation(0.0);
    if (r->Remesh() && r->BuildOutput(1)) {
      out = r->GetOutput();
      labels = r->GetClustering();
      int64_t nnull = 0;
      for (int32_t l : *labels) nnull += l == K;
      std::printf("ok %lld %lld %lld\n", (long long)out->GetNumberOfPoints(), (long long)out->GetNumberOfCells(),
                  (long long)nnull);
      r->Delete();
      return 0;
    }
    err = r->GetLastError();
    r->Delete();
  }
  std::printf("fail %s\n", err.c_str());
  return 1;
}
