// Synthetic surface generators for the clustering workloads (host C++).
//
// These build the seeded input meshes that the benchmarks and tests run on:
//   - icospheres (optionally radially displaced by N(0, sigma) noise or by a
//     random real spherical-harmonic radius field),
//   - the cube [-1,1]^3 with an n x n quad grid per face (tangential jitter),
//   - the displaced torus on a (nu x nv) parametric grid,
// plus the discrete mean-curvature indicator |H| (cotangent Laplacian) that
// the curvature-graded workload injects as the per-vertex density, i.e. the
// array the reference passes to Metric::SetCurvatureInfo
// (DiscreteRemeshing/vtkQEMetricForClustering.h:75-83, used as
// area * indicator^gradation at :329-336).
//
// The random stream is our own (splitmix64-seeded xoshiro256**, Box-Muller
// normals in FP64) so that every host produces the same meshes.
//
// All entry points use the "query sizes, then fill caller buffers" pattern so
// that they are callable from ctypes without ownership transfer.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>
#include <algorithm>

namespace {

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s[i] = splitmix(x);
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform in [0,1)
  double uniform() { return double(next() >> 11) * 0x1.0p-53; }
  // standard normal (Box-Muller, one value per call)
  double normal() {
    double u1 = 1.0 - uniform();  // (0,1]
    double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
  // integer in [lo, hi]
  int64_t range(int64_t lo, int64_t hi) {
    uint64_t span = uint64_t(hi - lo + 1);
    return lo + int64_t(next() % span);
  }
};

// ---------------------------------------------------------------- icosphere
struct Ico {
  std::vector<double> xyz;   // unit vectors
  std::vector<int32_t> tri;
};

void build_icosphere(int level, Ico& m) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
                              {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
                              {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  const int faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                            {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                            {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                            {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
  m.xyz.clear();
  m.tri.clear();
  for (auto& p : base) {
    double n = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    m.xyz.push_back(p[0] / n);
    m.xyz.push_back(p[1] / n);
    m.xyz.push_back(p[2] / n);
  }
  for (auto& f : faces) {
    m.tri.push_back(f[0]);
    m.tri.push_back(f[1]);
    m.tri.push_back(f[2]);
  }
  for (int l = 0; l < level; ++l) {
    std::unordered_map<uint64_t, int32_t> mid;
    mid.reserve(m.tri.size());
    std::vector<int32_t> nt;
    nt.reserve(m.tri.size() * 4);
    auto midpoint = [&](int32_t a, int32_t b) -> int32_t {
      uint64_t key = a < b ? (uint64_t(a) << 32 | uint32_t(b)) : (uint64_t(b) << 32 | uint32_t(a));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      double x = m.xyz[3 * a] + m.xyz[3 * b];
      double y = m.xyz[3 * a + 1] + m.xyz[3 * b + 1];
      double z = m.xyz[3 * a + 2] + m.xyz[3 * b + 2];
      double n = std::sqrt(x * x + y * y + z * z);
      int32_t id = int32_t(m.xyz.size() / 3);
      m.xyz.push_back(x / n);
      m.xyz.push_back(y / n);
      m.xyz.push_back(z / n);
      mid.emplace(key, id);
      return id;
    };
    size_t nf = m.tri.size() / 3;
    for (size_t f = 0; f < nf; ++f) {
      int32_t a = m.tri[3 * f], b = m.tri[3 * f + 1], c = m.tri[3 * f + 2];
      int32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      const int32_t q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      nt.insert(nt.end(), q, q + 12);
    }
    m.tri.swap(nt);
  }
}

int64_t ico_nv(int level) { return 10 * (int64_t(1) << (2 * level)) + 2; }
int64_t ico_nf(int level) { return 20 * (int64_t(1) << (2 * level)); }

// ------------------------------------------------ real spherical harmonics
// Orthonormal real SH Y_lm (no Condon-Shortley phase), l <= 8.
double factorial_ratio(int l, int m) {  // (l-m)!/(l+m)!
  double r = 1.0;
  for (int k = l - m + 1; k <= l + m; ++k) r /= double(k);
  return r;
}

void legendre_all(int lmax, double x, double P[9][9]) {
  // P[l][m], m <= l, associated Legendre without the (-1)^m phase.
  double s = std::sqrt(std::max(0.0, 1.0 - x * x));
  P[0][0] = 1.0;
  for (int m = 1; m <= lmax; ++m) P[m][m] = P[m - 1][m - 1] * double(2 * m - 1) * s;
  for (int m = 0; m < lmax; ++m) P[m + 1][m] = x * double(2 * m + 1) * P[m][m];
  for (int m = 0; m <= lmax; ++m)
    for (int l = m + 2; l <= lmax; ++l)
      P[l][m] = (x * double(2 * l - 1) * P[l - 1][m] - double(l + m - 1) * P[l - 2][m]) / double(l - m);
}

// ---------------------------------------------------------------- cube grid
struct CubeFace {
  int axis;      // fixed axis
  int side;      // 0 -> coordinate -1, 1 -> +1
  int ua, va;    // in-face axes, ordered so that (u x v) points outward
};

}  // namespace

extern "C" {

// ---- icosphere, unit radius, FP64 -----------------------------------------
int acvd_gen_icosphere_sizes(int level, int64_t* nv, int64_t* nf) {
  if (level < 0 || level > 10) return -1;
  *nv = ico_nv(level);
  *nf = ico_nf(level);
  return 0;
}

// Unit icosphere radially displaced by x <- x * (1 + sigma * z_i), z_i ~ N(0,1)
// drawn in vertex order from the seeded stream (sigma = 0: exact sphere).
int acvd_gen_icosphere(int level, uint64_t seed, double sigma, double* xyz, int32_t* tri) {
  if (level < 0 || level > 10) return -1;
  Ico m;
  build_icosphere(level, m);
  Rng rng(seed);
  int64_t nv = int64_t(m.xyz.size() / 3);
  for (int64_t i = 0; i < nv; ++i) {
    double f = 1.0;
    if (sigma != 0.0) f = 1.0 + sigma * rng.normal();
    for (int k = 0; k < 3; ++k) xyz[3 * i + k] = m.xyz[3 * i + k] * f;
  }
  std::memcpy(tri, m.tri.data(), m.tri.size() * sizeof(int32_t));
  return 0;
}

// Star-shaped icosphere: radius 1 + sum_{l=1..8} sum_m a_lm Y_lm with
// a_lm ~ N(0, (amp/(l+1))^2); coefficients are redrawn while min r < rmin.
// FP32 output.  Returns the number of redraws (>= 0) or -1.
int acvd_gen_star_icosphere(int level, uint64_t seed, double amp, double rmin, float* xyz, int32_t* tri) {
  if (level < 0 || level > 10) return -1;
  Ico m;
  build_icosphere(level, m);
  Rng rng(seed);
  const int LMAX = 8;
  int64_t nv = int64_t(m.xyz.size() / 3);
  std::vector<double> r(nv);
  int redraws = 0;
  for (;;) {
    double a[LMAX + 1][2 * LMAX + 1];
    for (int l = 1; l <= LMAX; ++l)
      for (int mm = -l; mm <= l; ++mm) a[l][mm + LMAX] = rng.normal() * amp / double(l + 1);
    double rmin_seen = 1e300;
    for (int64_t i = 0; i < nv; ++i) {
      double x = m.xyz[3 * i], y = m.xyz[3 * i + 1], z = m.xyz[3 * i + 2];
      double phi = std::atan2(y, x);
      double P[9][9];
      legendre_all(LMAX, std::max(-1.0, std::min(1.0, z)), P);
      double rad = 1.0;
      for (int l = 1; l <= LMAX; ++l) {
        for (int mm = -l; mm <= l; ++mm) {
          int am = mm < 0 ? -mm : mm;
          double N = std::sqrt(double(2 * l + 1) / (4.0 * M_PI) * factorial_ratio(l, am));
          double y_lm;
          if (mm == 0) y_lm = N * P[l][0];
          else if (mm > 0) y_lm = std::sqrt(2.0) * N * P[l][am] * std::cos(am * phi);
          else y_lm = std::sqrt(2.0) * N * P[l][am] * std::sin(am * phi);
          rad += a[l][mm + LMAX] * y_lm;
        }
      }
      r[i] = rad;
      rmin_seen = std::min(rmin_seen, rad);
    }
    if (rmin_seen >= rmin) break;
    ++redraws;
    if (redraws > 1000) return -1;
  }
  for (int64_t i = 0; i < nv; ++i)
    for (int k = 0; k < 3; ++k) xyz[3 * i + k] = float(m.xyz[3 * i + k] * r[i]);
  std::memcpy(tri, m.tri.data(), m.tri.size() * sizeof(int32_t));
  return redraws;
}

// ---- cube [-1,1]^3, n x n quads per face, 2 triangles per quad ------------
int acvd_gen_cube_sizes(int n, int64_t* nv, int64_t* nf) {
  if (n < 1) return -1;
  int64_t p = n + 1;
  *nv = 6 * p * p - 12 * p + 8;
  *nf = 12 * int64_t(n) * n;
  return 0;
}

// Tangential jitter: each free (non-extreme) grid coordinate gets
// jitter * U(-1,1); coordinates at +-1 are kept, so cube edges and corners
// stay exact.  FP32 output.
int acvd_gen_cube(int n, uint64_t seed, double jitter, float* xyz, int32_t* tri) {
  if (n < 1) return -1;
  const CubeFace faces[6] = {{0, 1, 1, 2}, {0, 0, 2, 1}, {1, 1, 2, 0},
                             {1, 0, 0, 2}, {2, 1, 0, 1}, {2, 0, 1, 0}};
  // map from lattice point to vertex id for points with >= 2 extreme coords
  std::unordered_map<uint64_t, int32_t> shared;
  Rng rng(seed);
  int64_t nvtx = 0;
  auto key3 = [](int a, int b, int c) { return (uint64_t(a) << 40) | (uint64_t(b) << 20) | uint64_t(c); };
  auto emit = [&](const int g[3]) -> int32_t {
    int extreme = 0;
    for (int k = 0; k < 3; ++k) extreme += (g[k] == 0 || g[k] == n);
    if (extreme >= 2) {
      uint64_t key = key3(g[0], g[1], g[2]);
      auto it = shared.find(key);
      if (it != shared.end()) return it->second;
      shared.emplace(key, int32_t(nvtx));
    }
    for (int k = 0; k < 3; ++k) {
      double c = -1.0 + 2.0 * double(g[k]) / double(n);
      if (g[k] != 0 && g[k] != n) c += jitter * (2.0 * rng.uniform() - 1.0);
      else c = (g[k] == 0) ? -1.0 : 1.0;
      xyz[3 * nvtx + k] = float(c);
    }
    return int32_t(nvtx++);
  };
  int64_t nt = 0;
  std::vector<int32_t> ids(size_t(n + 1) * (n + 1));
  for (const CubeFace& cf : faces) {
    for (int j = 0; j <= n; ++j) {
      for (int i = 0; i <= n; ++i) {
        int g[3];
        g[cf.axis] = cf.side ? n : 0;
        g[cf.ua] = i;
        g[cf.va] = j;
        ids[size_t(j) * (n + 1) + i] = emit(g);
      }
    }
    for (int j = 0; j < n; ++j) {
      for (int i = 0; i < n; ++i) {
        int32_t p00 = ids[size_t(j) * (n + 1) + i], p10 = ids[size_t(j) * (n + 1) + i + 1];
        int32_t p11 = ids[size_t(j + 1) * (n + 1) + i + 1], p01 = ids[size_t(j + 1) * (n + 1) + i];
        const int32_t q[6] = {p00, p10, p11, p00, p11, p01};
        std::memcpy(tri + 3 * nt, q, sizeof(q));
        nt += 2;
      }
    }
  }
  int64_t ev, ef;
  acvd_gen_cube_sizes(n, &ev, &ef);
  return (nvtx == ev && nt == ef) ? 0 : -2;
}

// ---- displaced torus ------------------------------------------------------
int acvd_gen_torus_sizes(int nu, int nv, int64_t* V, int64_t* F) {
  if (nu < 3 || nv < 3) return -1;
  *V = int64_t(nu) * nv;
  *F = 2 * int64_t(nu) * nv;
  return 0;
}

// Torus: major radius R over u (nu samples), tube radius
// r(u,v) = r0 + sum_{k=1..nsin} (amp/k) sin(m_k u + n_k v + phi_k) over v (nv
// samples), m_k, n_k uniform in 1..fmax, phi_k ~ U(0, 2 pi).  Vertex (i, j)
// has id i * nv + j.  FP32 output.
int acvd_gen_torus(int nu, int nv, double R, double r0, int nsin, double amp, int fmax, uint64_t seed,
                   float* xyz, int32_t* tri) {
  if (nu < 3 || nv < 3) return -1;
  Rng rng(seed);
  std::vector<int> mk(nsin), nk(nsin);
  std::vector<double> ph(nsin);
  for (int k = 0; k < nsin; ++k) {
    mk[k] = int(rng.range(1, fmax));
    nk[k] = int(rng.range(1, fmax));
    ph[k] = 2.0 * M_PI * rng.uniform();
  }
  for (int i = 0; i < nu; ++i) {
    double u = 2.0 * M_PI * double(i) / double(nu);
    double cu = std::cos(u), su = std::sin(u);
    for (int j = 0; j < nv; ++j) {
      double v = 2.0 * M_PI * double(j) / double(nv);
      double r = r0;
      for (int k = 0; k < nsin; ++k) r += amp / double(k + 1) * std::sin(mk[k] * u + nk[k] * v + ph[k]);
      double rho = R + r * std::cos(v);
      int64_t id = int64_t(i) * nv + j;
      xyz[3 * id] = float(rho * cu);
      xyz[3 * id + 1] = float(rho * su);
      xyz[3 * id + 2] = float(r * std::sin(v));
    }
  }
  int64_t t = 0;
  for (int i = 0; i < nu; ++i) {
    int i1 = (i + 1) % nu;
    for (int j = 0; j < nv; ++j) {
      int j1 = (j + 1) % nv;
      int32_t p00 = int32_t(int64_t(i) * nv + j), p10 = int32_t(int64_t(i1) * nv + j);
      int32_t p11 = int32_t(int64_t(i1) * nv + j1), p01 = int32_t(int64_t(i) * nv + j1);
      const int32_t q[6] = {p00, p01, p11, p00, p11, p10};
      std::memcpy(tri + 3 * t, q, sizeof(q));
      t += 2;
    }
  }
  return 0;
}

// ---- discrete mean curvature magnitude ------------------------------------
// |H_i| = || sum_j (cot a_ij + cot b_ij) (x_j - x_i) || / (4 A_i), with A_i one
// third of the area of the faces around i (FP64 throughout).  Positions are
// FP32 (xyz32) or FP64 (xyz64); exactly one of them is non-null.
int acvd_mean_curvature(int64_t V, const float* xyz32, const double* xyz64, int64_t F, const int32_t* tri,
                        double* absH) {
  std::vector<double> L(3 * size_t(V), 0.0), A(size_t(V), 0.0);
  auto P = [&](int64_t v, int k) -> double { return xyz32 ? double(xyz32[3 * v + k]) : xyz64[3 * v + k]; };
  for (int64_t f = 0; f < F; ++f) {
    int32_t v[3] = {tri[3 * f], tri[3 * f + 1], tri[3 * f + 2]};
    double p[3][3];
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < 3; ++k) p[a][k] = P(v[a], k);
    double e1[3], e2[3], cr[3];
    for (int k = 0; k < 3; ++k) { e1[k] = p[1][k] - p[0][k]; e2[k] = p[2][k] - p[0][k]; }
    cr[0] = e1[1] * e2[2] - e1[2] * e2[1];
    cr[1] = e1[2] * e2[0] - e1[0] * e2[2];
    cr[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double dbl = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    if (dbl == 0.0) continue;
    for (int a = 0; a < 3; ++a) A[v[a]] += dbl / 6.0;
    // cotangent at corner a, weighting the opposite edge (b, c)
    for (int a = 0; a < 3; ++a) {
      int b = (a + 1) % 3, c = (a + 2) % 3;
      double u[3], w[3];
      for (int k = 0; k < 3; ++k) { u[k] = p[b][k] - p[a][k]; w[k] = p[c][k] - p[a][k]; }
      double dot = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
      double cot = dot / dbl;
      for (int k = 0; k < 3; ++k) {
        double d = p[c][k] - p[b][k];
        L[3 * v[b] + k] += cot * d;
        L[3 * v[c] + k] -= cot * d;
      }
    }
  }
  for (int64_t i = 0; i < V; ++i) {
    double n = std::sqrt(L[3 * i] * L[3 * i] + L[3 * i + 1] * L[3 * i + 1] + L[3 * i + 2] * L[3 * i + 2]);
    absH[i] = A[i] > 0 ? n / (4.0 * A[i]) : 0.0;
  }
  return 0;
}

// ---- seeded initial samples -------------------------------------------------
// K distinct items chosen by a partial Fisher-Yates shuffle of 0..V-1 from the
// seeded stream, labelled 0..K-1; every other item gets the null label K.
int acvd_gen_seed_labels(int64_t V, int32_t K, uint64_t seed, int32_t* labels) {
  if (K < 0 || K > V) return -1;
  std::vector<int32_t> perm(V);
  for (int64_t i = 0; i < V; ++i) perm[i] = int32_t(i);
  Rng rng(seed);
  for (int64_t i = 0; i < K; ++i) {
    int64_t j = i + int64_t(rng.next() % uint64_t(V - i));
    std::swap(perm[i], perm[j]);
  }
  for (int64_t i = 0; i < V; ++i) labels[i] = K;
  for (int64_t k = 0; k < K; ++k) labels[perm[k]] = int32_t(k);
  return 0;
}

// Draws the 384-mesh batch plan: level_j = lmin + floor((lmax-lmin+1) U).
int acvd_gen_batch_levels(int count, int lmin, int lmax, uint64_t seed, int32_t* levels) {
  Rng rng(seed);
  for (int j = 0; j < count; ++j) {
    int L = lmin + int(double(lmax - lmin + 1) * rng.uniform());
    levels[j] = std::min(L, lmax);
  }
  return 0;
}

}  // extern "C"
