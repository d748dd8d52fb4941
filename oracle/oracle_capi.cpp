// TEST INFRASTRUCTURE — CPU oracle, not part of the product.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load liboracle.so, and only as the checker / the CPU baseline.
//
// extern "C" entry points of the oracles:
//   A  serial restatement of the reference clustering (serial.h)
//   B  threaded restatement of vtkThreadedClustering (threaded.h)
//   C  replay of the GPU schedule (replay.h)
//   E  FP64 unshifted energy of a labelling (energy below)
//   K  the reference's per-vertex curvature indicator (curvature.h)
// plus the reference's quadric tools and ring check for unit tests, and a
// host build of the device sweep inputs (devinputs.h).

#include <algorithm>
#include <chrono>
#include <random>
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "curvature.h"
#include "devinputs.h"
#include "energy_ref.h"
#include "metrics.h"
#include "quadric.h"
#include "ref_mesh.h"
#include "replay.h"
#include "serial.h"
#include "threaded.h"

using namespace oracle;

namespace {
thread_local std::string g_err;

struct OrcStats {
  int64_t loops, convergences, total_moves, evaluated;
  double seconds;
  double shifted_energy;
};

// per-loop progress lines of the serial runs on stderr (orc_set_progress)
int g_progress = 0;
// fixed clusters of the next runs and samplings (orc_set_fixed): cluster k
// holds item g_fixed[k]
std::vector<int32_t> g_fixed;
// frozen clusters of the next serial runs and replays (orc_set_frozen)
std::vector<int32_t> g_frozen;
// energy log of the serial runs (orc_set_energy_log / orc_serial_energy_log)
int g_energy_log = 0;
std::vector<long double> g_serial_log;
long double g_serial_final = 0.0L;

template <class M>
int run_serial(const RefMesh& m, int K, double gradation, const double* ind, int level, int unconstrained,
               const int32_t* init, int max_loops, int32_t* out, OrcStats* st,
               std::vector<EdgeTrace>* etrace = nullptr, size_t ecap = 0, std::vector<LoopTrace>* ltrace = nullptr) {
  Clustering<M> cl(m);
  cl.progress = g_progress;
  cl.fixed = g_fixed;
  if (g_frozen.size() == size_t(K)) cl.frozen = g_frozen;
  cl.edge_trace = etrace;
  cl.edge_trace_cap = ecap;
  cl.loop_trace = ltrace;
  g_serial_log.clear();
  if (g_energy_log) cl.energy_log = &g_serial_log;
  if constexpr (M::kQEM) cl.metric.level = level;
  cl.set_number_of_clusters(K);
  // QEM: GetCluster(k)->AnchorItem = fixed item k (Examples/ACVDQ.cxx:333-334)
  if constexpr (M::kQEM)
    for (size_t k = 0; k < g_fixed.size() && k < size_t(K); ++k) cl.clusters[k].anchor = g_fixed[k];
  cl.unconstrained_init = unconstrained;
  cl.max_loops = max_loops;
  if (init) {
    cl.initial_sampling_type = 2;
    cl.initial_labels = init;
  }
  cl.process(ind, gradation);
  g_serial_final = cl.final_energy;
  for (int64_t i = 0; i < m.nv; ++i) out[i] = cl.labels[i];
  if (st) {
    st->loops = cl.stats.loops;
    st->convergences = cl.stats.convergences;
    st->total_moves = cl.stats.total_modifications;
    st->evaluated = cl.stats.evaluated_edges;
    st->seconds = cl.stats.seconds;
    st->shifted_energy = cl.stats.shifted_energy;
  }
  return 0;
}

template <class M>
int run_threaded(const RefMesh& m, int K, double gradation, const double* ind, int level, int unconstrained,
                 const int32_t* init, int max_loops, int nthreads, int pooling, int32_t* out, OrcStats* st) {
  ThreadedClustering<M> cl(m);
  if constexpr (M::kQEM) cl.metric.level = level;
  cl.set_number_of_clusters(K);
  cl.unconstrained_init = unconstrained;
  cl.max_loops = max_loops;
  cl.num_threads = nthreads;
  cl.pooling_ratio = pooling;
  if (init) {
    cl.initial_sampling_type = 2;
    cl.initial_labels = init;
  }
  cl.process(ind, gradation);
  for (int64_t i = 0; i < m.nv; ++i) out[i] = cl.labels[i];
  if (st) {
    st->loops = cl.stats.loops;
    st->convergences = cl.stats.convergences;
    st->total_moves = cl.stats.total_modifications;
    st->evaluated = cl.stats.evaluated_edges;
    st->seconds = cl.stats.seconds;
    st->shifted_energy = cl.stats.shifted_energy;
  }
  return 0;
}

// Unshifted FP64 energy sum_i w_i |x_i - p_c|^2 with p_c the cluster point of
// the metric (centroid, or the rank-truncated quadric point at `level`),
// cluster sums taken in item order.  out[0] = unshifted energy, out[1] =
// shifted energy sum_c E_c (reference form), out[2] = sum_i w_i |x_i|^2.
template <class M>
void energy(const RefMesh& m, const double* ind, double gradation, int level, int K, const int32_t* labels,
            double* out) {
  M met;
  if constexpr (M::kQEM) met.level = level;
  met.build(m, ind, gradation);
  std::vector<typename M::Cluster> cl(K);
  for (auto& c : cl) M::reset(&c);
  if constexpr (M::kQEM)  // fixed clusters: the point of cluster k is its item's position
    for (size_t k = 0; k < g_fixed.size() && k < size_t(K); ++k) cl[k].anchor = g_fixed[k];
  for (int64_t i = 0; i < m.nv; ++i)
    if (labels[i] >= 0 && labels[i] < K) met.add_item(i, &cl[labels[i]]);
  double shifted = 0.0;
  std::vector<double> pts(3 * size_t(K));
  for (int c = 0; c < K; ++c) {
    met.compute_centroid(&cl[c]);
    M::compute_energy(&cl[c]);
    if (cl[c].W > 0) shifted += M::energy(&cl[c]);
    M::centroid(&cl[c], &pts[3 * c]);
  }
  double e = 0.0, sx = 0.0;
  for (int64_t i = 0; i < m.nv; ++i) {
    const double w = met.item_weight(i);
    const double* x = &m.P[3 * i];
    sx += w * (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    const int32_t c = labels[i];
    if (c < 0 || c >= K) continue;
    const double* p = &pts[3 * c];
    e += w * ((x[0] - p[0]) * (x[0] - p[0]) + (x[1] - p[1]) * (x[1] - p[1]) + (x[2] - p[2]) * (x[2] - p[2]));
  }
  out[0] = e;
  out[1] = shifted;
  out[2] = sx;
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void* orc_mesh_create(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri) {
  try {
    RefMesh* m = new RefMesh();
    m->build(nv, xyz, nf, tri);
    return m;
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void orc_mesh_destroy(void* m) { delete static_cast<RefMesh*>(m); }

int64_t orc_mesh_num_edges(void* mh) { return static_cast<RefMesh*>(mh)->num_edges(); }

void orc_mesh_edges(void* mh, int32_t* ev1, int32_t* ev2) {
  RefMesh* m = static_cast<RefMesh*>(mh);
  std::memcpy(ev1, m->ev1.data(), m->ev1.size() * sizeof(int32_t));
  std::memcpy(ev2, m->ev2.data(), m->ev2.size() * sizeof(int32_t));
}

// K: curvature indicator of every vertex (curvature.h).
void orc_curvature_indicator(void* mh, double* out) {
  std::vector<double> k;
  curvature_indicator(*static_cast<RefMesh*>(mh), k);
  std::memcpy(out, k.data(), k.size() * sizeof(double));
}

// Face list of vertex v for the curvature fit; returns its length (at most
// cap entries are written).
int64_t orc_curvature_face_list(void* mh, int64_t v, int32_t* out, int64_t cap) {
  std::vector<int32_t> fl;
  curvature_face_list(*static_cast<RefMesh*>(mh), v, fl);
  for (int64_t k = 0; k < int64_t(fl.size()) && k < cap; ++k) out[k] = fl[k];
  return int64_t(fl.size());
}

void orc_mesh_vertex_areas(void* mh, double* out) {
  std::vector<double> a;
  static_cast<RefMesh*>(mh)->vertex_areas(a);
  std::memcpy(out, a.data(), a.size() * sizeof(double));
}

// Items of the metric (BuildMetric): weights w[nv]; for QEM also Q[9*nv].
int orc_build_items(void* mh, int metric, const double* ind, double gradation, double* w, double* q9) {
  const RefMesh& m = *static_cast<RefMesh*>(mh);
  if (metric == 0) {
    IsoMetric met;
    met.build(m, ind, gradation);
    for (int64_t i = 0; i < m.nv; ++i) w[i] = met.items[i].W;
  } else {
    QEMMetric met;
    met.build(m, ind, gradation);
    for (int64_t i = 0; i < m.nv; ++i) {
      w[i] = met.items[i].W;
      if (q9)
        for (int k = 0; k < 9; ++k) q9[9 * i + k] = met.items[i].Q[k];
    }
  }
  return 0;
}

int orc_serial_run(void* mh, int metric, int K, double gradation, const double* ind, int level, int unconstrained,
                   const int32_t* init, int max_loops, int32_t* out, OrcStats* st) {
  try {
    const RefMesh& m = *static_cast<RefMesh*>(mh);
    if (metric == 0) return run_serial<IsoMetric>(m, K, gradation, ind, level, unconstrained, init, max_loops, out, st);
    return run_serial<QEMMetric>(m, K, gradation, ind, level, unconstrained, init, max_loops, out, st);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

void orc_set_progress(int on) { g_progress = on; }

void orc_set_fixed(const int32_t* ids, int64_t n) { g_fixed.assign(ids, ids + (ids ? n : 0)); }

void orc_set_energy_log(int on) { g_energy_log = on; }

void orc_set_frozen(const int32_t* flags, int64_t n) { g_frozen.assign(flags, flags + (flags ? n : 0)); }

// The energy log of the last serial run (long double values as double).
int64_t orc_serial_energy_log(int64_t cap, double* energy, double* final_energy) {
  const int64_t n = int64_t(g_serial_log.size());
  for (int64_t i = 0; i < n && i < cap; ++i) energy[i] = double(g_serial_log[size_t(i)]);
  if (final_energy) *final_energy = double(g_serial_final);
  return n;
}

// orc_serial_run with its trace: edge records (loop, decision, edge, i1, i2 in
// edge_i[5k..]; try1..try3 in edge_d[3k..]) up to edge_cap, loop records
// (loop, mods, connexity, constrained, convergence, disconnected in
// loop_i[6k..]) up to loop_cap; n_out = {edge records, loop records}.
int orc_serial_trace(void* mh, int metric, int K, double gradation, const double* ind, int level, int unconstrained,
                     const int32_t* init, int max_loops, int32_t* out, OrcStats* st, int64_t edge_cap,
                     int64_t* edge_i, double* edge_d, int64_t loop_cap, int32_t* loop_i, int64_t* n_out) {
  try {
    const RefMesh& m = *static_cast<RefMesh*>(mh);
    std::vector<EdgeTrace> et;
    std::vector<LoopTrace> lt;
    int rc = metric == 0 ? run_serial<IsoMetric>(m, K, gradation, ind, level, unconstrained, init, max_loops, out, st,
                                                 &et, size_t(edge_cap), &lt)
                         : run_serial<QEMMetric>(m, K, gradation, ind, level, unconstrained, init, max_loops, out,
                                                 st, &et, size_t(edge_cap), &lt);
    for (size_t k = 0; k < et.size(); ++k) {
      const EdgeTrace& r = et[k];
      const int64_t v[5] = {r.loop, r.decision, r.edge, r.i1, r.i2};
      for (int j = 0; j < 5; ++j) edge_i[5 * k + j] = v[j];
      edge_d[3 * k] = r.try1;
      edge_d[3 * k + 1] = r.try2;
      edge_d[3 * k + 2] = r.try3;
    }
    const size_t nl = std::min(lt.size(), size_t(loop_cap));
    for (size_t k = 0; k < nl; ++k) {
      const LoopTrace& r = lt[k];
      const int32_t v[6] = {r.loop, r.mods, r.connexity, r.constrained, r.convergence, r.disconnected};
      for (int j = 0; j < 6; ++j) loop_i[6 * k + j] = v[j];
    }
    n_out[0] = int64_t(et.size());
    n_out[1] = int64_t(nl);
    return rc;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int orc_threaded_run(void* mh, int metric, int K, double gradation, const double* ind, int level, int unconstrained,
                     const int32_t* init, int max_loops, int nthreads, int pooling, int32_t* out, OrcStats* st) {
  try {
    const RefMesh& m = *static_cast<RefMesh*>(mh);
    if (metric == 0)
      return run_threaded<IsoMetric>(m, K, gradation, ind, level, unconstrained, init, max_loops, nthreads, pooling,
                                     out, st);
    return run_threaded<QEMMetric>(m, K, gradation, ind, level, unconstrained, init, max_loops, nthreads, pooling,
                                   out, st);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Reference initial sampling (type 1, ComputeInitialRandomSampling).
int orc_initial_sampling(void* mh, int metric, int K, double gradation, const double* ind, int32_t* out) {
  const RefMesh& m = *static_cast<RefMesh*>(mh);
  std::vector<int32_t> lab;
  if (metric == 0) {
    Clustering<IsoMetric> cl(m);
    cl.set_number_of_clusters(K);
    cl.fixed = g_fixed;
    cl.init(ind, gradation);
    cl.initial_random_sampling(lab, K);
  } else {
    Clustering<QEMMetric> cl(m);
    cl.set_number_of_clusters(K);
    cl.fixed = g_fixed;
    cl.init(ind, gradation);
    cl.initial_random_sampling(lab, K);
  }
  std::memcpy(out, lab.data(), lab.size() * sizeof(int32_t));
  return 0;
}

// InitSamples type 0 (vtkUniformClustering.h:1126-1159) on a mesh: labels
// 0..K-1, K = null.
int orc_seed_sampling(void* mh, int K, int32_t* out) {
  const RefMesh& m = *static_cast<RefMesh*>(mh);
  Clustering<IsoMetric> cl(m);
  cl.set_number_of_clusters(K);
  cl.allocate();
  cl.initial_sampling_type = 0;
  cl.init_samples();
  for (int64_t i = 0; i < m.nv; ++i) out[i] = cl.labels[i];
  return 0;
}

// Reference FillHolesInClustering (connexity off) on a labelling (in place).
int orc_fill_holes(void* mh, int K, int32_t* labels) {
  const RefMesh& m = *static_cast<RefMesh*>(mh);
  Clustering<IsoMetric> cl(m);
  cl.set_number_of_clusters(K);
  cl.allocate();
  std::vector<int32_t> lab(labels, labels + m.nv);
  cl.labels = lab;
  cl.fill_holes();
  std::memcpy(labels, cl.labels.data(), m.nv * sizeof(int32_t));
  return 0;
}

// Reference CleanClustering on a labelling (in place); returns the count.
int orc_clean(void* mh, int K, int32_t* labels) {
  const RefMesh& m = *static_cast<RefMesh*>(mh);
  Clustering<IsoMetric> cl(m);
  cl.set_number_of_clusters(K);
  cl.fixed = g_fixed;
  cl.allocate();
  cl.labels.assign(labels, labels + m.nv);
  int n = cl.clean_clustering();
  std::memcpy(labels, cl.labels.data(), m.nv * sizeof(int32_t));
  return n;
}

int orc_energy(void* mh, int metric, const double* ind, double gradation, int level, int K, const int32_t* labels,
               double* out3) {
  const RefMesh& m = *static_cast<RefMesh*>(mh);
  if (metric == 0) energy<IsoMetric>(m, ind, gradation, level, K, labels, out3);
  else energy<QEMMetric>(m, ind, gradation, level, K, labels, out3);
  return 0;
}

// vtkQuadricTools::ComputeRepresentativePoint on (Q[9], P[3]) -> P, rank def.
int orc_rep_point(const double* Q9, double* P, int level, double thr) {
  return compute_representative_point(Q9, P, level, thr);
}

void orc_sym_eigen(const double* A9, double* lam3, double* V9) {
  double A[3][3], V[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = A9[3 * i + j];
  sym_eigen3(A, lam3, V);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V9[3 * i + j] = V[i][j];
}

void orc_triangle_quadric(const double* x1, const double* x2, const double* x3, double* Q10) {
  for (int k = 0; k < 10; ++k) Q10[k] = 0.0;
  add_triangle_quadric(Q10, x1, x2, x3, true);
}

// Reference ring check (ConnexityConstraintProblemLocal) on the oracle mesh.
int orc_ring_problem(void* mh, const int32_t* labels, int64_t item, int32_t cluster) {
  const RefMesh& m = *static_cast<RefMesh*>(mh);
  Clustering<IsoMetric> cl(m);
  cl.set_number_of_clusters(1);
  cl.labels.assign(labels, labels + m.nv);
  cl.connexity = 1;
  return cl.connexity_problem(item, cluster);
}

// Shared-header functions evaluated on the host (compared with the oracle).
// V0 (nullable): eigenvector columns to start the Jacobi solve from; Vout
// (nullable) receives the eigenvector columns of the solve.
int orc_hdr_rep_point_delta(const double* Q9, const double* c, int level, double* D, const double* V0,
                            double* Vout) {
  return acvd::rep_point_delta(Q9, c, level, D, V0, Vout);
}

int orc_hdr_ring_problem(const int32_t* rp, const int32_t* col, const int32_t* labels, int32_t item,
                         int32_t cluster) {
  return acvd::ring_problem(rp, col, labels, item, cluster) ? 1 : 0;
}

// Ring check through the ring adjacency mask (-1 for rows longer than the
// mask covers).
int orc_hdr_ring_problem_mask(const int32_t* rp, const int32_t* col, const int32_t* labels, int32_t item,
                              int32_t cluster) {
  const int32_t r0 = rp[item], deg = rp[item + 1] - r0;
  if (deg > acvd::kRingAdjMax) return -1;
  int32_t nb[acvd::kRingAdjMax];
  for (int p = 0; p < deg; ++p) nb[p] = labels[col[r0 + p]];
  return acvd::ring_problem_mask(acvd::ring_adjacency(rp, col, item), nb, deg, cluster) ? 1 : 0;
}

uint64_t orc_hdr_accept_key(uint32_t seed, int32_t loop, int round, uint32_t item) {
  return acvd::accept_key(seed, loop, round, item);
}

// Hartigan terms of the shared header: out2 = {leave term w.r.t. (ca, Wa),
// join term w.r.t. (cb, Wb)}.
void orc_hdr_terms(const float* xw4, const double* ca, double Wa, const double* cb, double Wb, float* out2) {
  acvd::ClusterGeom a{ca[0], ca[1], ca[2], Wa}, b{cb[0], cb[1], cb[2], Wb};
  out2[0] = acvd::leave_term(xw4[0], xw4[1], xw4[2], xw4[3], a);
  out2[1] = acvd::join_term(xw4[0], xw4[1], xw4[2], xw4[3], b);
}

// T = W |q - c|^2 and q of a cluster from fixed accumulators (shared header).
double orc_hdr_qem_T(const int64_t* ws4, const int64_t* q10, const double* scales3, int level, double* q_out) {
  acvd::Scales sc{scales3[0], scales3[1], scales3[2], 1.0 / scales3[0], 1.0 / scales3[1], 1.0 / scales3[2]};
  acvd::ClusterGeom g = acvd::geom_from_fixed(ws4[0], ws4[1], ws4[2], ws4[3], sc);
  return acvd::qem_T_from_fixed(q10, g, sc, level, q_out);
}

// Fixed-point image of one item (shared header): out = {w, wx, wy, wz, q0..q9}.
void orc_hdr_fix_item(const float* xw4, const float* q10, const double* scales3, int64_t* out14) {
  acvd::Scales sc{scales3[0], scales3[1], scales3[2], 1.0 / scales3[0], 1.0 / scales3[1], 1.0 / scales3[2]};
  out14[0] = acvd::fix_w(xw4[3], sc);
  for (int k = 0; k < 3; ++k) out14[1 + k] = acvd::fix_s(xw4[3], xw4[k], sc);
  for (int k = 0; k < 10; ++k) out14[4 + k] = q10 ? acvd::fix_q(q10[k], sc) : 0;
}

// The C++ standard fixes the 10000th output of a default-seeded std::mt19937
// (4123659995); the reference's initial sampling draws from this engine.
// Park and Miller's published check of the minimal standard generator: from
// x = 1, the 10000th value is 1043618065 (the oracle's VtkMinStd step).
int64_t orc_minstd_10000(void) {
  VtkMinStd g;
  g.x = 1;
  for (int i = 0; i < 10000; ++i) g.step();
  return g.x;
}

uint32_t orc_mt19937_10000(void) {
  std::mt19937 g;
  uint32_t v = 0;
  for (int i = 0; i < 10000; ++i) v = g();
  return v;
}

// Replay C.  params: K, metric, level, seed, unconstrained, max_loops,
// max_convergences.  scales3 = {sw, ss, sq}.  init/out in the reference
// convention (K = null).
// Energy log of the last replay run with path bit 16 set.
static std::vector<double> g_replay_log;
static double g_replay_final = 0.0;

int64_t orc_replay_energy_log(int64_t cap, double* energy, double* final_energy) {
  const int64_t n = int64_t(g_replay_log.size());
  for (int64_t i = 0; i < n && i < cap; ++i) energy[i] = g_replay_log[size_t(i)];
  if (final_energy) *final_energy = g_replay_final;
  return n;
}

// path: bit 0 generic QEM rule, bit 1 every round (no end after an empty round),
// bit 16 record the energy log
// moves of the last replay to a cluster none of the item's neighbours held
static int64_t g_replay_nonadj = 0;
int64_t orc_replay_nonadjacent() { return g_replay_nonadj; }

int orc_replay_run_path(int path, int64_t n, const int32_t* rp, const int32_t* col, const float* xw,
                        const float* q, const double* scales3, int K, int metric, int level, uint32_t seed,
                        int unconstrained, int max_loops, int max_conv, const int32_t* init, int32_t* out,
                        int64_t* stats6) {
  Replay r;
  if (!g_fixed.empty()) {
    r.anchor.assign(size_t(K), -1);
    for (size_t k = 0; k < g_fixed.size() && k < size_t(K); ++k) r.anchor[k] = g_fixed[k];
  }
  if (g_frozen.size() == size_t(K)) r.frozen = g_frozen;
  r.generic_qem = (path & 1) != 0;
  r.all_rounds = (path & 2) != 0;
  // bits 8..15: accept rounds per pass before the first convergence (0: default)
  if ((path >> 8) & 0xFF) r.first_rounds = std::min((path >> 8) & 0xFF, acvd::kMatchRounds);
  r.N = n;
  r.K = K;
  r.metric = metric;
  r.level = level;
  r.seed = seed;
  r.unconstrained_init = unconstrained;
  r.max_loops = max_loops;
  r.max_conv = max_conv;
  r.rp = rp;
  r.col = col;
  r.xw = xw;
  r.q = q;
  r.sc.sw = scales3[0];
  r.sc.ss = scales3[1];
  r.sc.sq = scales3[2];
  r.sc.isw = 1.0 / r.sc.sw;
  r.sc.iss = 1.0 / r.sc.ss;
  r.sc.isq = 1.0 / r.sc.sq;
  r.log_energy = (path >> 16) & 1;
  r.run(init);
  g_replay_log = r.energy_log;
  g_replay_final = r.final_energy;
  g_replay_nonadj = r.stats.non_adjacent_moves;
  for (int64_t i = 0; i < n; ++i) out[i] = r.labels[i] == acvd::kNull ? K : r.labels[i];
  if (stats6) {
    stats6[0] = r.stats.loops;
    stats6[1] = r.stats.convergences;
    stats6[2] = r.stats.total_moves;
    stats6[3] = r.stats.evaluated;
    stats6[4] = r.stats.proposed;
    stats6[5] = r.stats.last_disconnected;
  }
  return 0;
}

int orc_replay_run(int64_t n, const int32_t* rp, const int32_t* col, const float* xw, const float* q,
                   const double* scales3, int K, int metric, int level, uint32_t seed, int unconstrained,
                   int max_loops, int max_conv, const int32_t* init, int32_t* out, int64_t* stats6) {
  return orc_replay_run_path(0, n, rp, col, xw, q, scales3, K, metric, level, seed, unconstrained, max_loops, max_conv,
                             init, out, stats6);
}

// Host build of the device sweep inputs (devinputs.h).
void* orc_devinputs_create(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, int metric,
                           const double* ind, double gradation) {
  DevInputs* d = new DevInputs();
  build_device_inputs(nv, xyz, nf, tri, metric, ind, gradation, *d);
  return d;
}
int64_t orc_devinputs_ncol(void* h) { return int64_t(static_cast<DevInputs*>(h)->col.size()); }
void orc_devinputs_get(void* h, int32_t* rp, int32_t* col, float* xw, float* q, double* scales3) {
  DevInputs* d = static_cast<DevInputs*>(h);
  std::memcpy(rp, d->row_ptr.data(), d->row_ptr.size() * sizeof(int32_t));
  std::memcpy(col, d->col.data(), d->col.size() * sizeof(int32_t));
  std::memcpy(xw, d->xw.data(), d->xw.size() * sizeof(float));
  if (q && !d->q.empty()) std::memcpy(q, d->q.data(), d->q.size() * sizeof(float));
  for (int k = 0; k < 3; ++k) scales3[k] = d->scales[k];
}
void orc_devinputs_destroy(void* h) { delete static_cast<DevInputs*>(h); }

}  // extern "C"

// ---- fixed-assignment energy reference
extern "C" void* orc_energy_create(int64_t nv, int64_t nf, const int32_t* tri, const int32_t* labels, int K,
                                   int metric, int level) {
  return new oracle::EnergyRef(nv, nf, tri, labels, K, metric, level);
}
extern "C" void orc_energy_destroy(void* h) { delete static_cast<oracle::EnergyRef*>(h); }
extern "C" double orc_energy_total(void* h, const double* x, const double* w) {
  return static_cast<oracle::EnergyRef*>(h)->total(x, w);
}
// coords: 3*i+k for x_i[k], 3*nv+i for w_i
extern "C" void orc_energy_derivatives(void* h, const double* x, const double* w, int64_t n, const int64_t* coords,
                                       double* out) {
  auto* e = static_cast<oracle::EnergyRef*>(h);
  for (int64_t k = 0; k < n; ++k) out[k] = e->derivative(x, w, coords[k]);
}
extern "C" void orc_energy_central_differences(void* h, const double* x, const double* w, int64_t n,
                                               const int64_t* coords, double step, double* out) {
  auto* e = static_cast<oracle::EnergyRef*>(h);
  for (int64_t k = 0; k < n; ++k) out[k] = e->central_difference(x, w, coords[k], step);
}
