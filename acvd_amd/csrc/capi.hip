// C-ABI implementation: contexts, device build, and the MinimizeEnergy
// driver (vtkUniformClustering.h:654-830) over the sweep kernels.
//
// Host work per ProcessClustering is limited to launching kernels and, between
// chunks of passes, reading back a small control block (and the per-mesh
// states when a mesh reached a convergence).  The chunk of passes
// (propose -> apply -> controller) x kChunk (kShortChunk near the end of a
// phase) is captured once into a hipGraph and replayed; all pass-dependent
// values live in device memory.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/acvd_c.h"
#include "acvd_ctx.h"
#include "host_mesh.h"
#include "host_sampling.h"
#include "host_parallel.h"

using namespace acvd;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return fail(ACVD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));            \
  } while (0)

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
    if (e == hipSuccess) n = count;
    else p = nullptr;
    return e;
  }
};

struct HostMesh {
  int64_t nv = 0, nf = 0;
  std::vector<double> xyz;
  std::vector<int32_t> tri;
  std::vector<double> indicator;  // set by the caller (SetCurvatureInfo)
  std::vector<double> curvature;  // computed by the last build (curvature_indicator)
  std::vector<int32_t> init_labels;
  std::vector<int32_t> fixed;  // SetFixedClusters: cluster k < fixed.size() holds item fixed[k]
  std::vector<int32_t> frozen;  // IsClusterFreezed per cluster (ForceManifold restarts); empty: none
  acvd_params params;
  // layout in the packed context
  int64_t item_off = 0, clus_off = 0, face_off = 0, n_adj = 0, adj_off = 0;
  // output mesh (BuildDelaunayTriangulation)
  std::vector<double> out_xyz;
  std::vector<int32_t> out_tri;
  bool clustered = false;
  acvd_stats stats;
  // energy log of the last run (write_energy_log): per loop, seconds since the
  // start of the minimisation and the energy; the energy after the run
  std::vector<double> log_seconds, log_energy;
  double log_final = 0.0;
  bool logged = false;
};

constexpr int kChunk = 16;        // passes per graph launch
constexpr int kShortChunk = 4;    // passes per graph launch near the end of a phase
static_assert(kPassLogCap > kChunk, "the pass log must hold a chunk of passes");
constexpr int kReduceItems = 16384;  // items per block of the deterministic reductions
constexpr int kEnergyLogRing = 2 * kChunk;  // energy log entries per mesh on the device, drained per chunk

}  // namespace

struct acvd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<HostMesh> meshes;
  bool built = false;
  bool any_iso = false, any_qem = false;
  int64_t N = 0, K = 0, F = 0, E2 = 0;
  // device arrays
  DBuf<DevMesh> d_meshes;
  DBuf<MeshState> d_state;
  DBuf<Ctrl> d_ctrl;
  DBuf<int32_t> d_item_mesh, d_clus_mesh;
  DBuf<int32_t> d_row_ptr, d_col, d_tri;
  // kept from the build for the reference's initial sampling: FP64 positions,
  // vertex -> face rows (faces in increasing id)
  DBuf<double> d_pos;
  DBuf<int32_t> d_vfptr, d_vfcol;
  DBuf<float4> d_xw;
  DBuf<float> d_q;
  DBuf<int32_t> d_labels, d_sizes, d_stamp, d_cand_pass, d_pending;
  DBuf<unsigned long long> d_acc_ws, d_acc_q, d_prio, d_ring_adj;
  DBuf<ClusterGeom> d_geom;
  DBuf<int32_t> d_anchor;  // per cluster: global id of its fixed item or -1
  bool any_fixed = false;
  DBuf<int32_t> d_frozen;  // per cluster: nonzero when frozen
  bool any_frozen = false;
  DBuf<double> d_T, d_evec;
  DBuf<int32_t> d_cand_tgt, d_cand_lab, d_cand_since, d_cl_round;
  DBuf<int32_t> d_work, d_work_count, d_cand, d_seg_cand, d_moved, d_seg_moved;
  int32_t work_cap = 0, n_seg = 0, cand_cap = 0;
  // launch-shape knobs (acvd_set_tuning); none changes the labels
  acvd_tuning tuning = {kSweepSegs, kTileOnDivisor, kChunk, kShortChunk};
  // scan tiles
  int32_t n_tiles = 0;
  DBuf<int32_t> d_tile_items, d_item_tile, d_tile_mesh, d_tile_cnt, d_tile_set, d_tile_prop;
  DBuf<int32_t> d_mask_all, d_mask, d_ndisc;
  DBuf<unsigned long long> d_mesh_count;
  // clean scratch
  DBuf<int32_t> d_parent, d_comp_size, d_ncomp;
  DBuf<unsigned long long> d_best;
  // energy log (meshes with write_energy_log): cluster energies, a ring of
  // kEnergyLogRing entries per mesh, the start clock, the final energies
  bool energy_log = false;
  DBuf<double> d_log_cluster, d_log_final;
  DBuf<EnergyLogEntry> d_log_ring;
  DBuf<unsigned long long> d_log_clock0;
  // pinned host mirrors
  Ctrl* h_ctrl = nullptr;
  MeshState* h_state = nullptr;
  std::vector<DevMesh> h_meshes;
  void* cub_tmp = nullptr;
  size_t cub_bytes = 0;
  // pass-chunk graphs by (QEM meshes unconstrained, tile sets maintained,
  // short chunk): index qem_unc + 2 * tiles + 4 * short
  // captured chunks by (QEM unconstrained, tile sets on, short, rounds - 1)
  static constexpr int kChunkKinds = 8 * kMatchRounds;
  hipGraphExec_t chunk_exec[kChunkKinds] = {};
  hipGraph_t chunk_graph[kChunkKinds] = {};
  double build_ms = 0.0, converge_ms = 0.0;
  Ctrl last_ctrl{};
  bool timing = false;
  bool count_rounds = false;  // timing mode 1: the accept-round counters as well
  bool labels_seeded = false;  // device labels hold the initial sampling
  bool keep_frozen = false;    // the ForceManifold restarts keep the frozen flags
  double kernel_ms[4] = {0, 0, 0, 0};
  int64_t kernel_launches[4] = {0, 0, 0, 0};

  ~acvd_ctx() {
    drop_graphs();
    if (cub_tmp) (void)hipFree(cub_tmp);
    if (h_ctrl) (void)hipHostFree(h_ctrl);
    if (h_state) (void)hipHostFree(h_state);
    if (stream) (void)hipStreamDestroy(stream);
  }

  SweepArgs args() const {
    SweepArgs a{};
    a.n_items = N;
    a.n_clusters = K;
    a.n_meshes = int32_t(meshes.size());
    a.n_tiles = n_tiles;
    a.meshes = d_meshes.p;
    a.state = d_state.p;
    a.item_mesh = meshes.size() > 1 ? d_item_mesh.p : nullptr;
    a.clus_mesh = meshes.size() > 1 ? d_clus_mesh.p : nullptr;
    a.row_ptr = d_row_ptr.p;
    a.col = d_col.p;
    a.ring_adj = d_ring_adj.p;
    a.xw = d_xw.p;
    a.q = d_q.p;
    a.labels = d_labels.p;
    a.acc_ws = d_acc_ws.p;
    a.acc_q = any_qem ? d_acc_q.p : nullptr;
    a.geom = d_geom.p;
    a.anchor = any_fixed ? d_anchor.p : nullptr;
    a.frozen = any_frozen ? d_frozen.p : nullptr;
    a.count_rounds = count_rounds ? 1 : 0;
    a.T = any_qem ? d_T.p : nullptr;
    a.evec = any_qem ? d_evec.p : nullptr;
    a.sizes = d_sizes.p;
    a.stamp = d_stamp.p;
    a.prio = d_prio.p;
    a.cand_pass = d_cand_pass.p;
    a.cand_tgt = d_cand_tgt.p;
    a.cand_lab = d_cand_lab.p;
    a.cand_since = d_cand_since.p;
    a.cl_round = d_cl_round.p;
    a.work = d_work.p;
    a.work_count = d_work_count.p;
    a.work_cap = work_cap;
    a.n_seg = n_seg;
    a.cand_cap = cand_cap;
    a.seg_cand = d_seg_cand.p;
    a.moved = d_moved.p;
    a.seg_moved = d_seg_moved.p;
    a.cand = d_cand.p;
    a.ctrl = d_ctrl.p;
    a.tile_items = d_tile_items.p;
    a.item_tile = d_item_tile.p;
    a.tile_mesh = d_tile_mesh.p;
    a.tile_cnt = d_tile_cnt.p;
    a.tile_set = d_tile_set.p;
    a.tile_prop = d_tile_prop.p;
    return a;
  }

  void drop_graphs() {
    for (int k = 0; k < kChunkKinds; ++k) {
      if (chunk_exec[k]) (void)hipGraphExecDestroy(chunk_exec[k]);
      if (chunk_graph[k]) (void)hipGraphDestroy(chunk_graph[k]);
      chunk_exec[k] = nullptr;
      chunk_graph[k] = nullptr;
    }
  }

  void invalidate() {
    built = false;
    drop_graphs();
  }
};

namespace {

int check_ctx(const acvd_ctx* c) {
  if (!c) return fail(ACVD_ERR_ARG, "null context");
  return ACVD_OK;
}

int check_mesh(const acvd_ctx* c, int32_t m) {
  if (!c) return fail(ACVD_ERR_ARG, "null context");
  if (m < 0 || m >= int32_t(c->meshes.size())) return fail(ACVD_ERR_ARG, "mesh id out of range");
  return ACVD_OK;
}

template <class P>
int add_mesh(acvd_ctx* c, int64_t nv, const P* xyz, int64_t nf, const int32_t* tri, int32_t* mesh_id) {
  if (!c) return fail(ACVD_ERR_ARG, "null context");
  if (nv <= 0 || nf <= 0 || !xyz || !tri) return fail(ACVD_ERR_ARG, "empty mesh or null buffer");
  if (nv >= kMaxItemsPerMesh) return fail(ACVD_ERR_LIMIT, "mesh has more than 2^27 vertices");
  for (int64_t k = 0; k < 3 * nf; ++k)
    if (tri[k] < 0 || tri[k] >= nv)
      return fail(ACVD_ERR_ARG, "triangle " + std::to_string(k / 3) + " references vertex " + std::to_string(tri[k]) +
                                    " but only " + std::to_string(nv) + " vertices exist");
  HostMesh h;
  h.nv = nv;
  h.nf = nf;
  h.xyz.resize(3 * nv);
  for (int64_t k = 0; k < 3 * nv; ++k) h.xyz[k] = double(xyz[k]);
  h.tri.assign(tri, tri + 3 * nf);
  acvd_params_init(&h.params);
  std::memset(&h.stats, 0, sizeof(h.stats));
  c->meshes.push_back(std::move(h));
  c->invalidate();
  if (mesh_id) *mesh_id = int32_t(c->meshes.size() - 1);
  return ACVD_OK;
}

bool morton_tiles() {
  const char* e = std::getenv("ACVD_TILE_ORDER");
  return e && std::strcmp(e, "morton") == 0;
}

// eval grid (blocks of kBlock threads, grid-stride over the work list)
int64_t sweep_grid_for(int64_t n, int32_t cap) {
  return std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, cap));
}

// ------------------------------------------------------------------- build
int build(acvd_ctx* c) {
  if (c->built) return ACVD_OK;
  if (c->meshes.empty()) return fail(ACVD_ERR_STATE, "no input mesh");
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int M = int(c->meshes.size());
  int64_t N = 0, K = 0, F = 0;
  c->any_iso = c->any_qem = false;
  for (auto& h : c->meshes) {
    if (h.params.number_of_clusters <= 0)
      return fail(ACVD_ERR_ARG, "must set NumberOfClusters to more than zero");
    if (h.params.number_of_clusters > h.nv) return fail(ACVD_ERR_ARG, "more clusters than vertices");
    h.item_off = N;
    h.clus_off = K;
    N += h.nv;
    K += h.params.number_of_clusters;
    F += h.nf;
    if (h.params.metric == ACVD_METRIC_QEM) c->any_qem = true;
    else c->any_iso = true;
  }
  if (N >= (int64_t(1) << 31) - 1 || K >= (int64_t(1) << 31) - 1 || 3 * F >= (int64_t(1) << 31))
    return fail(ACVD_ERR_LIMIT, "context exceeds 2^31 items, clusters or face corners");
  if (M >= (1 << 16)) return fail(ACVD_ERR_LIMIT, "context holds more than 65535 meshes");
  int64_t n_tiles = 0;
  for (auto& h : c->meshes) n_tiles += (h.nv + kTileItems - 1) / kTileItems;
  if (n_tiles * kTileItems >= (int64_t(1) << 31)) return fail(ACVD_ERR_LIMIT, "context exceeds 2^31 tile slots");
  c->n_tiles = int32_t(n_tiles);
  c->N = N;
  c->K = K;
  c->F = F;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, s));

  // per-mesh descriptors
  c->h_meshes.assign(M, DevMesh{});
  std::vector<int32_t> item_mesh, clus_mesh;
  if (M > 1) {
    item_mesh.resize(N);
    clus_mesh.resize(K);
  }
  std::vector<double> pos(3 * N), ind(N, 0.0);
  std::vector<int32_t> tri(3 * F);
  std::vector<int32_t> tile_mesh(c->n_tiles);
  std::vector<float> bbox(6 * M);
  bool any_ind = false;
  // meshes whose indicator is the reference's curvature measure, computed
  // below once the adjacency exists (vtkDiscreteRemeshing::SamplingPreProcessing,
  // DiscreteRemeshing/vtkDiscreteRemeshing.h:583-653, when the metric needs it:
  // gradation > 0, vtkQEMetricForClustering.h:63-67)
  std::vector<char> curv(M, 0);
  bool any_curv = false;
  for (int m = 0; m < M; ++m) {
    const HostMesh& h = c->meshes[m];
    curv[m] = h.indicator.empty() && h.params.curvature_indicator && h.params.gradation > 0.0;
    any_curv |= curv[m] != 0;
  }
  int64_t foff = 0, toff = 0;
  for (int m = 0; m < M; ++m) {
    HostMesh& h = c->meshes[m];
    DevMesh& d = c->h_meshes[m];
    d.item_off = h.item_off;
    d.n_items = h.nv;
    d.clus_off = h.clus_off;
    d.K = h.params.number_of_clusters;
    d.metric = h.params.metric == ACVD_METRIC_QEM ? 1 : 0;
    d.level = h.params.quadrics_optimization_level;
    d.seed = h.params.seed;
    d.has_indicator = (h.indicator.empty() && !curv[m]) ? 0 : 1;
    d.gradation = h.params.gradation;
    d.clamp_ratio = d.metric == 1 ? 10000.0 : 100000.0;
    d.tile_off = int32_t(toff);
    const int64_t nt = (h.nv + kTileItems - 1) / kTileItems;
    std::fill(tile_mesh.begin() + toff, tile_mesh.begin() + toff + nt, m);
    toff += nt;
    {
      double lo[3] = {h.xyz[0], h.xyz[1], h.xyz[2]}, hi[3] = {lo[0], lo[1], lo[2]};
      for (int64_t v = 0; v < h.nv; ++v)
        for (int k = 0; k < 3; ++k) {
          lo[k] = std::min(lo[k], h.xyz[3 * v + k]);
          hi[k] = std::max(hi[k], h.xyz[3 * v + k]);
        }
      // Tiles follow the vertex order (contiguous items: coalesced reads, and
      // spatially coherent when the input order is, as for grid-generated and
      // most scanned meshes) unless ACVD_TILE_ORDER=morton asks for spatial
      // (Morton) tiles, whose cluster sets are tighter for arbitrary orders.
      const bool morton = morton_tiles();
      for (int k = 0; k < 3; ++k) {
        bbox[6 * m + k] = float(lo[k]);
        bbox[6 * m + 3 + k] = (morton && hi[k] > lo[k]) ? float(65535.0 / (hi[k] - lo[k])) : 0.0f;
      }
    }
    std::copy(h.xyz.begin(), h.xyz.end(), pos.begin() + 3 * h.item_off);
    if (!h.indicator.empty()) {
      any_ind = true;
      std::copy(h.indicator.begin(), h.indicator.end(), ind.begin() + h.item_off);
    }
    for (int64_t k = 0; k < 3 * h.nf; ++k) tri[3 * foff + k] = int32_t(h.tri[k] + h.item_off);
    h.face_off = foff;
    foff += h.nf;
    if (M > 1) {
      std::fill(item_mesh.begin() + h.item_off, item_mesh.begin() + h.item_off + h.nv, m);
      std::fill(clus_mesh.begin() + h.clus_off, clus_mesh.begin() + h.clus_off + d.K, m);
    }
  }
  // allocations
  HIP_TRY(c->d_meshes.alloc(M));
  HIP_TRY(c->d_state.alloc(M));
  HIP_TRY(c->d_ctrl.alloc(1));
  HIP_TRY(c->d_mask_all.alloc(M));
  HIP_TRY(c->d_mask.alloc(M));
  HIP_TRY(c->d_ndisc.alloc(M));
  HIP_TRY(c->d_mesh_count.alloc(M));
  if (M > 1) {
    HIP_TRY(c->d_item_mesh.alloc(N));
    HIP_TRY(c->d_clus_mesh.alloc(K));
    HIP_TRY(hipMemcpyAsync(c->d_item_mesh.p, item_mesh.data(), N * sizeof(int32_t), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->d_clus_mesh.p, clus_mesh.data(), K * sizeof(int32_t), hipMemcpyHostToDevice, s));
  }
  HIP_TRY(hipMemcpyAsync(c->d_meshes.p, c->h_meshes.data(), M * sizeof(DevMesh), hipMemcpyHostToDevice, s));
  std::vector<int32_t> ones(M, 1);
  HIP_TRY(hipMemcpyAsync(c->d_mask_all.p, ones.data(), M * sizeof(int32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(c->d_tri.alloc(3 * F));
  HIP_TRY(hipMemcpyAsync(c->d_tri.p, tri.data(), 3 * F * sizeof(int32_t), hipMemcpyHostToDevice, s));
  DBuf<double> d_ind, d_raw;
  DBuf<double>& d_pos = c->d_pos;
  HIP_TRY(d_pos.alloc(3 * N));
  HIP_TRY(hipMemcpyAsync(d_pos.p, pos.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice, s));
  if (any_ind || any_curv) {
    HIP_TRY(d_ind.alloc(N));
    HIP_TRY(hipMemcpyAsync(d_ind.p, ind.data(), N * sizeof(double), hipMemcpyHostToDevice, s));
  }
  // ---- CSR
  {
    DBuf<int32_t> deg2, vfdeg, tptr, tfill, vffill, tcol, deg;
    DBuf<int32_t>& vfptr = c->d_vfptr;
    DBuf<int32_t>& vfcol = c->d_vfcol;
    HIP_TRY(deg2.alloc(N + 1));
    HIP_TRY(vfdeg.alloc(N + 1));
    HIP_TRY(tptr.alloc(N + 1));
    HIP_TRY(vfptr.alloc(N + 1));
    HIP_TRY(tfill.alloc(N));
    HIP_TRY(vffill.alloc(N));
    HIP_TRY(hipMemsetAsync(deg2.p, 0, (N + 1) * sizeof(int32_t), s));
    HIP_TRY(hipMemsetAsync(vfdeg.p, 0, (N + 1) * sizeof(int32_t), s));
    HIP_TRY(hipMemsetAsync(tfill.p, 0, N * sizeof(int32_t), s));
    HIP_TRY(hipMemsetAsync(vffill.p, 0, N * sizeof(int32_t), s));
    HIP_TRY(launch_csr_count(s, F, c->d_tri.p, deg2.p, vfdeg.p));
    HIP_TRY(exclusive_scan_i32(s, deg2.p, tptr.p, N + 1, &c->cub_tmp, &c->cub_bytes));
    HIP_TRY(exclusive_scan_i32(s, vfdeg.p, vfptr.p, N + 1, &c->cub_tmp, &c->cub_bytes));
    int32_t tot[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(&tot[0], tptr.p + N, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&tot[1], vfptr.p + N, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(tcol.alloc(tot[0]));
    HIP_TRY(vfcol.alloc(tot[1]));
    HIP_TRY(launch_csr_fill(s, F, c->d_tri.p, tptr.p, tfill.p, tcol.p, vfptr.p, vffill.p, vfcol.p));
    HIP_TRY(deg.alloc(N + 1));
    HIP_TRY(hipMemsetAsync(deg.p, 0, (N + 1) * sizeof(int32_t), s));
    HIP_TRY(launch_csr_sort_unique(s, N, tptr.p, tcol.p, deg.p, vfptr.p, vfcol.p));
    HIP_TRY(c->d_row_ptr.alloc(N + 1));
    HIP_TRY(exclusive_scan_i32(s, deg.p, c->d_row_ptr.p, N + 1, &c->cub_tmp, &c->cub_bytes));
    int32_t e2 = 0;
    HIP_TRY(hipMemcpyAsync(&e2, c->d_row_ptr.p + N, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    c->E2 = e2;
    HIP_TRY(c->d_col.alloc(e2));
    HIP_TRY(launch_csr_compact(s, N, tptr.p, tcol.p, c->d_row_ptr.p, c->d_col.p));
    HIP_TRY(c->d_ring_adj.alloc(N));
    HIP_TRY(launch_ring_adj(s, N, c->d_row_ptr.p, c->d_col.p, c->d_ring_adj.p));
    // ---- curvature indicators (before the weights that use them)
    for (int m = 0; m < M; ++m) {
      HostMesh& h = c->meshes[m];
      h.curvature.clear();
      if (!curv[m]) continue;
      HIP_TRY(launch_curvature(s, h.item_off, h.nv, d_pos.p, c->d_tri.p, c->d_row_ptr.p, c->d_col.p, vfptr.p,
                               vfcol.p, d_ind.p));
      h.curvature.resize(h.nv);
      HIP_TRY(hipMemcpyAsync(h.curvature.data(), d_ind.p + h.item_off, h.nv * sizeof(double),
                             hipMemcpyDeviceToHost, s));
    }
    if (any_curv) HIP_TRY(hipStreamSynchronize(s));
    // ---- items
    HIP_TRY(c->d_xw.alloc(N));
    HIP_TRY(c->d_q.alloc(10 * N));  // vertex quadrics of every mesh
    HIP_TRY(d_raw.alloc(N));
    SweepArgs a = c->args();
    HIP_TRY(launch_item_build(s, a, d_pos.p, c->d_tri.p, vfptr.p, vfcol.p, (any_ind || any_curv) ? d_ind.p : nullptr,
                              d_raw.p, c->d_xw.p, c->d_q.p));
  }
  // ---- scan tiles in spatial order
  {
    const int64_t nt = c->n_tiles;
    HIP_TRY(c->d_tile_items.alloc(nt * kTileItems));
    HIP_TRY(c->d_item_tile.alloc(N));
    HIP_TRY(c->d_tile_mesh.alloc(nt));
    HIP_TRY(c->d_tile_cnt.alloc(nt));
    HIP_TRY(c->d_tile_set.alloc(nt * kTileSet));
    HIP_TRY(c->d_tile_prop.alloc(nt));
    // no cluster id equals the fill value (free set slots, see k_tile_sets)
    HIP_TRY(launch_fill_i32(s, c->d_tile_set.p, nt * kTileSet, -2));
    HIP_TRY(hipMemcpyAsync(c->d_tile_mesh.p, tile_mesh.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice, s));
    DBuf<float> d_bbox;
    HIP_TRY(d_bbox.alloc(6 * M));
    HIP_TRY(hipMemcpyAsync(d_bbox.p, bbox.data(), 6 * M * sizeof(float), hipMemcpyHostToDevice, s));
    SweepArgs a = c->args();
    HIP_TRY(launch_tile_order(s, a, d_bbox.p, c->d_tile_items.p, c->d_item_tile.p, &c->cub_tmp, &c->cub_bytes));
  }
  // ---- deterministic reductions: clamp then scales
  {
    std::vector<int2> tab;
    for (auto& h : c->meshes)
      for (int64_t b = h.item_off; b < h.item_off + h.nv; b += kReduceItems)
        tab.push_back(make_int2(int32_t(b), int32_t(std::min<int64_t>(b + kReduceItems, h.item_off + h.nv))));
    DBuf<int2> d_tab;
    DBuf<double> d_part;
    HIP_TRY(d_tab.alloc(tab.size()));
    HIP_TRY(d_part.alloc(4 * tab.size() + M));
    HIP_TRY(hipMemcpyAsync(d_tab.p, tab.data(), tab.size() * sizeof(int2), hipMemcpyHostToDevice, s));
    SweepArgs a = c->args();
    HIP_TRY(launch_seg_reduce(s, int(tab.size()), d_tab.p, a, d_raw.p, 0, d_part.p));
    HIP_TRY(launch_clamp_and_scales(s, a, int(tab.size()), d_tab.p, d_part.p, 0, d_raw.p, c->d_meshes.p));
    HIP_TRY(launch_seg_reduce(s, int(tab.size()), d_tab.p, a, d_raw.p, 1, d_part.p));
    HIP_TRY(launch_clamp_and_scales(s, a, int(tab.size()), d_tab.p, d_part.p, 1, d_raw.p, c->d_meshes.p));
    HIP_TRY(hipMemcpyAsync(c->h_meshes.data(), c->d_meshes.p, M * sizeof(DevMesh), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  // ---- sweep state
  HIP_TRY(c->d_labels.alloc(N));
  HIP_TRY(c->d_sizes.alloc(K));
  HIP_TRY(c->d_stamp.alloc(K));
  HIP_TRY(c->d_cand_pass.alloc(N));
  HIP_TRY(c->d_pending.alloc(N));
  HIP_TRY(c->d_acc_ws.alloc(4 * K));
  if (c->any_qem) {
    HIP_TRY(c->d_acc_q.alloc(10 * K));
    HIP_TRY(c->d_T.alloc(K));
    HIP_TRY(c->d_evec.alloc(9 * K));
  }
  HIP_TRY(c->d_prio.alloc(K));
  HIP_TRY(c->d_geom.alloc(K));
  // fixed items per cluster (a change of the fixed clusters rebuilds the
  // context, so the captured chunks never see a stale pointer)
  c->any_fixed = false;
  for (auto& h : c->meshes) c->any_fixed |= !h.fixed.empty();
  if (c->any_fixed) HIP_TRY(c->d_anchor.alloc(K));
  c->any_frozen = false;
  for (auto& h : c->meshes) c->any_frozen |= !h.frozen.empty();
  if (c->any_frozen) HIP_TRY(c->d_frozen.alloc(K));
  HIP_TRY(c->d_cl_round.alloc(K));
  {
    // work slices: scan tiles are dealt round-robin to the slices
    c->work_cap = int32_t(((int64_t(c->n_tiles) + kWorkSlices - 1) / kWorkSlices) * kTileItems);
    HIP_TRY(c->d_work.alloc(size_t(c->work_cap) * kWorkSlices));
    HIP_TRY(c->d_work_count.alloc(kWorkSlices));
    HIP_TRY(hipMemsetAsync(c->d_work_count.p, 0, kWorkSlices * sizeof(int32_t), s));
    // candidate segments: evaluate block b takes work positions
    // b * kBlock + j * n_seg * kBlock (at most N of them) and proposes at most
    // once per item; apply block b appends one moved item per candidate
    // (once per pass)
    c->n_seg = int32_t(sweep_grid_for(N, c->tuning.sweep_segments));
    const int64_t steps = (N + int64_t(c->n_seg) * kBlock - 1) / (int64_t(c->n_seg) * kBlock);
    const int64_t cap = std::max<int64_t>(1, steps) * kBlock;
    if (cap * c->n_seg > (int64_t(1) << 31) - 1) return fail(ACVD_ERR_ARG, "too many items for one context");
    c->cand_cap = int32_t(cap);
    const size_t nc = size_t(c->cand_cap) * size_t(c->n_seg);
    HIP_TRY(c->d_cand.alloc(nc));
    HIP_TRY(c->d_cand_tgt.alloc(nc));
    HIP_TRY(c->d_cand_lab.alloc(nc));
    HIP_TRY(c->d_cand_since.alloc(nc));
    HIP_TRY(c->d_moved.alloc(nc));
    HIP_TRY(c->d_seg_cand.alloc(c->n_seg));
    HIP_TRY(c->d_seg_moved.alloc(c->n_seg));
  }
  HIP_TRY(c->d_parent.alloc(N));
  HIP_TRY(c->d_comp_size.alloc(N));
  HIP_TRY(c->d_ncomp.alloc(K));
  HIP_TRY(c->d_best.alloc(K));
  HIP_TRY(hipMemsetAsync(c->d_ctrl.p, 0, sizeof(Ctrl), s));
  HIP_TRY(hipEventRecord(e1, s));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  c->build_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (auto& h : c->meshes) {
    h.n_adj = 0;
  }
  // per-mesh adjacency sizes (row_ptr at mesh boundaries)
  {
    std::vector<int32_t> rp(N + 1);
    HIP_TRY(hipMemcpy(rp.data(), c->d_row_ptr.p, (N + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (auto& h : c->meshes) {
      h.adj_off = rp[h.item_off];
      h.n_adj = int64_t(rp[h.item_off + h.nv]) - h.adj_off;
    }
  }
  c->built = true;
  return ACVD_OK;
}

// ------------------------------------------------------------ minimisation

int fill_until_stable(acvd_ctx* c, const int32_t* mask, int max_rounds) {
  hipStream_t s = c->stream;
  SweepArgs a = c->args();
  for (int r = 0; r < max_rounds; ++r) {
    HIP_TRY(hipMemsetAsync(&c->d_ctrl.p->fill_changed, 0, sizeof(int32_t), s));
    HIP_TRY(launch_fill_round(s, a, mask, c->d_pending.p));
    int32_t changed = 0;
    HIP_TRY(hipMemcpyAsync(&changed, &c->d_ctrl.p->fill_changed, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (!changed) break;
  }
  return ACVD_OK;
}

int recompute_statistics(acvd_ctx* c, const int32_t* mask, int set_stamp, unsigned long long* mesh_count) {
  hipStream_t s = c->stream;
  SweepArgs a = c->args();
  HIP_TRY(launch_zero_clusters(s, a, mask));
  HIP_TRY(launch_accumulate(s, a, mask, mesh_count));
  HIP_TRY(launch_derive(s, a, mask, set_stamp));
  return ACVD_OK;
}

int enqueue_chunk(acvd_ctx* c, bool qem_unc, bool tiles, int n_pass, int n_rounds) {
  hipStream_t s = c->stream;
  SweepArgs a = c->args();
  for (int p = 0; p < n_pass; ++p) {
    HIP_TRY(launch_scan(s, a, tiles));
    HIP_TRY(launch_eval(s, a, c->any_iso, c->any_qem, qem_unc));
    HIP_TRY(launch_apply(s, a, 0, !qem_unc));
    for (int r = 1; r < n_rounds; ++r) {
      HIP_TRY(launch_reeval(s, a, c->any_iso, c->any_qem, qem_unc, r));
      HIP_TRY(launch_apply(s, a, r, !qem_unc));
    }
    if (tiles) HIP_TRY(launch_tile_update(s, a));
    if (c->energy_log)
      HIP_TRY(launch_energy_log(s, a, c->d_log_cluster.p, c->d_log_ring.p, kEnergyLogRing, nullptr));
    HIP_TRY(launch_controller(s, a));
  }
  return ACVD_OK;
}

// Passes launched one by one with events around each kernel.
int run_chunk_timed(acvd_ctx* c, bool qem_unc, bool tiles, int n_pass, int n_rounds) {
  hipStream_t s = c->stream;
  SweepArgs a = c->args();
  std::vector<hipEvent_t> ev(5 * n_pass);
  for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
  for (int p = 0; p < n_pass; ++p) {
    HIP_TRY(hipEventRecord(ev[5 * p], s));
    HIP_TRY(launch_scan(s, a, tiles));
    HIP_TRY(hipEventRecord(ev[5 * p + 1], s));
    HIP_TRY(launch_eval(s, a, c->any_iso, c->any_qem, qem_unc));
    HIP_TRY(hipEventRecord(ev[5 * p + 2], s));
    HIP_TRY(launch_apply(s, a, 0, !qem_unc));
    for (int r = 1; r < n_rounds; ++r) {
      HIP_TRY(launch_reeval(s, a, c->any_iso, c->any_qem, qem_unc, r));
      HIP_TRY(launch_apply(s, a, r, !qem_unc));
    }
    if (tiles) HIP_TRY(launch_tile_update(s, a));
    HIP_TRY(hipEventRecord(ev[5 * p + 3], s));
    if (c->energy_log)  // counted with the controller
      HIP_TRY(launch_energy_log(s, a, c->d_log_cluster.p, c->d_log_ring.p, kEnergyLogRing, nullptr));
    HIP_TRY(launch_controller(s, a));
    HIP_TRY(hipEventRecord(ev[5 * p + 4], s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  for (int p = 0; p < n_pass; ++p)
    for (int k = 0; k < 4; ++k) {
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, ev[5 * p + k], ev[5 * p + k + 1]));
      c->kernel_ms[k] += ms;
      c->kernel_launches[k] += (k == 1 && c->any_iso && c->any_qem) ? 2 : 1;  // [2]: one accept-round group
    }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return ACVD_OK;
}

// Accept rounds a chunk launches: the most any sweeping mesh runs per pass
// (a mesh's count changes only at a convergence, which ends the chunk's work
// for that mesh until the host cleans it).
int chunk_rounds(const acvd_ctx* c) {
  int r = 1;
  for (size_t m = 0; m < c->meshes.size(); ++m)
    if (c->h_state[m].phase == kPhaseSweep) r = std::max<int>(r, pass_rounds(c->h_state[m]));
  return r;
}

int run_chunk(acvd_ctx* c, bool use_graph, bool qem_unc, bool tiles, bool short_chunk) {
  hipStream_t s = c->stream;
  const int n_pass = short_chunk ? c->tuning.short_chunk_passes : c->tuning.chunk_passes;
  const int n_rounds = chunk_rounds(c);
  if (c->timing) return run_chunk_timed(c, qem_unc, tiles, n_pass, n_rounds);
  if (use_graph) {
    const int k = int(qem_unc) + 2 * int(tiles) + 4 * int(short_chunk) + 8 * (n_rounds - 1);
    if (!c->chunk_exec[k]) {
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int rc = enqueue_chunk(c, qem_unc, tiles, n_pass, n_rounds);
      hipGraph_t g = nullptr;
      hipError_t ec = hipStreamEndCapture(s, &g);
      if (rc != ACVD_OK) return rc;
      HIP_TRY(ec);
      c->chunk_graph[k] = g;
      HIP_TRY(hipGraphInstantiate(&c->chunk_exec[k], g, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(c->chunk_exec[k], s));
    return ACVD_OK;
  }
  return enqueue_chunk(c, qem_unc, tiles, n_pass, n_rounds);
}

// True when no QEM mesh of the context is constrained (the state flips only
// at a convergence, after which the host runs the clean and re-reads it):
// QEM meshes then take the isotropic rule and every derive is a centroid.
bool qem_unconstrained(const acvd_ctx* c) {
  for (size_t m = 0; m < c->meshes.size(); ++m)
    if (c->meshes[m].params.metric == ACVD_METRIC_QEM && (c->h_state[m].constrained || c->h_state[m].anchored))
      return false;  // an anchored mesh carries T terms in every phase
  return true;
}

// Prints the logged passes in [*printed, ctrl.passes) and advances *printed.
void print_pass_log(const Ctrl& ctrl, unsigned long long* printed) {
  unsigned long long p = *printed;
  if (ctrl.passes > kPassLogCap && p < ctrl.passes - kPassLogCap) p = ctrl.passes - kPassLogCap;
  for (; p < ctrl.passes; ++p) {
    const int32_t* e = ctrl.pass_log[p % kPassLogCap];
    std::printf("%sLoop %d: %d items evaluated, %d proposals, %d Modifications (per round:", (e[4 + kMatchRounds] & 2) ? "*" : "",
                e[0], e[1], e[2], e[3]);
    for (int r = 0; r < kMatchRounds; ++r) std::printf(" %d", e[4 + r]);
    std::printf(")\n");
  }
  std::fflush(stdout);
  *printed = ctrl.passes;
}

// Fixed items per cluster on the device (global ids; -1 for free clusters).
int upload_anchors(acvd_ctx* c) {
  if (c->any_frozen) {
    std::vector<int32_t> fr(size_t(c->K), 0);
    for (auto& h : c->meshes)
      for (size_t k = 0; k < h.frozen.size(); ++k) fr[size_t(h.clus_off) + k] = h.frozen[k];
    HIP_TRY(hipMemcpy(c->d_frozen.p, fr.data(), size_t(c->K) * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  if (!c->any_fixed) return ACVD_OK;
  std::vector<int32_t> an(size_t(c->K), -1);
  for (auto& h : c->meshes)
    for (size_t k = 0; k < h.fixed.size(); ++k) an[size_t(h.clus_off) + k] = int32_t(h.item_off + h.fixed[k]);
  HIP_TRY(hipMemcpy(c->d_anchor.p, an.data(), size_t(c->K) * sizeof(int32_t), hipMemcpyHostToDevice));
  return ACVD_OK;
}

// The point of a fixed QEM cluster is its item's input position, exactly
// (GetClusterCentroid of an anchored cluster, vtkQEMetricForClustering.h:270-274).
void exact_anchor_points(const HostMesh& h, double* xyz) {
  if (h.params.metric != ACVD_METRIC_QEM) return;
  for (size_t k = 0; k < h.fixed.size(); ++k)
    for (int d = 0; d < 3; ++d) xyz[3 * k + d] = h.xyz[3 * size_t(h.fixed[k]) + d];
}

// Energy log (SetWriteToGlobalEnergyLog, vtkUniformClustering.h:123, 670-699,
// 1383-1395): buffers for the meshes that asked for it, and the start clock
// (StartTime, taken just before MinimizeEnergy, :689).
int start_energy_log(acvd_ctx* c) {
  c->energy_log = false;
  for (auto& h : c->meshes) {
    h.logged = h.params.write_energy_log != 0;
    h.log_seconds.clear();
    h.log_energy.clear();
    h.log_final = 0.0;
    c->energy_log |= h.logged;
  }
  if (!c->energy_log) return ACVD_OK;
  const size_t M = c->meshes.size();
  if (c->d_log_cluster.n < size_t(c->K)) HIP_TRY(c->d_log_cluster.alloc(size_t(c->K)));
  if (c->d_log_ring.n < M * kEnergyLogRing) HIP_TRY(c->d_log_ring.alloc(M * kEnergyLogRing));
  if (c->d_log_final.n < M) HIP_TRY(c->d_log_final.alloc(M));
  if (!c->d_log_clock0.p) HIP_TRY(c->d_log_clock0.alloc(1));
  HIP_TRY(launch_clock_stamp(c->stream, c->d_log_clock0.p));
  return ACVD_OK;
}

double wall_clock_hz(int device) {
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0) return 0.0;
  return double(khz) * 1e3;
}

// Moves the log entries of the passes run since the last call to the host
// (each pass of a sweeping mesh wrote entry `loops % kEnergyLogRing`; a chunk
// runs fewer passes than the ring holds).
int drain_energy_log(acvd_ctx* c) {
  if (!c->energy_log) return ACVD_OK;
  hipStream_t s = c->stream;
  const int M = int(c->meshes.size());
  std::vector<EnergyLogEntry> ring(size_t(M) * kEnergyLogRing);
  std::vector<MeshState> st(M);
  unsigned long long clock0 = 0;
  HIP_TRY(hipMemcpyAsync(ring.data(), c->d_log_ring.p, ring.size() * sizeof(EnergyLogEntry), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(st.data(), c->d_state.p, M * sizeof(MeshState), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(&clock0, c->d_log_clock0.p, sizeof(clock0), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const double hz = wall_clock_hz(c->device);
  for (int m = 0; m < M; ++m) {
    HostMesh& h = c->meshes[m];
    if (!h.logged) continue;
    for (int64_t L = int64_t(h.log_energy.size()); L < st[m].loops; ++L) {
      const EnergyLogEntry& x = ring[size_t(m) * kEnergyLogRing + size_t(L % kEnergyLogRing)];
      h.log_energy.push_back(x.energy);
      h.log_seconds.push_back(hz > 0.0 ? double(x.clock - clock0) / hz : 0.0);
    }
  }
  return ACVD_OK;
}

// The energy after the run ("Final Energy :", vtkUniformClustering.h:695-699).
int finish_energy_log(acvd_ctx* c) {
  if (!c->energy_log) return ACVD_OK;
  hipStream_t s = c->stream;
  const int M = int(c->meshes.size());
  HIP_TRY(launch_energy_log(s, c->args(), c->d_log_cluster.p, c->d_log_ring.p, kEnergyLogRing, c->d_log_final.p));
  std::vector<double> fin(M);
  HIP_TRY(hipMemcpyAsync(fin.data(), c->d_log_final.p, M * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int m = 0; m < M; ++m)
    if (c->meshes[m].logged) c->meshes[m].log_final = fin[m];
  return ACVD_OK;
}

int minimize_energy(acvd_ctx* c) {
  hipStream_t s = c->stream;
  const int M = int(c->meshes.size());
  SweepArgs a = c->args();
  int rc0 = start_energy_log(c);
  if (rc0) return rc0;
  // the energy log launches its kernels in every pass, outside the captured chunks
  const bool use_graph = std::getenv("ACVD_NO_GRAPH") == nullptr && !c->energy_log;
  // FillHolesInClustering, then ReComputeStatistics + SetAllClustersToModified
  int rc = fill_until_stable(c, c->d_mask_all.p, 1 << 20);
  if (rc) return rc;
  for (int m = 0; m < M; ++m) {
    const acvd_params& p = c->meshes[m].params;
    MeshState& st = c->h_state[m];
    std::memset(&st, 0, sizeof(st));
    st.phase = kPhaseSweep;
    st.max_loops = std::min<int32_t>(p.max_number_of_loops, kMaxLoops - 2);
    st.max_convergences = p.max_number_of_convergences;
    st.unconstrained_init = p.unconstrained_initialization;
    st.anchored = p.metric == ACVD_METRIC_QEM && !c->meshes[m].fixed.empty();
    st.first_rounds = p.first_phase_rounds > 0 ? p.first_phase_rounds : auto_first_rounds(c->meshes[m].nv);
    // ProcessClustering: SetConstrainedClustering(0) when unconstrained
    // initialisation is requested (vtkUniformClustering.h:683-688)
    st.constrained = p.unconstrained_initialization ? 0 : 1;
  }
  HIP_TRY(hipMemcpyAsync(c->d_state.p, c->h_state, M * sizeof(MeshState), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemsetAsync(c->d_mesh_count.p, 0, M * sizeof(unsigned long long), s));
  rc = recompute_statistics(c, c->d_mask_all.p, 1, c->d_mesh_count.p);
  if (rc) return rc;
  std::vector<unsigned long long> cnt(M);
  HIP_TRY(hipMemcpyAsync(cnt.data(), c->d_mesh_count.p, M * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int m = 0; m < M; ++m) c->h_state[m].early_stop_items = int64_t(cnt[m]);
  HIP_TRY(hipMemcpyAsync(c->d_state.p, c->h_state, M * sizeof(MeshState), hipMemcpyHostToDevice, s));
  HIP_TRY(launch_fill_i32(s, c->d_cand_pass.p, c->N, -2));
  HIP_TRY(launch_fill_i32(s, c->d_tile_prop.p, c->n_tiles, -2));
  // tile sets off: every tile is scanned until the moves per pass are few
  bool tiles = false;
  HIP_TRY(launch_fill_i32(s, c->d_tile_cnt.p, c->n_tiles, kTileSet + 1));
  HIP_TRY(launch_fill_i32(s, c->d_cl_round.p, c->K, -1));
  HIP_TRY(launch_fill_u64(s, c->d_prio.p, c->K, ~0ull));
  HIP_TRY(hipMemsetAsync(c->d_ctrl.p, 0, sizeof(Ctrl), s));
  HIP_TRY(hipMemsetAsync(c->d_work_count.p, 0, kWorkSlices * sizeof(int32_t), s));
  for (int k = 0; k < 4; ++k) {
    c->kernel_ms[k] = 0.0;
    c->kernel_launches[k] = 0;
  }
  std::vector<int32_t> mask(M), ndisc(M);
  int64_t guard = 0;
  // per-loop console output (SetConsoleOutput(2), vtkUniformClustering.h:751-758)
  bool loop_output = false;
  for (int m = 0; m < M; ++m) loop_output |= c->meshes[m].params.console_output > 1;
  unsigned long long printed = 0;
  // early-stop threshold of the proposals summed over the sweeping meshes
  // that can still stop early (controller_step); set again after each clean
  auto early_stop_props = [&]() {
    int64_t t = 0;
    for (int m = 0; m < M; ++m) {
      const MeshState& st = c->h_state[m];
      if (st.phase == kPhaseSweep && st.convergences <= 1) t += st.early_stop_items / kEarlyStopDivisor;
    }
    return t;
  };
  int64_t stop_props = early_stop_props();
  bool near_end = false;  // short chunks: the phase is about to end
  while (true) {
    int rc2 = run_chunk(c, use_graph, qem_unconstrained(c), tiles, near_end);
    if (rc2) return rc2;
    HIP_TRY(hipMemcpyAsync(c->h_ctrl, c->d_ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    rc = drain_energy_log(c);
    if (rc) return rc;
    if (loop_output) print_pass_log(*c->h_ctrl, &printed);
    if (c->h_ctrl->any_need_clean) {
      HIP_TRY(hipMemcpyAsync(c->h_state, c->d_state.p, M * sizeof(MeshState), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      for (int m = 0; m < M; ++m) mask[m] = c->h_state[m].phase == kPhaseNeedClean ? 1 : 0;
      HIP_TRY(hipMemcpyAsync(c->d_mask.p, mask.data(), M * sizeof(int32_t), hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemsetAsync(c->d_ndisc.p, 0, M * sizeof(int32_t), s));
      // CleanClustering, FillHolesInClustering
      HIP_TRY(launch_cc(s, a, c->d_mask.p, c->d_parent.p, c->d_comp_size.p, c->d_best.p, c->d_ncomp.p,
                        c->d_ndisc.p));
      rc = fill_until_stable(c, c->d_mask.p, 1 << 20);
      if (rc) return rc;
      HIP_TRY(hipMemcpyAsync(ndisc.data(), c->d_ndisc.p, M * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      for (int m = 0; m < M; ++m) {
        if (!mask[m]) continue;
        MeshState& st = c->h_state[m];
        if (c->meshes[m].params.console_output)
          std::printf("\nConvergence: %d disconnected classes (mesh %d, loop %d)\n", ndisc[m], m, st.loops);
        st.phase = finish_after_clean(st, ndisc[m]) ? kPhaseDone : kPhaseSweep;
      }
      // statistics of the cleaned clusterings; continuing meshes restart
      // with every cluster marked modified
      rc = recompute_statistics(c, c->d_mask.p, 1, nullptr);
      if (rc) return rc;
      // a phase starts with many moves per pass: tile sets off again
      tiles = false;
      near_end = false;
      stop_props = early_stop_props();
      HIP_TRY(launch_fill_i32(s, c->d_tile_cnt.p, c->n_tiles, kTileSet + 1));
      HIP_TRY(hipMemcpyAsync(c->d_state.p, c->h_state, M * sizeof(MeshState), hipMemcpyHostToDevice, s));
      HIP_TRY(hipStreamSynchronize(s));
      bool any = false;
      for (int m = 0; m < M; ++m) any |= c->h_state[m].phase == kPhaseSweep;
      if (!any) break;
      continue;
    }
    if (!c->h_ctrl->any_sweep) break;
    // the last pass proposed at most 3/2 of the early-stop threshold: the
    // phase ends within a few passes, and short chunks keep the passes run
    // after its end (which change nothing) few
    if (c->h_ctrl->passes > 0 && stop_props > 0)
      near_end = 2 * int64_t(c->h_ctrl->pass_log[(c->h_ctrl->passes - 1) % kPassLogCap][2]) <= 3 * stop_props;
    // tile sets on once the last pass moved at most K / kTileOnDivisor items:
    // most tiles then hold no changed cluster and the scan skips them
    if (!tiles && c->h_ctrl->passes > 0 &&
        int64_t(c->h_ctrl->pass_log[(c->h_ctrl->passes - 1) % kPassLogCap][3]) * c->tuning.tile_on_divisor <= c->K) {
      HIP_TRY(launch_tile_sets(s, a, nullptr));
      tiles = true;
    }
    if (++guard > (int64_t(1) << 40)) return fail(ACVD_ERR_STATE, "sweep did not terminate");
  }
  HIP_TRY(hipMemcpyAsync(c->h_state, c->d_state.p, M * sizeof(MeshState), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(c->h_ctrl, c->d_ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  c->last_ctrl = *c->h_ctrl;
  return finish_energy_log(c);
}

// The reference's initial sampling (type 1) of mesh h on a built context:
// the rings in edge-creation order and the vertex areas come from the device
// (k_sample_topology over the build's adjacency and vertex -> face rows), the
// weights (indicator^gradation, clamp) and the region growing, sequential by
// nature, run on the host.  Same labels as reference_initial_sampling.
int device_reference_sampling(acvd_ctx* c, const HostMesh& h, const double* ind, int32_t* lab) {
  hipStream_t s = c->stream;
  const int64_t nv = h.nv;
  // the shuffled order depends on nv alone: drawn on another thread while
  // the device builds the rings and areas
  SamplingOrder order;
  std::thread shuffler([&order, nv] { sampling_order(nv, &order); });
  struct Join {
    std::thread& t;
    ~Join() {
      if (t.joinable()) t.join();
    }
  } join{shuffler};
  DBuf<int32_t> ring, len, key;
  DBuf<double> area;
  HIP_TRY(ring.alloc(size_t(c->E2)));
  HIP_TRY(len.alloc(size_t(nv)));
  HIP_TRY(key.alloc(c->d_vfcol.n));
  HIP_TRY(area.alloc(size_t(nv)));
  HIP_TRY(launch_sample_topology(s, h.item_off, nv, c->d_pos.p, c->d_tri.p, c->d_vfptr.p, c->d_vfcol.p,
                                 c->d_row_ptr.p, ring.p, len.p, key.p, area.p));
  huge_vector<int32_t> h_ring(static_cast<size_t>(h.n_adj)), h_len(static_cast<size_t>(nv)),
      rp(static_cast<size_t>(nv + 1));
  huge_vector<double> w(static_cast<size_t>(nv));
  huge_vector<int64_t> rbeg(static_cast<size_t>(nv));
  HIP_TRY(hipMemcpyAsync(h_ring.data(), ring.p + h.adj_off, size_t(h.n_adj) * sizeof(int32_t), hipMemcpyDeviceToHost,
                         s));
  HIP_TRY(hipMemcpyAsync(h_len.data(), len.p, size_t(nv) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(w.data(), area.p, size_t(nv) * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(rp.data(), c->d_row_ptr.p + h.item_off, size_t(nv + 1) * sizeof(int32_t),
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  parallel_for(nv, [&](int64_t v0, int64_t v1) {
    for (int64_t v = v0; v < v1; ++v) rbeg[size_t(v)] = int64_t(rp[size_t(v)]) - h.adj_off;
  });
  sampling_weights(nv, h.params.metric, ind, h.params.gradation, w.data());
  shuffler.join();
  if (!region_growing(nv, rbeg.data(), h_len.data(), h_ring.data(), w.data(), h.params.number_of_clusters, lab,
                      int64_t(h.fixed.size()), h.fixed.data(), &order))
    return fail(ACVD_ERR_STATE, "initial sampling failed");
  return ACVD_OK;
}

int initial_labels(acvd_ctx* c) {
  hipStream_t s = c->stream;
  SweepArgs a = c->args();
  for (auto& h : c->meshes) {
    const int32_t K = h.params.number_of_clusters;
    if (!h.init_labels.empty() || h.params.initial_sampling != 1) {
      huge_vector<int32_t> lab(h.nv);
      if (h.init_labels.empty() && h.params.initial_sampling >= 2) {
        // the reference's sampling types 0 (one random item per cluster) and
        // -1 (item k for cluster k)
        if (int64_t(K) > h.nv) return fail(ACVD_ERR_ARG, "more clusters than items");
        if (h.params.initial_sampling == 2) {
          reference_seed_sampling(h.nv, K, lab.data());
        } else {
          for (int64_t i = 0; i < h.nv; ++i) lab[i] = i < K ? int32_t(i) : K;
        }
        for (int64_t i = 0; i < h.nv; ++i) lab[i] = lab[i] == K ? kNull : int32_t(lab[i] + h.clus_off);
      } else if (h.init_labels.empty()) {
        // the reference's region growing, then local ids -> global
        // the indicator of the weights: the set one, else the computed curvature
        const std::vector<double>& ind = !h.indicator.empty() ? h.indicator : h.curvature;
        if (int rc = device_reference_sampling(c, h, ind.empty() ? nullptr : ind.data(), lab.data())) return rc;
        for (int64_t i = 0; i < h.nv; ++i) lab[i] = lab[i] == K ? kNull : int32_t(lab[i] + h.clus_off);
      } else {
        for (int64_t i = 0; i < h.nv; ++i) {
          int32_t l = h.init_labels[i];
          lab[i] = (l < 0 || l >= K) ? kNull : int32_t(l + h.clus_off);
        }
      }
      HIP_TRY(hipMemcpyAsync(c->d_labels.p + h.item_off, lab.data(), h.nv * sizeof(int32_t), hipMemcpyHostToDevice,
                             s));
      HIP_TRY(hipStreamSynchronize(s));
    } else {
      DBuf<uint32_t> k0, k1;
      DBuf<int32_t> v0, v1;
      HIP_TRY(k0.alloc(h.nv));
      HIP_TRY(k1.alloc(h.nv));
      HIP_TRY(v0.alloc(h.nv));
      HIP_TRY(v1.alloc(h.nv));
      HIP_TRY(launch_fill_i32(s, c->d_labels.p + h.item_off, h.nv, kNull));
      HIP_TRY(launch_seed_keys(s, a, h.item_off, h.nv, h.params.seed, k0.p, v0.p));
      HIP_TRY(sort_pairs_u32(s, k0.p, k1.p, v0.p, v1.p, h.nv, &c->cub_tmp, &c->cub_bytes));
      HIP_TRY(launch_seed_assign(s, c->d_labels.p, h.item_off, h.clus_off, K, v1.p));
      HIP_TRY(hipStreamSynchronize(s));
    }
  }
  return ACVD_OK;
}

}  // namespace

// ============================================================== C interface
extern "C" {

const char* acvd_version(void) { return "acvd_amd 0.1 (gfx950)"; }

const char* acvd_last_error(void) { return g_last_error.c_str(); }

int acvd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void acvd_params_init(acvd_params* p) {
  if (!p) return;
  p->number_of_clusters = 0;
  p->metric = ACVD_METRIC_ISOTROPIC;
  p->quadrics_optimization_level = 3;
  p->unconstrained_initialization = 0;
  p->max_number_of_loops = 5000000;
  p->max_number_of_convergences = 1000000000;
  p->console_output = 0;
  p->seed = 0;
  p->initial_sampling = 0;
  p->curvature_indicator = 0;
  p->first_phase_rounds = 0;
  p->write_energy_log = 0;
  p->gradation = 0.0;
}

int acvd_create(acvd_ctx** out, int device) {
  if (!out) return fail(ACVD_ERR_ARG, "null output pointer");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return fail(ACVD_ERR_HIP, std::string("no HIP device available: ") + hipGetErrorString(e));
  if (device < 0 || device >= n) return fail(ACVD_ERR_ARG, "device index out of range");
  std::unique_ptr<acvd_ctx> c(new acvd_ctx());
  c->device = device;
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&c->h_ctrl), sizeof(Ctrl), hipHostMallocDefault));
  *out = c.release();
  return ACVD_OK;
}

int acvd_destroy(acvd_ctx* ctx) {
  if (!ctx) return ACVD_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return ACVD_OK;
}

int acvd_set_input_f64(acvd_ctx* ctx, int64_t nv, const double* xyz, int64_t nf, const int32_t* tri) {
  if (!ctx) return fail(ACVD_ERR_ARG, "null context");
  ctx->meshes.clear();
  return add_mesh(ctx, nv, xyz, nf, tri, nullptr);
}
int acvd_set_input_f32(acvd_ctx* ctx, int64_t nv, const float* xyz, int64_t nf, const int32_t* tri) {
  if (!ctx) return fail(ACVD_ERR_ARG, "null context");
  ctx->meshes.clear();
  return add_mesh(ctx, nv, xyz, nf, tri, nullptr);
}
int acvd_add_input_f64(acvd_ctx* ctx, int64_t nv, const double* xyz, int64_t nf, const int32_t* tri,
                       int32_t* mesh_id) {
  return add_mesh(ctx, nv, xyz, nf, tri, mesh_id);
}
int acvd_add_input_f32(acvd_ctx* ctx, int64_t nv, const float* xyz, int64_t nf, const int32_t* tri,
                       int32_t* mesh_id) {
  return add_mesh(ctx, nv, xyz, nf, tri, mesh_id);
}
int acvd_num_meshes(const acvd_ctx* ctx) { return ctx ? int(ctx->meshes.size()) : 0; }

int acvd_set_params(acvd_ctx* ctx, int32_t mesh_id, const acvd_params* p) {
  if (!ctx || !p) return fail(ACVD_ERR_ARG, "null argument");
  if (p->metric != ACVD_METRIC_ISOTROPIC && p->metric != ACVD_METRIC_QEM) return fail(ACVD_ERR_ARG, "unknown metric");
  if (p->quadrics_optimization_level < 0 || p->quadrics_optimization_level > 3)
    return fail(ACVD_ERR_ARG, "quadrics optimization level must be 0..3");
  if (p->initial_sampling < 0 || p->initial_sampling > 3) return fail(ACVD_ERR_ARG, "initial sampling must be 0..3");
  if (p->max_number_of_loops < 0) return fail(ACVD_ERR_ARG, "negative loop limit");
  if (p->first_phase_rounds < 0 || p->first_phase_rounds > kMatchRounds)
    return fail(ACVD_ERR_ARG, "first phase rounds must be 0..12");
  if (mesh_id == -1) {
    for (auto& h : ctx->meshes) h.params = *p;
  } else {
    if (int rc = check_mesh(ctx, mesh_id)) return rc;
    ctx->meshes[mesh_id].params = *p;
  }
  ctx->invalidate();
  return ACVD_OK;
}

int acvd_get_params(const acvd_ctx* ctx, int32_t mesh_id, acvd_params* p) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (!p) return fail(ACVD_ERR_ARG, "null argument");
  *p = ctx->meshes[mesh_id].params;
  return ACVD_OK;
}

int acvd_set_indicator(acvd_ctx* ctx, int32_t mesh_id, const double* indicator) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  HostMesh& h = ctx->meshes[mesh_id];
  if (!indicator) h.indicator.clear();
  else h.indicator.assign(indicator, indicator + h.nv);
  ctx->invalidate();
  return ACVD_OK;
}

int acvd_get_indicator(acvd_ctx* ctx, int32_t mesh_id, double* out) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (!out) return fail(ACVD_ERR_ARG, "null argument");
  const HostMesh& h = ctx->meshes[mesh_id];
  const std::vector<double>& src = !h.indicator.empty() ? h.indicator : h.curvature;
  if (src.empty())
    return fail(ACVD_ERR_STATE, "no indicator: none was set and no build computed the curvature indicator");
  std::copy(src.begin(), src.end(), out);
  return ACVD_OK;
}

int acvd_set_initial_labels(acvd_ctx* ctx, int32_t mesh_id, const int32_t* labels) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  HostMesh& h = ctx->meshes[mesh_id];
  if (!labels) h.init_labels.clear();
  else h.init_labels.assign(labels, labels + h.nv);
  return ACVD_OK;
}

int acvd_subdivide(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, int64_t* out_nv, int64_t* out_nf,
                   double* out_xyz, int32_t* out_tri, int32_t* parents) {
  if (nv <= 0 || nf <= 0 || !xyz || !tri || !out_nv || !out_nf) return fail(ACVD_ERR_ARG, "empty mesh or null buffer");
  for (int64_t k = 0; k < 3 * nf; ++k)
    if (tri[k] < 0 || tri[k] >= nv) return fail(ACVD_ERR_ARG, "triangle references a missing vertex");
  std::vector<double> x;
  std::vector<int32_t> t, par;
  const int64_t ne = subdivide_mesh(nv, xyz, nf, tri, &x, &t, &par);
  if (nv + ne >= (int64_t(1) << 31) - 1) return fail(ACVD_ERR_LIMIT, "subdivided mesh exceeds 2^31 vertices");
  *out_nv = nv + ne;
  *out_nf = int64_t(t.size() / 3);
  if (out_xyz) std::memcpy(out_xyz, x.data(), x.size() * sizeof(double));
  if (out_tri) std::memcpy(out_tri, t.data(), t.size() * sizeof(int32_t));
  if (parents) std::memcpy(parents, par.data(), par.size() * sizeof(int32_t));
  return ACVD_OK;
}

int acvd_subdivide_input(acvd_ctx* ctx, int32_t mesh_id, int64_t min_vertices, int32_t* levels) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  HostMesh& h = ctx->meshes[mesh_id];
  if (levels) *levels = 0;
  if (h.nv >= min_vertices) return ACVD_OK;
  if (!h.init_labels.empty())
    return fail(ACVD_ERR_STATE, "initial labels are set for the input's vertices; the input needs subdivision");
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const bool with_ind = !h.indicator.empty();
  int64_t nv = h.nv, nf = h.nf;
  DBuf<double> xyz, ind;
  DBuf<int32_t> tri;
  HIP_TRY(xyz.alloc(size_t(3 * nv)));
  HIP_TRY(tri.alloc(size_t(3 * nf)));
  HIP_TRY(hipMemcpyAsync(xyz.p, h.xyz.data(), sizeof(double) * 3 * nv, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(tri.p, h.tri.data(), sizeof(int32_t) * 3 * nf, hipMemcpyHostToDevice, s));
  if (with_ind) {
    HIP_TRY(ind.alloc(size_t(nv)));
    HIP_TRY(hipMemcpyAsync(ind.p, h.indicator.data(), sizeof(double) * nv, hipMemcpyHostToDevice, s));
  }
  int32_t done = 0;
  while (nv < min_vertices) {
    const int64_t n3 = 3 * nf;
    if (n3 + 1 >= int64_t(INT32_MAX)) return fail(ACVD_ERR_LIMIT, "subdivision: more than 2^31 half-edges");
    DBuf<unsigned long long> k0, k1;
    DBuf<int32_t> v0, v1, flag, num, face_edge, active, fpos;
    HIP_TRY(k0.alloc(size_t(n3)));
    HIP_TRY(k1.alloc(size_t(n3)));
    HIP_TRY(v0.alloc(size_t(n3)));
    HIP_TRY(v1.alloc(size_t(n3)));
    HIP_TRY(flag.alloc(size_t(n3 + 1)));
    HIP_TRY(num.alloc(size_t(n3 + 1)));
    HIP_TRY(face_edge.alloc(size_t(n3)));
    HIP_TRY(active.alloc(size_t(nf + 1)));
    HIP_TRY(fpos.alloc(size_t(nf + 1)));
    HIP_TRY(hipMemsetAsync(flag.p + n3, 0, sizeof(int32_t), s));
    HIP_TRY(hipMemsetAsync(active.p + nf, 0, sizeof(int32_t), s));
    HIP_TRY(launch_sub_keys(s, nf, nv, tri.p, k0.p, v0.p, face_edge.p, active.p));
    HIP_TRY(sort_sub_keys(s, n3, nv, k0.p, k1.p, v0.p, v1.p, &ctx->cub_tmp, &ctx->cub_bytes));
    HIP_TRY(launch_sub_heads(s, n3, nv, k1.p, v1.p, flag.p));
    HIP_TRY(exclusive_scan_i32(s, flag.p, num.p, n3 + 1, &ctx->cub_tmp, &ctx->cub_bytes));
    HIP_TRY(exclusive_scan_i32(s, active.p, fpos.p, nf + 1, &ctx->cub_tmp, &ctx->cub_bytes));
    int32_t ne = 0, nact = 0;
    HIP_TRY(hipMemcpyAsync(&ne, num.p + n3, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&nact, fpos.p + nf, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (nv + ne >= kMaxItemsPerMesh) return fail(ACVD_ERR_LIMIT, "subdivided mesh has more than 2^27 vertices");
    if (nact == 0) return fail(ACVD_ERR_ARG, "subdivision: no active face");
    const int64_t nv2 = nv + ne, nf2 = 4 * int64_t(nact);
    DBuf<double> xyz2, ind2;
    DBuf<int32_t> tri2;
    HIP_TRY(xyz2.alloc(size_t(3 * nv2)));
    HIP_TRY(tri2.alloc(size_t(3 * nf2)));
    HIP_TRY(hipMemcpyAsync(xyz2.p, xyz.p, sizeof(double) * 3 * nv, hipMemcpyDeviceToDevice, s));
    if (with_ind) {
      HIP_TRY(ind2.alloc(size_t(nv2)));
      HIP_TRY(hipMemcpyAsync(ind2.p, ind.p, sizeof(double) * nv, hipMemcpyDeviceToDevice, s));
    }
    HIP_TRY(launch_sub_edges(s, n3, nv, k1.p, v1.p, num.p, tri.p, xyz.p, xyz2.p, with_ind ? ind.p : nullptr,
                             with_ind ? ind2.p : nullptr, face_edge.p));
    HIP_TRY(launch_sub_faces(s, nf, nv, tri.p, face_edge.p, fpos.p, tri2.p));
    HIP_TRY(hipStreamSynchronize(s));
    std::swap(xyz.p, xyz2.p);
    std::swap(xyz.n, xyz2.n);
    std::swap(tri.p, tri2.p);
    std::swap(tri.n, tri2.n);
    std::swap(ind.p, ind2.p);
    std::swap(ind.n, ind2.n);
    nv = nv2;
    nf = nf2;
    ++done;
  }
  h.xyz.resize(size_t(3 * nv));
  h.tri.resize(size_t(3 * nf));
  HIP_TRY(hipMemcpyAsync(h.xyz.data(), xyz.p, sizeof(double) * 3 * nv, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(h.tri.data(), tri.p, sizeof(int32_t) * 3 * nf, hipMemcpyDeviceToHost, s));
  if (with_ind) {
    h.indicator.resize(size_t(nv));
    HIP_TRY(hipMemcpyAsync(h.indicator.data(), ind.p, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  h.nv = nv;
  h.nf = nf;
  h.curvature.clear();
  h.out_xyz.clear();
  h.out_tri.clear();
  h.clustered = false;
  h.logged = false;
  ctx->invalidate();
  if (levels) *levels = done;
  return ACVD_OK;
}

int acvd_get_input_sizes(const acvd_ctx* ctx, int32_t mesh_id, int64_t* nv, int64_t* nf) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (nv) *nv = ctx->meshes[mesh_id].nv;
  if (nf) *nf = ctx->meshes[mesh_id].nf;
  return ACVD_OK;
}

int acvd_get_input(const acvd_ctx* ctx, int32_t mesh_id, double* xyz, int32_t* tri) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  const HostMesh& h = ctx->meshes[mesh_id];
  if (xyz) std::copy(h.xyz.begin(), h.xyz.end(), xyz);
  if (tri) std::copy(h.tri.begin(), h.tri.end(), tri);
  return ACVD_OK;
}

int acvd_set_fixed_clusters(acvd_ctx* ctx, int32_t mesh_id, int64_t n, const int32_t* items) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (n < 0 || (n > 0 && !items)) return fail(ACVD_ERR_ARG, "null fixed-item list");
  HostMesh& h = ctx->meshes[mesh_id];
  std::vector<int32_t> ids(items, items + n);
  for (int32_t v : ids)
    if (v < 0 || v >= h.nv) return fail(ACVD_ERR_ARG, "fixed item out of range");
  std::vector<int32_t> sorted(ids);
  std::sort(sorted.begin(), sorted.end());
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
    return fail(ACVD_ERR_ARG, "an item is fixed to two clusters");
  h.fixed.swap(ids);
  ctx->invalidate();
  return ACVD_OK;
}

int acvd_build(acvd_ctx* ctx) {
  if (int rc = check_ctx(ctx)) return rc;
  return build(ctx);
}

int acvd_process_clustering(acvd_ctx* ctx) {
  if (int rc = check_ctx(ctx)) return rc;
  if (ctx->meshes.empty()) return fail(ACVD_ERR_STATE, "no input mesh");
  // Init() clears IsClusterFreezed (vtkUniformClustering.h:1370-1375): only
  // the ForceManifold restarts run with frozen clusters
  if (!ctx->keep_frozen)
    for (auto& h : ctx->meshes)
      if (!h.frozen.empty()) {
        h.frozen.clear();
        ctx->invalidate();
      }
  for (auto& h : ctx->meshes) {
    if (h.params.number_of_clusters <= 0) return fail(ACVD_ERR_ARG, "must set NumberOfClusters to more than zero");
    if (int64_t(h.fixed.size()) > h.params.number_of_clusters)
      return fail(ACVD_ERR_ARG, "more fixed clusters than clusters");
    if (!h.fixed.empty() && h.params.initial_sampling != 0 && h.init_labels.empty())
      return fail(ACVD_ERR_ARG, "fixed clusters need the reference's initial sampling (0) or initial labels");
    if (!h.frozen.empty() && int64_t(h.frozen.size()) != h.params.number_of_clusters)
      return fail(ACVD_ERR_ARG, "frozen flags do not match the number of clusters");
  }
  int rc = build(ctx);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  rc = upload_anchors(ctx);
  if (rc) return rc;
  const int M = int(ctx->meshes.size());
  // (re)allocate the pinned state mirror for M meshes
  if (ctx->h_state) (void)hipHostFree(ctx->h_state);
  ctx->h_state = nullptr;
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&ctx->h_state), M * sizeof(MeshState), hipHostMallocDefault));
  for (auto& h : ctx->meshes)
    if (h.params.console_output) std::printf("Clustering......\n");
  rc = initial_labels(ctx);
  if (rc) return rc;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, ctx->stream));
  rc = minimize_energy(ctx);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(e1, ctx->stream));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->converge_ms = ms;
  for (int m = 0; m < M; ++m) {
    HostMesh& h = ctx->meshes[m];
    h.clustered = true;
    h.out_xyz.clear();
    h.out_tri.clear();
    const MeshState& st = ctx->h_state[m];
    acvd_stats& S = h.stats;
    S.loops = st.loops;
    S.convergences = st.convergences;
    S.total_moves = st.total_mods;
    S.last_disconnected = st.last_disconnected;
    S.finished = st.phase == kPhaseDone ? 1 : 0;
    S.ctx_passes = int64_t(ctx->last_ctrl.passes);
    S.ctx_evaluated = int64_t(ctx->last_ctrl.evaluated);
    S.ctx_proposed = int64_t(ctx->last_ctrl.proposed);
    S.ctx_accepted = int64_t(ctx->last_ctrl.accepted);
    S.ctx_rounds = int64_t(ctx->last_ctrl.rounds);
    S.ctx_walked = int64_t(ctx->last_ctrl.walked);
    S.ctx_reeval_live = int64_t(ctx->last_ctrl.reeval_live);
    S.ctx_reevaluated = int64_t(ctx->last_ctrl.reevaluated);
    S.ctx_rechecked = int64_t(ctx->last_ctrl.rechecked);
    S.ctx_accepted_constrained = int64_t(ctx->last_ctrl.accepted_constrained);
    S.converge_ms = ms;
    S.build_ms = ctx->build_ms;
    S.n_items = h.nv;
    S.n_adjacency = h.n_adj;
    for (int k = 0; k < 4; ++k) {
      S.kernel_ms[k] = ctx->kernel_ms[k];
      S.kernel_launches[k] = ctx->kernel_launches[k];
    }
    if (h.params.console_output) {
      std::printf("The clustering took :%g seconds.\n", ms / 1000.0);
      std::printf("Number of loops:\t%lld\n", (long long)st.loops);
    }
  }
  return ACVD_OK;
}

int acvd_initial_sampling(acvd_ctx* ctx) {
  if (int rc = check_ctx(ctx)) return rc;
  int rc = build(ctx);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  rc = initial_labels(ctx);
  if (rc) return rc;
  // the first FillHolesInClustering of MinimizeEnergy (vtkUniformClustering.h:727)
  return fill_until_stable(ctx, ctx->d_mask_all.p, 1 << 20);
}

int acvd_set_timing(acvd_ctx* ctx, int enable) {
  if (int rc = check_ctx(ctx)) return rc;
  if (enable < 0 || enable > 2) return fail(ACVD_ERR_ARG, "timing mode must be 0, 1 or 2");
  ctx->timing = enable != 0;
  ctx->count_rounds = enable == 1;
  return ACVD_OK;
}

int acvd_reference_initial_sampling(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, int32_t metric,
                                    const double* indicator, double gradation, int32_t n_clusters, int32_t* labels) {
  if (nv <= 0 || nf <= 0 || !xyz || !tri || !labels) return fail(ACVD_ERR_ARG, "empty mesh or null buffer");
  if (n_clusters <= 0 || n_clusters > nv) return fail(ACVD_ERR_ARG, "number of clusters must be in 1..nv");
  if (metric != ACVD_METRIC_ISOTROPIC && metric != ACVD_METRIC_QEM) return fail(ACVD_ERR_ARG, "unknown metric");
  for (int64_t k = 0; k < 3 * nf; ++k)
    if (tri[k] < 0 || tri[k] >= nv) return fail(ACVD_ERR_ARG, "triangle references a missing vertex");
  if (!reference_initial_sampling(nv, xyz, nf, tri, metric, indicator, gradation, n_clusters, labels))
    return fail(ACVD_ERR_STATE, "initial sampling failed");
  return ACVD_OK;
}

int acvd_build_output(acvd_ctx* ctx, int32_t mesh_id, int32_t quadrics_level) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  HostMesh& h = ctx->meshes[mesh_id];
  if (!ctx->built || !h.clustered) return fail(ACVD_ERR_STATE, "no clustering computed");
  if (quadrics_level < 0 || quadrics_level > 3) return fail(ACVD_ERR_ARG, "quadrics level must be 0..3");
  if (h.params.number_of_clusters >= (1 << 21)) return fail(ACVD_ERR_ARG, "output mesh limited to 2^21 - 1 clusters");
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int32_t K = h.params.number_of_clusters;
  DBuf<double> pts;
  DBuf<int32_t> first, otri;
  HIP_TRY(pts.alloc(3 * size_t(K)));
  HIP_TRY(first.alloc(1));
  HIP_TRY(otri.alloc(3 * size_t(h.nf)));
  HIP_TRY(launch_cluster_points(s, ctx->args(), mesh_id, pts.p));
  HIP_TRY(launch_fill_i32(s, first.p, 1, K));
  HIP_TRY(fix_empty_cluster_points(s, ctx->d_sizes.p, h.clus_off, K, first.p, pts.p));
  if (quadrics_level > 0) {
    DBuf<unsigned long long> acc;
    HIP_TRY(acc.alloc(10 * size_t(K)));
    HIP_TRY(quadric_placement(s, h.nv, h.item_off, ctx->d_labels.p, h.clus_off, K, ctx->d_q.p,
                              ctx->h_meshes[mesh_id].scales.sq, acc.p, quadrics_level, pts.p));
    HIP_TRY(hipStreamSynchronize(s));
  }
  int64_t nf_out = 0;
  HIP_TRY(build_dual_faces(s, h.nf, ctx->d_tri.p + 3 * h.face_off, ctx->d_labels.p, h.item_off, h.clus_off, otri.p,
                           &nf_out));
  h.out_xyz.resize(3 * size_t(K));
  h.out_tri.resize(3 * size_t(nf_out));
  HIP_TRY(hipMemcpyAsync(h.out_xyz.data(), pts.p, 3 * size_t(K) * sizeof(double), hipMemcpyDeviceToHost, s));
  if (nf_out)
    HIP_TRY(hipMemcpyAsync(h.out_tri.data(), otri.p, 3 * size_t(nf_out) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  exact_anchor_points(h, h.out_xyz.data());
  return ACVD_OK;
}

int acvd_set_tuning(acvd_ctx* ctx, const acvd_tuning* t) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!t) return fail(ACVD_ERR_ARG, "null tuning");
  if (t->sweep_segments < 1 || t->sweep_segments > kMaxSeg) return fail(ACVD_ERR_ARG, "sweep_segments must be 1..2048");
  if (t->tile_on_divisor < 1 || t->tile_on_divisor > 1024) return fail(ACVD_ERR_ARG, "tile_on_divisor must be 1..1024");
  if (t->chunk_passes < 1 || t->chunk_passes > kChunk) return fail(ACVD_ERR_ARG, "chunk_passes must be 1..16");
  if (t->short_chunk_passes < 1 || t->short_chunk_passes > t->chunk_passes)
    return fail(ACVD_ERR_ARG, "short_chunk_passes must be 1..chunk_passes");
  const bool rebuild = t->sweep_segments != ctx->tuning.sweep_segments;
  ctx->tuning = *t;
  if (rebuild) ctx->invalidate();
  else ctx->drop_graphs();
  return ACVD_OK;
}

int acvd_get_tuning(acvd_ctx* ctx, acvd_tuning* t) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!t) return fail(ACVD_ERR_ARG, "null tuning");
  *t = ctx->tuning;
  return ACVD_OK;
}

int acvd_autotune(acvd_ctx* ctx, int32_t repeats, acvd_tuning* best_out, double* best_ms_out) {
  if (int rc = check_ctx(ctx)) return rc;
  if (ctx->meshes.empty()) return fail(ACVD_ERR_STATE, "no input mesh");
  if (repeats < 1 || repeats > 16) return fail(ACVD_ERR_ARG, "repeats must be 1..16");
  // one timed point: the best of `repeats` clusterings, and the labels
  auto measure = [&](const acvd_tuning& t, double* ms, std::vector<int32_t>* lab) -> int {
    if (int rc = acvd_set_tuning(ctx, &t)) return rc;
    double best = 1e300;
    for (int r = 0; r < repeats; ++r) {
      if (int rc = acvd_process_clustering(ctx)) return rc;
      best = std::min(best, ctx->converge_ms);
    }
    lab->resize(size_t(ctx->N));
    HIP_TRY(hipMemcpy(lab->data(), ctx->d_labels.p, size_t(ctx->N) * sizeof(int32_t), hipMemcpyDeviceToHost));
    *ms = best;
    return ACVD_OK;
  };
  const acvd_tuning start = ctx->tuning;
  acvd_tuning cur = start;
  double cur_ms = 0.0;
  std::vector<int32_t> ref, lab;
  if (int rc = measure(cur, &cur_ms, &ref)) return rc;
  static const int32_t kSegs[] = {256, 512, 768, 1024, 1152, 1280, 1536, 2048};
  static const int32_t kDivs[] = {4, 8, 16, 32, 64};
  static const int32_t kChunks[] = {4, 8, 12, 16};
  static const int32_t kShorts[] = {1, 2, 4, 8};
  const int32_t* cand[4] = {kSegs, kDivs, kChunks, kShorts};
  const int ncand[4] = {8, 5, 4, 4};
  for (int knob = 0; knob < 4; ++knob)
    for (int j = 0; j < ncand[knob]; ++j) {
      acvd_tuning t = cur;
      int32_t* f[4] = {&t.sweep_segments, &t.tile_on_divisor, &t.chunk_passes, &t.short_chunk_passes};
      if (*f[knob] == cand[knob][j]) continue;
      *f[knob] = cand[knob][j];
      if (t.short_chunk_passes > t.chunk_passes) continue;
      double ms = 0.0;
      if (int rc = measure(t, &ms, &lab)) return rc;
      if (lab != ref) {
        (void)acvd_set_tuning(ctx, &start);
        return fail(ACVD_ERR_STATE, "the labels changed with the launch shape");
      }
      if (ms < 0.99 * cur_ms) {
        cur = t;
        cur_ms = ms;
      }
    }
  if (int rc = measure(cur, &cur_ms, &lab)) return rc;
  if (best_out) *best_out = cur;
  if (best_ms_out) *best_ms_out = cur_ms;
  return ACVD_OK;
}

int acvd_set_frozen_clusters(acvd_ctx* ctx, int32_t mesh_id, int64_t n, const int32_t* flags) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  HostMesh& h = ctx->meshes[mesh_id];
  if (n != 0 && (!flags || n != h.params.number_of_clusters))
    return fail(ACVD_ERR_ARG, "frozen flags: one per cluster (or none)");
  h.frozen.assign(flags, flags + n);
  ctx->keep_frozen = n != 0;
  ctx->invalidate();
  return ACVD_OK;
}

int acvd_vertex_manifold(int64_t nv, int64_t nf, const int32_t* tri, uint8_t* manifold) {
  if (nv <= 0 || nf < 0 || (nf && !tri) || !manifold) return fail(ACVD_ERR_ARG, "empty mesh or null buffer");
  for (int64_t k = 0; k < 3 * nf; ++k)
    if (tri[k] < 0 || tri[k] >= nv) return fail(ACVD_ERR_ARG, "triangle references a missing vertex");
  std::vector<uint8_t> ok;
  vertex_manifold(nv, nf, tri, &ok);
  std::memcpy(manifold, ok.data(), size_t(nv));
  return ACVD_OK;
}

int acvd_force_manifold_step(int64_t nv, int64_t nf, const int32_t* tri, int64_t n_out, const int32_t* out_tri,
                             int32_t* n_clusters, int32_t* labels, int32_t* frozen, int64_t frozen_capacity,
                             int64_t* issues) {
  if (nv <= 0 || nf <= 0 || !tri || n_out < 0 || (n_out && !out_tri) || !n_clusters || !labels || !issues)
    return fail(ACVD_ERR_ARG, "empty mesh or null buffer");
  const int32_t K = *n_clusters;
  if (K <= 0) return fail(ACVD_ERR_ARG, "no clusters");
  for (int64_t k = 0; k < 3 * nf; ++k)
    if (tri[k] < 0 || tri[k] >= nv) return fail(ACVD_ERR_ARG, "triangle references a missing vertex");
  for (int64_t k = 0; k < 3 * n_out; ++k)
    if (out_tri[k] < 0 || out_tri[k] >= K) return fail(ACVD_ERR_ARG, "output triangle references a missing cluster");
  std::vector<uint8_t> in_ok;
  vertex_manifold(nv, nf, tri, &in_ok);
  VertexTopology topo;
  vertex_topology(nv, nf, tri, &topo);
  std::vector<int32_t> fr;
  int32_t Kn = K;
  std::vector<int32_t> lab(labels, labels + nv);
  const int64_t n = force_manifold_step(nv, topo, in_ok, n_out, out_tri, &Kn, lab.data(), &fr, false);
  if (frozen && int64_t(fr.size()) > frozen_capacity) return fail(ACVD_ERR_ARG, "frozen buffer too small");
  std::memcpy(labels, lab.data(), size_t(nv) * sizeof(int32_t));
  if (frozen) std::memcpy(frozen, fr.data(), fr.size() * sizeof(int32_t));
  *n_clusters = Kn;
  *issues = n;
  return ACVD_OK;
}

int acvd_force_manifold(acvd_ctx* ctx, int32_t mesh_id, int32_t max_restarts, int32_t* restarts,
                        int64_t* remaining_issues) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (ctx->meshes.size() != 1) return fail(ACVD_ERR_ARG, "ForceManifold runs on a context holding one mesh");
  if (max_restarts < 0) return fail(ACVD_ERR_ARG, "negative restart limit");
  HostMesh& h = ctx->meshes[0];
  if (!ctx->built || !h.clustered) return fail(ACVD_ERR_STATE, "no clustering computed");
  const bool verbose = h.params.console_output > 0;
  std::vector<uint8_t> in_ok;
  vertex_manifold(h.nv, h.nf, h.tri.data(), &in_ok);
  VertexTopology topo;
  vertex_topology(h.nv, h.nf, h.tri.data(), &topo);
  const acvd_params user = h.params;
  std::vector<int32_t> user_init;
  user_init.swap(h.init_labels);
  std::vector<int32_t> lab(size_t(h.nv));
  int32_t n = 0;
  int64_t issues = 0;
  int rc = ACVD_OK;
  for (;;) {
    rc = acvd_build_output(ctx, 0, 0);
    if (rc) break;
    rc = acvd_get_labels(ctx, 0, lab.data());
    if (rc) break;
    int32_t K = h.params.number_of_clusters;
    for (auto& l : lab)
      if (l == K) l = -1;  // null: out of range whatever K becomes
    std::vector<int32_t> frozen;
    issues = force_manifold_step(h.nv, topo, in_ok, int64_t(h.out_tri.size() / 3), h.out_tri.data(), &K,
                                 lab.data(), &frozen, verbose);
    if (issues == 0 || n >= max_restarts) break;
    if (verbose) std::printf("%lld topology issues, restarting minimization\n", (long long)issues);
    // MinimizeEnergy again from the current labels with ConnexityConstraint
    // = 0 (:941-943): the metric stays constrained, clusters added
    h.params.number_of_clusters = K;
    h.params.unconstrained_initialization = 0;
    h.params.write_energy_log = 0;
    h.init_labels = lab;
    h.frozen.swap(frozen);
    ctx->keep_frozen = true;
    ctx->invalidate();
    rc = acvd_process_clustering(ctx);
    ctx->keep_frozen = false;
    if (rc) break;
    ++n;
  }
  h.params.unconstrained_initialization = user.unconstrained_initialization;
  h.params.write_energy_log = user.write_energy_log;
  h.init_labels.swap(user_init);
  if (restarts) *restarts = n;
  if (remaining_issues) *remaining_issues = issues;
  return rc;
}

int acvd_reference_seed_sampling(int64_t nv, int32_t n_clusters, int32_t* labels) {
  if (nv <= 0 || !labels) return fail(ACVD_ERR_ARG, "empty mesh or null buffer");
  if (n_clusters <= 0 || int64_t(n_clusters) > nv) return fail(ACVD_ERR_ARG, "clusters must be 1..items");
  reference_seed_sampling(nv, n_clusters, labels);
  return ACVD_OK;
}

int acvd_mesh_boundary_fix(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, const int32_t* labels,
                           int32_t n_clusters, int64_t* nv_added, int64_t* nf_added, double* added_xyz,
                           int32_t* added_tri) {
  if (nv <= 0 || nf <= 0 || !xyz || !tri || !labels || !nv_added || !nf_added || n_clusters <= 0)
    return fail(ACVD_ERR_ARG, "empty mesh or null buffer");
  for (int64_t k = 0; k < 3 * nf; ++k)
    if (tri[k] < 0 || tri[k] >= nv) return fail(ACVD_ERR_ARG, "triangle references a missing vertex");
  std::vector<double> ox(3 * size_t(n_clusters), 0.0);
  std::vector<int32_t> ot;
  fix_mesh_boundaries(nv, xyz, nf, tri, labels, n_clusters, &ox, &ot);
  *nv_added = int64_t(ox.size() / 3) - n_clusters;
  *nf_added = int64_t(ot.size() / 3);
  if (added_xyz) std::memcpy(added_xyz, ox.data() + 3 * size_t(n_clusters), size_t(*nv_added) * 3 * sizeof(double));
  if (added_tri) std::memcpy(added_tri, ot.data(), ot.size() * sizeof(int32_t));
  return ACVD_OK;
}

int acvd_fix_mesh_boundaries(acvd_ctx* ctx, int32_t mesh_id, int32_t* changed) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  HostMesh& h = ctx->meshes[mesh_id];
  if (h.out_xyz.empty()) return fail(ACVD_ERR_STATE, "output mesh not built (acvd_build_output)");
  std::vector<int32_t> lab(size_t(h.nv));
  if (int rc = acvd_get_labels(ctx, mesh_id, lab.data())) return rc;
  const bool c = fix_mesh_boundaries(h.nv, h.xyz.data(), h.nf, h.tri.data(), lab.data(), h.params.number_of_clusters,
                                     &h.out_xyz, &h.out_tri);
  if (changed) *changed = c ? 1 : 0;
  return ACVD_OK;
}

int acvd_get_output_sizes(acvd_ctx* ctx, int32_t mesh_id, int64_t* nv, int64_t* nf) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  const HostMesh& h = ctx->meshes[mesh_id];
  if (h.out_xyz.empty()) return fail(ACVD_ERR_STATE, "output mesh not built (acvd_build_output)");
  *nv = int64_t(h.out_xyz.size() / 3);
  *nf = int64_t(h.out_tri.size() / 3);
  return ACVD_OK;
}

int acvd_get_output(acvd_ctx* ctx, int32_t mesh_id, double* xyz, int32_t* tri) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  const HostMesh& h = ctx->meshes[mesh_id];
  if (h.out_xyz.empty()) return fail(ACVD_ERR_STATE, "output mesh not built (acvd_build_output)");
  if (xyz) std::memcpy(xyz, h.out_xyz.data(), h.out_xyz.size() * sizeof(double));
  if (tri) std::memcpy(tri, h.out_tri.data(), h.out_tri.size() * sizeof(int32_t));
  return ACVD_OK;
}

int acvd_get_labels(acvd_ctx* ctx, int32_t mesh_id, int32_t* labels) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (!labels) return fail(ACVD_ERR_ARG, "null buffer");
  if (!ctx->built || !ctx->d_labels.p) return fail(ACVD_ERR_STATE, "no clustering computed");
  const HostMesh& h = ctx->meshes[mesh_id];
  HIP_TRY(hipMemcpy(labels, ctx->d_labels.p + h.item_off, h.nv * sizeof(int32_t), hipMemcpyDeviceToHost));
  const int32_t K = h.params.number_of_clusters;
  for (int64_t i = 0; i < h.nv; ++i) labels[i] = labels[i] == kNull ? K : int32_t(labels[i] - h.clus_off);
  return ACVD_OK;
}

int acvd_get_cluster_sizes(acvd_ctx* ctx, int32_t mesh_id, int32_t* sizes) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (!sizes) return fail(ACVD_ERR_ARG, "null buffer");
  if (!ctx->built) return fail(ACVD_ERR_STATE, "no clustering computed");
  const HostMesh& h = ctx->meshes[mesh_id];
  HIP_TRY(hipMemcpy(sizes, ctx->d_sizes.p + h.clus_off, h.params.number_of_clusters * sizeof(int32_t),
                    hipMemcpyDeviceToHost));
  return ACVD_OK;
}

int acvd_get_cluster_points(acvd_ctx* ctx, int32_t mesh_id, double* xyz) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (!xyz) return fail(ACVD_ERR_ARG, "null buffer");
  if (!ctx->built) return fail(ACVD_ERR_STATE, "no clustering computed");
  const HostMesh& h = ctx->meshes[mesh_id];
  DBuf<double> d;
  HIP_TRY(d.alloc(3 * size_t(h.params.number_of_clusters)));
  HIP_TRY(launch_cluster_points(ctx->stream, ctx->args(), mesh_id, d.p));
  HIP_TRY(hipMemcpyAsync(xyz, d.p, 3 * size_t(h.params.number_of_clusters) * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  exact_anchor_points(h, xyz);
  return ACVD_OK;
}

int acvd_get_energy(acvd_ctx* ctx, int32_t mesh_id, double* energy) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (!energy) return fail(ACVD_ERR_ARG, "null buffer");
  if (!ctx->built) return fail(ACVD_ERR_STATE, "no clustering computed");
  hipStream_t s = ctx->stream;
  DBuf<double> pts, e;
  HIP_TRY(pts.alloc(3 * size_t(ctx->K)));
  HIP_TRY(e.alloc(ctx->N));
  for (int m = 0; m < int(ctx->meshes.size()); ++m)
    HIP_TRY(launch_cluster_points(s, ctx->args(), m, pts.p + 3 * ctx->meshes[m].clus_off));
  HIP_TRY(launch_item_energy(s, ctx->args(), pts.p, e.p));
  const HostMesh& h = ctx->meshes[mesh_id];
  std::vector<double> he(h.nv);
  HIP_TRY(hipMemcpyAsync(he.data(), e.p + h.item_off, h.nv * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  // fixed-order pairwise sum on the host (diagnostic only)
  while (he.size() > 1) {
    size_t half = (he.size() + 1) / 2;
    for (size_t k = 0; k + half < he.size(); ++k) he[k] += he[k + half];
    he.resize(half);
  }
  *energy = he.empty() ? 0.0 : he[0];
  return ACVD_OK;
}

int acvd_get_energy_log(acvd_ctx* ctx, int32_t mesh_id, int64_t capacity, int64_t* n_loops, double* seconds,
                        double* energy, double* final_energy) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (!n_loops || capacity < 0) return fail(ACVD_ERR_ARG, "null count or negative capacity");
  const HostMesh& h = ctx->meshes[mesh_id];
  if (!h.logged || !h.clustered) return fail(ACVD_ERR_STATE, "no energy log recorded (write_energy_log off)");
  const int64_t n = int64_t(h.log_energy.size());
  *n_loops = n;
  const int64_t w = std::min(n, capacity);
  if (seconds) std::copy(h.log_seconds.begin(), h.log_seconds.begin() + w, seconds);
  if (energy) std::copy(h.log_energy.begin(), h.log_energy.begin() + w, energy);
  if (final_energy) *final_energy = h.log_final;
  return ACVD_OK;
}

int acvd_get_stats(acvd_ctx* ctx, int32_t mesh_id, acvd_stats* st) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (!st) return fail(ACVD_ERR_ARG, "null buffer");
  *st = ctx->meshes[mesh_id].stats;
  return ACVD_OK;
}

int acvd_get_adjacency_size(acvd_ctx* ctx, int32_t mesh_id, int64_t* n_adj) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (!ctx->built) return fail(ACVD_ERR_STATE, "device inputs not built");
  *n_adj = ctx->meshes[mesh_id].n_adj;
  return ACVD_OK;
}

int acvd_get_adjacency(acvd_ctx* ctx, int32_t mesh_id, int32_t* row_ptr, int32_t* col) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (!ctx->built) return fail(ACVD_ERR_STATE, "device inputs not built");
  const HostMesh& h = ctx->meshes[mesh_id];
  HIP_TRY(hipMemcpy(row_ptr, ctx->d_row_ptr.p + h.item_off, (h.nv + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i <= h.nv; ++i) row_ptr[i] -= int32_t(h.adj_off);
  HIP_TRY(hipMemcpy(col, ctx->d_col.p + h.adj_off, h.n_adj * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int64_t k = 0; k < h.n_adj; ++k) col[k] -= int32_t(h.item_off);
  return ACVD_OK;
}

int acvd_get_items(acvd_ctx* ctx, int32_t mesh_id, float* xw, float* q) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (!ctx->built) return fail(ACVD_ERR_STATE, "device inputs not built");
  const HostMesh& h = ctx->meshes[mesh_id];
  if (xw) HIP_TRY(hipMemcpy(xw, ctx->d_xw.p + h.item_off, h.nv * sizeof(float4), hipMemcpyDeviceToHost));
  if (q) HIP_TRY(hipMemcpy(q, ctx->d_q.p + 10 * h.item_off, 10 * h.nv * sizeof(float), hipMemcpyDeviceToHost));
  return ACVD_OK;
}

int acvd_get_scales(acvd_ctx* ctx, int32_t mesh_id, double* scales3) {
  if (int rc = check_mesh(ctx, mesh_id)) return rc;
  if (!ctx->built) return fail(ACVD_ERR_STATE, "device inputs not built");
  const Scales& s = ctx->h_meshes[mesh_id].scales;
  scales3[0] = s.sw;
  scales3[1] = s.ss;
  scales3[2] = s.sq;
  return ACVD_OK;
}

}  // extern "C"

// ====================================================== fixed-assignment energy
#include "acvd_energy.h"

struct acvd_energy {
  int device = 0;
  struct Mesh {
    int64_t nv, nf;
    std::vector<int32_t> tri;
    int32_t K, metric, level;
  };
  std::vector<Mesh> meshes;
  bool built = false;
  int64_t N = 0, K = 0, F = 0;
  DBuf<EnergyMesh> d_meshes;
  DBuf<int32_t> d_item_mesh, d_clus_mesh, d_tri, d_vfptr, d_vfcol;
  DBuf<double> d_acc, d_pts, d_adj;
  bool have_forward = false;
  EnergyArgs args() const {
    EnergyArgs a{};
    a.n_items = N;
    a.n_clusters = K;
    a.meshes = d_meshes.p;
    a.item_mesh = meshes.size() > 1 ? d_item_mesh.p : nullptr;
    a.clus_mesh = meshes.size() > 1 ? d_clus_mesh.p : nullptr;
    a.tri = d_tri.p;
    a.vfptr = d_vfptr.p;
    a.vfcol = d_vfcol.p;
    a.acc = d_acc.p;
    a.pts = d_pts.p;
    a.adj = d_adj.p;
    return a;
  }
};

namespace {

int energy_build(acvd_energy* e) {
  if (e->built) return ACVD_OK;
  if (e->meshes.empty()) return fail(ACVD_ERR_STATE, "no mesh added");
  HIP_TRY(hipSetDevice(e->device));
  const int M = int(e->meshes.size());
  int64_t N = 0, K = 0, F = 0;
  for (auto& m : e->meshes) {
    N += m.nv;
    K += m.K;
    F += m.nf;
  }
  if (N >= (int64_t(1) << 31) - 1 || K >= (int64_t(1) << 31) - 1 || 3 * F >= (int64_t(1) << 31))
    return fail(ACVD_ERR_LIMIT, "energy context exceeds 2^31 items, clusters or face corners");
  std::vector<EnergyMesh> hm(M);
  std::vector<int32_t> tri(3 * F), item_mesh, clus_mesh, vfptr(N + 1, 0), vfcol(3 * F);
  if (M > 1) {
    item_mesh.resize(N);
    clus_mesh.resize(K);
  }
  int64_t io = 0, co = 0, fo = 0;
  for (int m = 0; m < M; ++m) {
    const auto& h = e->meshes[m];
    hm[m] = EnergyMesh{io, h.nv, co, fo, h.K, h.metric, h.level, 0};
    for (int64_t k = 0; k < 3 * h.nf; ++k) tri[3 * fo + k] = int32_t(h.tri[k] + io);
    // vertex -> faces (mesh-local face ids), faces by increasing id
    for (int64_t k = 0; k < 3 * h.nf; ++k) vfptr[io + h.tri[k] + 1]++;
    if (M > 1) {
      std::fill(item_mesh.begin() + io, item_mesh.begin() + io + h.nv, m);
      std::fill(clus_mesh.begin() + co, clus_mesh.begin() + co + h.K, m);
    }
    io += h.nv;
    co += h.K;
    fo += h.nf;
  }
  for (int64_t i = 0; i < N; ++i) vfptr[i + 1] += vfptr[i];
  {
    std::vector<int32_t> fill(vfptr.begin(), vfptr.end() - 1);
    for (int m = 0; m < M; ++m) {
      const auto& h = e->meshes[m];
      for (int64_t f = 0; f < h.nf; ++f)
        for (int j = 0; j < 3; ++j) vfcol[fill[hm[m].item_off + h.tri[3 * f + j]]++] = int32_t(f);
    }
  }
  HIP_TRY(e->d_meshes.alloc(M));
  HIP_TRY(hipMemcpy(e->d_meshes.p, hm.data(), M * sizeof(EnergyMesh), hipMemcpyHostToDevice));
  if (M > 1) {
    HIP_TRY(e->d_item_mesh.alloc(N));
    HIP_TRY(e->d_clus_mesh.alloc(K));
    HIP_TRY(hipMemcpy(e->d_item_mesh.p, item_mesh.data(), N * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_clus_mesh.p, clus_mesh.data(), K * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  HIP_TRY(e->d_tri.alloc(3 * F));
  HIP_TRY(hipMemcpy(e->d_tri.p, tri.data(), 3 * F * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_TRY(e->d_vfptr.alloc(N + 1));
  HIP_TRY(hipMemcpy(e->d_vfptr.p, vfptr.data(), (N + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_TRY(e->d_vfcol.alloc(3 * F));
  HIP_TRY(hipMemcpy(e->d_vfcol.p, vfcol.data(), 3 * F * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_TRY(e->d_acc.alloc(13 * K));
  HIP_TRY(e->d_pts.alloc(6 * K));
  HIP_TRY(e->d_adj.alloc(12 * K));
  e->N = N;
  e->K = K;
  e->F = F;
  e->built = true;
  return ACVD_OK;
}

}  // namespace

extern "C" {

int acvd_energy_create(acvd_energy** out, int device) {
  if (!out) return fail(ACVD_ERR_ARG, "null output pointer");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(ACVD_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= n) return fail(ACVD_ERR_ARG, "device index out of range");
  auto* e = new (std::nothrow) acvd_energy();
  if (!e) return fail(ACVD_ERR_ALLOC, "out of host memory");
  e->device = device;
  *out = e;
  return ACVD_OK;
}

int acvd_energy_destroy(acvd_energy* e) {
  if (!e) return fail(ACVD_ERR_ARG, "null energy context");
  (void)hipSetDevice(e->device);
  delete e;
  return ACVD_OK;
}

int acvd_energy_add_mesh(acvd_energy* e, int64_t nv, int64_t nf, const int32_t* tri, int32_t n_clusters,
                         int32_t metric, int32_t level, int32_t* mesh_id) {
  if (!e) return fail(ACVD_ERR_ARG, "null energy context");
  if (nv <= 0 || nf <= 0 || !tri) return fail(ACVD_ERR_ARG, "empty mesh or null buffer");
  if (n_clusters <= 0) return fail(ACVD_ERR_ARG, "number of clusters must be positive");
  if (metric != ACVD_METRIC_ISOTROPIC && metric != ACVD_METRIC_QEM) return fail(ACVD_ERR_ARG, "unknown metric");
  if (level < 0 || level > 3) return fail(ACVD_ERR_ARG, "quadrics level must be 0..3");
  for (int64_t k = 0; k < 3 * nf; ++k)
    if (tri[k] < 0 || tri[k] >= nv) return fail(ACVD_ERR_ARG, "triangle references a missing vertex");
  acvd_energy::Mesh m;
  m.nv = nv;
  m.nf = nf;
  m.tri.assign(tri, tri + 3 * nf);
  m.K = n_clusters;
  m.metric = metric;
  m.level = level;
  e->meshes.push_back(std::move(m));
  e->built = false;
  e->have_forward = false;
  if (mesh_id) *mesh_id = int32_t(e->meshes.size() - 1);
  return ACVD_OK;
}

int acvd_energy_num_items(const acvd_energy* e, int64_t* n_items, int32_t* n_meshes) {
  if (!e) return fail(ACVD_ERR_ARG, "null energy context");
  int64_t n = 0;
  for (auto& m : e->meshes) n += m.nv;
  if (n_items) *n_items = n;
  if (n_meshes) *n_meshes = int32_t(e->meshes.size());
  return ACVD_OK;
}

int acvd_energy_forward(acvd_energy* e, const float* xyz, const float* w, const int32_t* labels, double* energy,
                        void* stream) {
  if (!e) return fail(ACVD_ERR_ARG, "null energy context");
  if (!xyz || !w || !labels || !energy) return fail(ACVD_ERR_ARG, "null device buffer");
  if (int rc = energy_build(e)) return rc;
  HIP_TRY(hipSetDevice(e->device));
  EnergyArgs a = e->args();
  a.xyz = xyz;
  a.w = w;
  a.labels = labels;
  HIP_TRY(energy_forward(static_cast<hipStream_t>(stream), a, energy, int(e->meshes.size())));
  e->have_forward = true;
  return ACVD_OK;
}

int acvd_energy_forward_cvt_noalloc(int device, int64_t n, int32_t n_clusters, const float* xyz, const float* w,
                                    const int32_t* labels, double* energy, void* stream) {
  if (n < 0 || n >= (int64_t(1) << 31) - 1) return fail(ACVD_ERR_LIMIT, "item count must be in 0..2^31-2");
  if (n_clusters <= 0) return fail(ACVD_ERR_ARG, "number of clusters must be positive");
  if (n_clusters > kNoAllocMaxClusters)
    return fail(ACVD_ERR_LIMIT, "acvd_energy_forward_cvt_noalloc takes at most 65536 clusters (its static scratch)");
  if ((n > 0 && (!xyz || !w || !labels)) || !energy) return fail(ACVD_ERR_ARG, "null device buffer");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(ACVD_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= count) return fail(ACVD_ERR_ARG, "device index out of range");
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(energy_forward_cvt_noalloc(static_cast<hipStream_t>(stream), n, n_clusters, xyz, w, labels, energy));
  return ACVD_OK;
}

int acvd_energy_backward(acvd_energy* e, const float* xyz, const float* w, const int32_t* labels,
                         const double* grad_energy, float* dxyz, float* dw, void* stream) {
  if (!e) return fail(ACVD_ERR_ARG, "null energy context");
  if (!e->have_forward) return fail(ACVD_ERR_STATE, "backward needs a forward on the same inputs first");
  if (!xyz || !w || !labels || !grad_energy || !dxyz || !dw) return fail(ACVD_ERR_ARG, "null device buffer");
  HIP_TRY(hipSetDevice(e->device));
  EnergyArgs a = e->args();
  a.xyz = xyz;
  a.w = w;
  a.labels = labels;
  HIP_TRY(energy_backward(static_cast<hipStream_t>(stream), a, grad_energy, dxyz, dw));
  return ACVD_OK;
}

}  // extern "C"
