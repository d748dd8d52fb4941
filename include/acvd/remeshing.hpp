// C++ drop-in for the reference's remeshing classes, over the C-ABI of
// libacvd_amd.so (include/acvd_c.h).  VTK-free: a Mesh value type replaces
// vtkSurface, an int32 label vector replaces vtkIntArray.
//
//   reference (DiscreteRemeshing/)                         here
//   vtkIsotropicDiscreteRemeshing (vtkIsotropicDiscreteRemeshing.h:42-69)  acvd::IsotropicDiscreteRemeshing
//   vtkQIsotropicDiscreteRemeshing (:71-102)                acvd::QIsotropicDiscreteRemeshing
//   ::New() / ->Delete()                                    New() / Delete()
//   SetInput(vtkSurface*) (vtkSurfaceClustering.h:178-211)  SetInput(const Mesh*)
//   SetNumberOfClusters (vtkUniformClustering.h:62-69)      same
//   SetConsoleOutput (:95), SetUnconstrainedInitialization (:81),
//   SetMaxNumberOfLoops (:103), SetMaxNumberOfConvergences (:99),
//   SetInitialClustering (:111)                             same
//   SetFixedClusters (:134) + GetCluster(i)->AnchorItem
//     (ACVDQ.cxx:327-334)                                   SetFixedClusters(const std::vector<int32_t>*)
//                                                           + GetAnchorItem(i)
//   GetMetric()->SetGradation / SetQuadricsOptimizationLevel
//     (vtkIsotropicMetricForClustering.h:73-76, vtkQEMetricForClustering.h:50-54)  same
//   SetSubsamplingThreshold (vtkDiscreteRemeshing.h:71)     same
//   SetWriteToGlobalEnergyLog (vtkUniformClustering.h:123),
//   SetOutputDirectory (:126)                               same (energy.txt written by ProcessClustering)
//   ProcessClustering() -> vtkIntArray* (vtkUniformClustering.h:75)
//                                                           ProcessClustering() -> const Labels*
//   GetClustering() (:78)                                   same (borrowed, owned by the object)
//   Remesh() (vtkDiscreteRemeshing.h:64, 877-954)           same (subdivides coarse inputs)
//   SetForceManifold (vtkDiscreteRemeshing.h:79)             same (at most 100 restarts)
//   SetBoundaryFixing (vtkDiscreteRemeshing.h:77)            same (FixMeshBoundaries after the output)
//   GetOutput() -> vtkSurface* (:61)                        GetOutput() -> const Mesh* (borrowed)
//
// Error behaviour follows the reference: no exceptions; a failing call
// returns 0 / nullptr and, with console output on, prints the reason
// (vtkUniformClustering.h:659-664).  GetLastError() returns it as well.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../acvd_c.h"

namespace acvd {

struct Mesh {
  std::vector<double> points;      // 3 per vertex
  std::vector<int32_t> triangles;  // 3 per face
  int64_t GetNumberOfPoints() const { return int64_t(points.size() / 3); }
  int64_t GetNumberOfCells() const { return int64_t(triangles.size() / 3); }
  void GetPoint(int64_t i, double p[3]) const {
    for (int k = 0; k < 3; ++k) p[k] = points[3 * i + k];
  }
};

using Labels = std::vector<int32_t>;  // 0..K-1, K = null

template <int METRIC>
class DiscreteRemeshing {
 public:
  class MetricHandle {
   public:
    void SetGradation(double g) { p_->gradation = g; }
    double GetGradation() const { return p_->gradation; }
    void SetQuadricsOptimizationLevel(int l) { p_->quadrics_optimization_level = l; }
    int GetQuadricsOptimizationLevel() const { return p_->quadrics_optimization_level; }
    // per-vertex curvature indicator (vtkIsotropicMetricForClustering.h:66,
    // vtkQEMetricForClustering.h:75 take it as a vtkDataArrayCollection)
    void SetCurvatureInfo(const std::vector<double>* ind) { *ind_ = ind; }

   private:
    friend class DiscreteRemeshing;
    MetricHandle(acvd_params* p, const std::vector<double>** ind) : p_(p), ind_(ind) {}
    acvd_params* p_;
    const std::vector<double>** ind_;
  };

  static DiscreteRemeshing* New() { return new DiscreteRemeshing(); }
  void Delete() { delete this; }

  void SetInput(const Mesh* m) { input_ = m; }
  const Mesh* GetInput() const { return input_; }
  void SetNumberOfClusters(int k) { params_.number_of_clusters = k; }
  int GetNumberOfClusters() const { return params_.number_of_clusters; }
  void SetConsoleOutput(int v) { params_.console_output = v; }
  void SetUnconstrainedInitialization(int v) { params_.unconstrained_initialization = v; }
  void SetMaxNumberOfLoops(int v) { params_.max_number_of_loops = v; }
  void SetMaxNumberOfConvergences(int v) { params_.max_number_of_convergences = v; }
  void SetInitialClustering(const Labels* l) { initial_ = l; }
  // cluster i is fixed at item (*ids)[i]: it starts there, keeps it and, with
  // the QEM metric, takes its position as the cluster point (the
  // AnchorItem ACVDQ sets to the same item); NumberOfClusters counts them
  void SetFixedClusters(const std::vector<int32_t>* ids) { fixed_ = ids; }
  int64_t GetAnchorItem(int i) const {
    return fixed_ && i >= 0 && size_t(i) < fixed_->size() ? (*fixed_)[size_t(i)] : -1;
  }
  void SetSubsamplingThreshold(int t) { subsampling_threshold_ = t; }
  // energy.txt in the output directory (prefixed as given, no separator
  // added, vtkUniformClustering.h:670-675): "loop seconds energy" per loop,
  // then "Final Energy :<energy>" (:695-699, 1383-1395)
  void SetWriteToGlobalEnergyLog(int v) { params_.write_energy_log = v; }
  void SetOutputDirectory(const char* d) { output_directory_ = d ? d : ""; }
  // SetInitialSamplingType (vtkUniformClustering.h:108, 1123-1175): 1 the
  // region growing (default), 0 one random item per cluster, -1 item k for
  // cluster k, 2 the SetInitialClustering labels
  void SetInitialSamplingType(int t) { params_.initial_sampling = t == 0 ? 2 : (t == -1 ? 3 : 0); }
  void SetSeed(uint32_t s) { params_.seed = s; }
  // accepted for drop-in use (vtkThreadedClustering.h:58-74, ACVDP / ACVDQP;
  // SetDisplay, vtkUniformClustering.h:87): the device runs the sweep and
  // nothing is rendered
  void SetNumberOfThreads(int) {}
  void SetPoolingRatio(int) {}
  void SetDisplay(int) {}
  void SetDevice(int d) { device_ = d; }
  MetricHandle* GetMetric() { return &metric_; }
  int64_t GetNumberOfItems() const { return input_ ? input_->GetNumberOfPoints() : 0; }
  int GetClusteringType() const { return 1; }  // items are vertices
  const std::string& GetLastError() const { return error_; }

  // ProcessClustering (vtkUniformClustering.h:654-714).  Returns the labels
  // (borrowed) or nullptr when K == 0 or on error.  curvature_indicator = 1
  // (set by Remesh) weights the items with the reference's curvature
  // indicator when the gradation is positive and no indicator was set.
  const Labels* ProcessClustering(int curvature_indicator = 0) {
    if (params_.number_of_clusters <= 0) return fail("NumberOfClusters is 0"), nullptr;
    if (!input_) return fail("no input"), nullptr;
    if (!ensure_ctx()) return nullptr;
    const int64_t nv = input_->GetNumberOfPoints(), nf = input_->GetNumberOfCells();
    if (!ok(acvd_set_input_f64(ctx_, nv, input_->points.data(), nf, input_->triangles.data()))) return nullptr;
    acvd_params p = params_;
    p.metric = METRIC;
    p.curvature_indicator = curvature_indicator;
    if (!ok(acvd_set_params(ctx_, 0, &p))) return nullptr;
    if (indicator_ && !ok(acvd_set_indicator(ctx_, 0, indicator_->data()))) return nullptr;
    if (initial_ && !ok(acvd_set_initial_labels(ctx_, 0, initial_->data()))) return nullptr;
    if (!ok(acvd_set_fixed_clusters(ctx_, 0, fixed_ ? int64_t(fixed_->size()) : 0,
                                    fixed_ && !fixed_->empty() ? fixed_->data() : nullptr)))
      return nullptr;
    if (!ok(acvd_process_clustering(ctx_))) return nullptr;
    clustering_.resize(size_t(nv));
    if (!ok(acvd_get_labels(ctx_, 0, clustering_.data()))) return nullptr;
    clustered_ = true;
    if (params_.write_energy_log && !write_energy_log()) return nullptr;
    return &clustering_;
  }
  const Labels* GetClustering() const { return clustered_ ? &clustering_ : nullptr; }

  // Remesh (vtkDiscreteRemeshing.h:877-954): SamplingPreProcessing (the
  // curvature indicator when the gradation is positive and none was set,
  // :583-653), clustering, then the output mesh.  Returns 1 on success, 0 on
  // failure.
  // CheckSubsamplingRatio (vtkDiscreteRemeshing.h:840-875): while the input
  // has fewer than SubsamplingThreshold * K vertices it is replaced by its
  // midpoint subdivision (acvd_subdivide_input, on the device), which
  // GetInput() then returns; the curvature indicator is computed on the
  // original input and interpolated to the midpoints as the mean of their
  // parents (:700-745).
  int Remesh() {
    if (!input_) return fail("no input"), 0;
    const int64_t min_vertices = int64_t(subsampling_threshold_) * params_.number_of_clusters;
    if (input_->GetNumberOfPoints() >= min_vertices) {
      if (!ProcessClustering(1)) return 0;
      return Finish();
    }
    if (initial_) return fail("the initial clustering labels the input's vertices; the input needs subdivision"), 0;
    acvd_ctx* c = nullptr;
    if (!ok(acvd_create(&c, device_))) return 0;
    const bool good = subdivide(c, min_vertices);
    acvd_destroy(c);
    if (!good) return 0;
    original_ = input_;
    input_ = &subdivided_;
    if (!interpolated_.empty()) indicator_ = &interpolated_;
    if (!ProcessClustering(0)) return 0;
    return Finish();
  }
  // ForceManifold (vtkDiscreteRemeshing.h:937-946): restarts with added
  // clusters while the output has non-manifold vertices (acvd_force_manifold)
  void SetForceManifold(bool v) { force_manifold_ = v; }
  // FixMeshBoundaries (vtkDiscreteRemeshing.h:425-579) after the output
  void SetBoundaryFixing(int v) { boundary_fixing_ = v; }
  int GetNumberOfRestarts() const { return restarts_; }
  // the input Remesh() replaced by its subdivision, or nullptr
  const Mesh* GetOriginalInput() const { return original_; }
  // Output mesh with vertices moved by the quadric placement of
  // Examples/ACVD.cxx:217-291 at `level` (0 = cluster points).
  // With boundary fixing on, the boundary vertices and triangles are added
  // after the cluster points (the reference's post-placement moves only the
  // cluster points, Examples/ACVD.cxx:270-275, and keeps them).
  int BuildOutput(int level) {
    if (!clustered_) return fail("no clustering computed"), 0;
    if (!ok(acvd_build_output(ctx_, 0, level))) return 0;
    if (boundary_fixing_) {
      int32_t changed = 0;
      if (!ok(acvd_fix_mesh_boundaries(ctx_, 0, &changed))) return 0;
    }
    return FetchOutput();
  }
  // the output as last built (and fixed along the boundaries)
  int FetchOutput() {
    int64_t nv = 0, nf = 0;
    if (!ok(acvd_get_output_sizes(ctx_, 0, &nv, &nf))) return 0;
    output_.points.resize(size_t(3 * nv));
    output_.triangles.resize(size_t(3 * nf));
    if (!ok(acvd_get_output(ctx_, 0, output_.points.data(), output_.triangles.data()))) return 0;
    return 1;
  }
  const Mesh* GetOutput() const { return clustered_ ? &output_ : nullptr; }

  int GetStats(acvd_stats* st) const { return ctx_ ? acvd_get_stats(ctx_, 0, st) == ACVD_OK : 0; }

 protected:
  DiscreteRemeshing() : metric_(&params_, &indicator_) { acvd_params_init(&params_); }
  ~DiscreteRemeshing() {
    if (ctx_) acvd_destroy(ctx_);
  }
  DiscreteRemeshing(const DiscreteRemeshing&) = delete;
  DiscreteRemeshing& operator=(const DiscreteRemeshing&) = delete;

 private:
  // the input (and the indicator it is weighted with) subdivided in context c
  bool subdivide(acvd_ctx* c, int64_t min_vertices) {
    const Mesh& in = *input_;
    if (!ok(acvd_set_input_f64(c, in.GetNumberOfPoints(), in.points.data(), in.GetNumberOfCells(),
                               in.triangles.data())))
      return false;
    std::vector<double> ind;
    if (indicator_) {
      ind = *indicator_;
    } else if (params_.gradation > 0) {  // curvature of the original input
      acvd_params p = params_;
      p.metric = METRIC;
      p.number_of_clusters = 1;
      p.curvature_indicator = 1;
      ind.resize(size_t(in.GetNumberOfPoints()));
      if (!ok(acvd_set_params(c, 0, &p)) || !ok(acvd_build(c)) || !ok(acvd_get_indicator(c, 0, ind.data())))
        return false;
    }
    if (!ind.empty() && !ok(acvd_set_indicator(c, 0, ind.data()))) return false;
    int32_t levels = 0;
    if (!ok(acvd_subdivide_input(c, 0, min_vertices, &levels))) return false;
    if (params_.console_output)
      for (int32_t k = 0; k < levels; ++k) std::printf("Subdividing mesh\n");
    int64_t nv = 0, nf = 0;
    if (!ok(acvd_get_input_sizes(c, 0, &nv, &nf))) return false;
    subdivided_.points.resize(size_t(3 * nv));
    subdivided_.triangles.resize(size_t(3 * nf));
    if (!ok(acvd_get_input(c, 0, subdivided_.points.data(), subdivided_.triangles.data()))) return false;
    interpolated_.clear();
    if (!ind.empty()) {
      interpolated_.resize(size_t(nv));
      if (!ok(acvd_get_indicator(c, 0, interpolated_.data()))) return false;
    }
    return true;
  }
  // The file of WriteToGlobalEnergyLog: loop, seconds (6 significant
  // digits), energy (15); a file that cannot be opened is skipped, as the
  // reference's ofstream does.
  bool write_energy_log() {
    int64_t n = 0;
    double fin = 0.0;
    if (!ok(acvd_get_energy_log(ctx_, 0, 0, &n, nullptr, nullptr, nullptr))) return false;
    std::vector<double> sec(static_cast<size_t>(n)), en(static_cast<size_t>(n));
    if (!ok(acvd_get_energy_log(ctx_, 0, n, &n, sec.data(), en.data(), &fin))) return false;
    const std::string name = output_directory_ + "energy.txt";
    std::FILE* f = std::fopen(name.c_str(), "w");
    if (!f) return true;
    for (int64_t i = 0; i < n; ++i) std::fprintf(f, "%lld %.6g %.15g\n", (long long)i, sec[size_t(i)], en[size_t(i)]);
    std::fprintf(f, "Final Energy :%.15g\n", fin);
    std::fclose(f);
    return true;
  }
  int Finish() {
    restarts_ = 0;
    if (force_manifold_) {
      int64_t left = 0;
      if (!ok(acvd_force_manifold(ctx_, 0, 100, &restarts_, &left))) return 0;
      acvd_params p;
      if (!ok(acvd_get_params(ctx_, 0, &p))) return 0;
      params_.number_of_clusters = p.number_of_clusters;
      clustering_.resize(size_t(input_->GetNumberOfPoints()));
      if (!ok(acvd_get_labels(ctx_, 0, clustering_.data()))) return 0;
    }
    return BuildOutput(0);
  }
  bool ensure_ctx() {
    if (ctx_) return true;
    return ok(acvd_create(&ctx_, device_));
  }
  bool ok(int rc) {
    if (rc == ACVD_OK) return true;
    fail(acvd_last_error());
    return false;
  }
  void fail(const std::string& msg) {
    error_ = msg;
    if (params_.console_output) std::fprintf(stdout, "%s\n", msg.c_str());
  }

  acvd_params params_{};
  MetricHandle metric_;
  const Mesh* input_ = nullptr;
  const Labels* initial_ = nullptr;
  const std::vector<int32_t>* fixed_ = nullptr;
  const Mesh* original_ = nullptr;
  Mesh subdivided_;
  std::vector<double> interpolated_;
  const std::vector<double>* indicator_ = nullptr;
  int subsampling_threshold_ = 10;
  std::string output_directory_;
  bool force_manifold_ = false;
  int boundary_fixing_ = 0;
  int32_t restarts_ = 0;
  int device_ = 0;
  acvd_ctx* ctx_ = nullptr;
  Labels clustering_;
  Mesh output_;
  bool clustered_ = false;
  std::string error_;
};

using IsotropicDiscreteRemeshing = DiscreteRemeshing<ACVD_METRIC_ISOTROPIC>;
using QIsotropicDiscreteRemeshing = DiscreteRemeshing<ACVD_METRIC_QEM>;

}  // namespace acvd
