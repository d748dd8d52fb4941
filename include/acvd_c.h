/* C-ABI of the MI355X vertex-clustering library (libacvd_amd.so).
 *
 * This is the drop-in boundary for the reference's clustering path:
 * vtkVerticesProcessing< vtkDiscreteRemeshing< Metric > > with Metric =
 * vtkIsotropicMetricForClustering (ACVD) or vtkQEMetricForClustering (ACVDQ),
 * i.e. vtkIsotropicDiscreteRemeshing / vtkQIsotropicDiscreteRemeshing
 * (DiscreteRemeshing/vtkIsotropicDiscreteRemeshing.h:42-102).  Plain
 * pointers and sizes only; every call returns ACVD_OK (0) or a negative
 * status, with a message available from acvd_last_error() (the reference
 * prints to cout and returns 0 / exits instead, vtkUniformClustering.h:659-664,
 * Common/vtkSurfaceBase.cxx:1248-1256).
 *
 * Threading: one context = one HIP stream; a context is not thread-safe.
 * Distinct contexts may be used concurrently (as distinct reference objects).
 *
 * Labels follow the reference convention: cluster ids 0..K-1 and the null
 * cluster K (vtkUniformClustering.h:1131-1132).
 *
 * A context holds one mesh (acvd_set_input_*) or a batch of meshes
 * (acvd_add_input_*), all clustered together by acvd_process_clustering().
 */
#ifndef ACVD_C_H
#define ACVD_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ACVD_OK 0
#define ACVD_ERR_ARG -1      /* invalid argument */
#define ACVD_ERR_STATE -2    /* call not valid in the current state */
#define ACVD_ERR_HIP -3      /* HIP runtime error (no device, launch failure...) */
#define ACVD_ERR_ALLOC -4    /* device or host allocation failed */
#define ACVD_ERR_LIMIT -5    /* input beyond a documented limit */

#define ACVD_METRIC_ISOTROPIC 0 /* vtkIsotropicMetricForClustering */
#define ACVD_METRIC_QEM 1       /* vtkQEMetricForClustering */

typedef struct acvd_ctx acvd_ctx;

/* Per-mesh clustering parameters.  Field <- reference setter:
 *   number_of_clusters          <- SetNumberOfClusters (vtkUniformClustering.h:62-69)
 *   metric                      <- template parameter Metric (vtkIsotropicDiscreteRemeshing.h:42-102)
 *   quadrics_optimization_level <- GetMetric()->SetQuadricsOptimizationLevel (vtkQEMetricForClustering.h:50-54), default 3 (:370)
 *   gradation                   <- GetMetric()->SetGradation (vtkIsotropicMetricForClustering.h:73-76, vtkQEMetricForClustering.h:86-90)
 *   unconstrained_initialization<- SetUnconstrainedInitialization (vtkUniformClustering.h:81)
 *   max_number_of_loops         <- SetMaxNumberOfLoops (:103), default 5000000 (:1407)
 *   max_number_of_convergences  <- SetMaxNumberOfConvergences (:99), default 1e9 (:1406)
 *   console_output              <- SetConsoleOutput (:95)
 *   seed                        <- (no counterpart) seed of the accept priorities and of the device sampling
 *   initial_sampling            <- InitSamples type (vtkUniformClustering.h:1123-1175): 0 = the reference's
 *                                  default region growing (type 1, :1177-1316, sequential, on the host),
 *                                  1 = weighted random seeds grown on the device (no counterpart; faster),
 *                                  2 = the reference's type 0 (:1126-1159): one item per cluster drawn with
 *                                  vtkMath::Random seeded 5000, the rest grown by the fill; 3 = its type -1
 *                                  (:1148): item k starts cluster k
 *   curvature_indicator         <- vtkDiscreteRemeshing::SamplingPreProcessing (vtkDiscreteRemeshing.h:583-653):
 *                                  1 = when gradation > 0 and no indicator is set, the build computes the
 *                                  reference's curvature indicator (vtkCurvatureMeasure, polynomial fit on
 *                                  the faces within two edges of each vertex, vtkCurvatureMeasure.cxx:188-509)
 *                                  on the device and weights the items with it; Remesh() of the class API
 *                                  sets it, as the reference's Remesh runs SamplingPreProcessing
 *   first_phase_rounds          <- (no counterpart) accept rounds per pass until the first convergence,
 *                                  1..12; 0 = the library's rule: 1 for meshes of at most 2^17 vertices
 *                                  (lower final energy), 2 up to 2^19, 12 above (fewer passes); later
 *                                  phases run 12
 *   write_energy_log            <- SetWriteToGlobalEnergyLog (vtkUniformClustering.h:123): nonzero = record the
 *                                  energy after every loop and after the run (acvd_get_energy_log)
 */
typedef struct acvd_params {
  int32_t number_of_clusters;
  int32_t metric;
  int32_t quadrics_optimization_level;
  int32_t unconstrained_initialization;
  int32_t max_number_of_loops;
  int32_t max_number_of_convergences;
  int32_t console_output;
  uint32_t seed;
  int32_t initial_sampling;
  int32_t curvature_indicator;
  int32_t first_phase_rounds;
  int32_t write_energy_log;
  double gradation;
} acvd_params;

/* Launch shape of a context (no reference counterpart: the reference's
 * nearest knob is vtkThreadedClustering's PoolingRatio / mesh slices,
 * DiscreteRemeshing/vtkThreadedClustering.h:58-74, 891-987).  None of the
 * knobs changes the labels; they trade launches, occupancy and host syncs:
 *   sweep_segments      evaluate / accept-round grid (blocks) cap, 1..2048 (default 1152)
 *   tile_on_divisor     the scan's tile sets are kept once a pass moves at most K / divisor
 *                       items, 1..1024 (default 16)
 *   chunk_passes        passes per captured graph launch, 1..16 (default 16)
 *   short_chunk_passes  passes per launch near the end of a phase, 1..chunk_passes (default 4) */
typedef struct acvd_tuning {
  int32_t sweep_segments;
  int32_t tile_on_divisor;
  int32_t chunk_passes;
  int32_t short_chunk_passes;
} acvd_tuning;

/* Statistics of the last acvd_process_clustering() for one mesh. */
typedef struct acvd_stats {
  int64_t loops;            /* passes run (NumberOfLoops) */
  int64_t convergences;     /* NumberOfConvergences */
  int64_t total_moves;      /* applied moves, all passes */
  int32_t last_disconnected;/* disconnected clusters found by the last clean */
  int32_t finished;         /* 1 when the mesh reached its stopping rule */
  /* context-wide counters (same for every mesh of a batch) */
  int64_t ctx_passes;       /* passes launched while any mesh was sweeping */
  int64_t ctx_evaluated;    /* boundary items evaluated */
  int64_t ctx_proposed;     /* candidate moves proposed */
  int64_t ctx_accepted;     /* candidate moves applied */
  double converge_ms;       /* device time from MinimizeEnergy start to the last mesh finishing */
  double build_ms;          /* device time of the CSR + item build */
  int64_t n_items;          /* items of this mesh */
  int64_t n_adjacency;      /* CSR entries of this mesh (2E) */
  /* per-kernel device time when timing is on (acvd_set_timing), context-wide:
   * [0] scan, [1] evaluate, [2] accept rounds (apply, re-evaluate, apply ...),
   * [3] controller; launches count one per pass for [2] */
  double kernel_ms[4];
  int64_t kernel_launches[4];
  /* accept-round work, context-wide, all passes (the re-registered,
   * re-evaluated and re-checked counts only with timing on,
   * acvd_set_timing; 0 otherwise) */
  int64_t ctx_rounds;       /* accept rounds that walked the candidate list */
  int64_t ctx_walked;       /* candidate entries read by the round kernels (dead ones included) */
  int64_t ctx_reeval_live;  /* live candidates registered again by re-evaluate rounds */
  int64_t ctx_reevaluated;  /* candidates evaluated again in full because one of their clusters moved */
  int64_t ctx_accepted_constrained; /* moves applied while mesh 0 was constrained (QEM derives solve) */
  int64_t ctx_rechecked;    /* stale candidates whose own move was checked again (the isotropic rule) */
} acvd_stats;

/* ---- library -------------------------------------------------------------- */
const char* acvd_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char* acvd_last_error(void);
/* Number of visible HIP devices (0 when none; never fails). */
int acvd_device_count(void);
/* Fills p with the reference defaults: K = 0, isotropic, level 3, gradation 0,
 * unconstrained initialisation off, max loops 5000000, max convergences 1e9,
 * console output 0, seed 0, the reference's initial sampling (0), the
 * library's default first-phase rounds (0). */
void acvd_params_init(acvd_params* p);

/* ---- context -------------------------------------------------------------- */
/* New() (vtkIsotropicDiscreteRemeshing.h:59-69, 84-94) on HIP device `device`. */
int acvd_create(acvd_ctx** out, int device);
/* Delete() (vtkUniformClustering.h:1421-1443). */
int acvd_destroy(acvd_ctx* ctx);

/* ---- input ---------------------------------------------------------------- */
/* SetInput(vtkSurface*) (vtkSurfaceClustering.h:58, 178-211): replaces the
 * context content by one triangle mesh.  xyz: 3*nv coordinates, tri: 3*nf
 * vertex ids.  The data is copied. */
int acvd_set_input_f64(acvd_ctx* ctx, int64_t nv, const double* xyz, int64_t nf, const int32_t* tri);
int acvd_set_input_f32(acvd_ctx* ctx, int64_t nv, const float* xyz, int64_t nf, const int32_t* tri);
/* Batching (no reference counterpart): appends one more mesh; *mesh_id gets
 * its index.  All meshes of a context are clustered by one call. */
int acvd_add_input_f64(acvd_ctx* ctx, int64_t nv, const double* xyz, int64_t nf, const int32_t* tri,
                       int32_t* mesh_id);
int acvd_add_input_f32(acvd_ctx* ctx, int64_t nv, const float* xyz, int64_t nf, const int32_t* tri,
                       int32_t* mesh_id);
int acvd_num_meshes(const acvd_ctx* ctx);

/* Parameters of mesh `mesh_id` (-1: every mesh). */
int acvd_set_params(acvd_ctx* ctx, int32_t mesh_id, const acvd_params* p);
int acvd_get_params(const acvd_ctx* ctx, int32_t mesh_id, acvd_params* p);
/* Metric::SetCurvatureInfo (vtkQEMetricForClustering.h:75-83,
 * vtkIsotropicMetricForClustering.h:66-72): per-vertex indicator used as
 * weight = area * indicator^gradation (:329-336 / :256-259).  NULL clears. */
int acvd_set_indicator(acvd_ctx* ctx, int32_t mesh_id, const double* indicator);
/* The indicator the last build weighted mesh `mesh_id` with: the one set by
 * acvd_set_indicator, else the curvature indicator computed by the build
 * (params.curvature_indicator; the CurvatureCollection of
 * vtkCurvatureMeasure::GetCurvatureIndicator, vtkCurvatureMeasure.cxx:721-760).
 * out[nv].  ACVD_ERR_STATE when there is neither. */
int acvd_get_indicator(acvd_ctx* ctx, int32_t mesh_id, double* out);
/* SetInitialClustering(vtkIntArray*) (vtkUniformClustering.h:336-342,
 * 1166-1169): labels in 0..K (K = null), one per vertex.  NULL clears (the
 * library then seeds its own initial clustering on the device). */
int acvd_set_initial_labels(acvd_ctx* ctx, int32_t mesh_id, const int32_t* labels);

/* SetFixedClusters(vtkIdList*) (vtkUniformClustering.h:134) together with
 * GetCluster(i)->AnchorItem = items[i] for the QEM metric, as ACVDQ's -fv / -ft
 * options do (DiscreteRemeshing/Examples/ACVDQ.cxx:327-334): cluster i
 * (i < n) starts at item items[i] in the reference's initial sampling
 * (vtkUniformClustering.h:1186-1191), keeps the component holding it when
 * cleaned (:441-447), and with the QEM metric its point is that item's input
 * position, which the item never leaves (vtkQEMetricForClustering.h:196-199,
 * 259, 270-274).  number_of_clusters counts these clusters.  n = 0 clears. */
int acvd_set_fixed_clusters(acvd_ctx* ctx, int32_t mesh_id, int64_t n, const int32_t* items);

/* ---- run ------------------------------------------------------------------ */
/* ProcessClustering() (vtkUniformClustering.h:654-714) for every mesh of the
 * context: device build of CSR + items (first call after an input change),
 * initial sampling, then MinimizeEnergy (:724-830) with the boundary sweep on
 * the GPU.  Returns ACVD_ERR_ARG if a mesh has number_of_clusters <= 0. */
int acvd_process_clustering(acvd_ctx* ctx);

/* InitSamples (vtkUniformClustering.h:1123-1175) followed by the first
 * FillHolesInClustering of MinimizeEnergy (:727), without clustering: meshes
 * without user labels get the device's seeded initial clustering (weighted
 * random seeds), then null items are grown into their neighbours' clusters.
 * acvd_get_labels then returns this starting labelling. */
int acvd_initial_sampling(acvd_ctx* ctx);

/* Per-kernel timing (no counterpart in the reference): when enabled, passes
 * are launched one by one with HIP events around each sweep kernel (instead
 * of replaying the captured chunk graph) and the times are reported in
 * acvd_stats.kernel_ms / kernel_launches.  enable: 0 off; 1 events and the
 * accept-round counters (acvd_stats.ctx_reeval_live / ctx_reevaluated; the
 * counting kernels are a few per cent slower); 2 events only. */
int acvd_set_timing(acvd_ctx* ctx, int enable);

/* ---- results (copies into caller buffers) --------------------------------- */
/* GetClustering() (vtkUniformClustering.h:78): nv labels in 0..K. */
int acvd_get_labels(acvd_ctx* ctx, int32_t mesh_id, int32_t* labels);
/* GetClusterSize (vtkUniformClustering.h:132) for all K clusters. */
int acvd_get_cluster_sizes(acvd_ctx* ctx, int32_t mesh_id, int32_t* sizes);
/* GetMetric()->GetClusterCentroid for all K clusters (3*K doubles):
 * isotropic S/W (vtkIsotropicMetricForClustering.h:178-183); QEM the
 * rank-truncated quadric point (vtkQEMetricForClustering.h:263-285,
 * Common/vtkQuadricTools.cxx:168-177).  Empty clusters give NaN. */
int acvd_get_cluster_points(acvd_ctx* ctx, int32_t mesh_id, double* xyz);
/* Unshifted clustering energy sum_i w_i |x_i - p(c_i)|^2 of the current
 * labels, accumulated in FP64 on the device (diagnostic). */
int acvd_get_energy(acvd_ctx* ctx, int32_t mesh_id, double* energy);
/* Energy log of the last acvd_process_clustering() of a mesh whose params
 * had write_energy_log set: what WriteToGlobalEnergyLog writes after each loop
 * (vtkUniformClustering.h:1383-1395; the loop number is the entry's index) and
 * the "Final Energy :" line after the run (:695-699).  Each energy is the sum
 * of the cluster energies GetClusterEnergy returns (-|S|^2/W isotropic,
 * W|p|^2 - 2 p.S QEM; vtkIsotropicMetricForClustering.h:114-121,
 * vtkQEMetricForClustering.h:194-208), summed in FP64 in a fixed order (the
 * reference: long double, ComputeGlobalEnergy :1318-1346); seconds is the
 * device wall clock since the start of the minimisation (Time - StartTime).
 * *n_loops receives the number of entries; min(n_loops, capacity) of them are
 * written to seconds / energy (either may be null); final_energy may be null.
 * ACVD_ERR_STATE when the last run did not record a log for this mesh. */
int acvd_get_energy_log(acvd_ctx* ctx, int32_t mesh_id, int64_t capacity, int64_t* n_loops, double* seconds,
                        double* energy, double* final_energy);
int acvd_get_stats(acvd_ctx* ctx, int32_t mesh_id, acvd_stats* st);

/* The reference's default initial sampling (InitSamples type 1,
 * vtkUniformClustering.h:1123-1175 -> ComputeInitialRandomSampling
 * :1177-1316) of one mesh, computed on the host without a context or device:
 * labels[nv] receives 0..K-1 and K for items left unassigned.  indicator may
 * be NULL.  acvd_initial_sampling() uses it when initial_sampling == 0. */
int acvd_reference_initial_sampling(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, int32_t metric,
                                    const double* indicator, double gradation, int32_t n_clusters, int32_t* labels);

/* vtkSurface::Subdivide (Common/vtkSurface.cxx:605-677), the step Remesh()
 * repeats while the input has fewer than SubsamplingThreshold * K vertices
 * (vtkDiscreteRemeshing::CheckSubsamplingRatio, vtkDiscreteRemeshing.h:840-875):
 * one level of midpoint subdivision on the host.  The output holds the nv
 * input vertices, then one vertex per edge at its midpoint, edges in the order
 * the reference creates them (face by face, edges (v0,v1), (v1,v2), (v2,v0),
 * Common/vtkSurfaceBase.cxx:1166-1221), and four faces per input face:
 * (v1,m12,m31), (m12,v2,m23), (m23,v3,m31), (m12,m23,m31).
 * parents[2k], parents[2k+1] receive the ends of the edge of new vertex nv+k
 * (VerticesParent1/2).  Pass NULL output buffers to get the sizes first
 * (*out_nv = nv + edges, *out_nf = 4 nf); out_xyz[3*out_nv], out_tri[3*out_nf],
 * parents[2*(out_nv - nv)]. */
int acvd_subdivide(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, int64_t* out_nv, int64_t* out_nf,
                   double* out_xyz, int32_t* out_tri, int32_t* parents);

/* CheckSubsamplingRatio (vtkDiscreteRemeshing.h:840-875) on the device: while
 * mesh mesh_id has fewer than min_vertices vertices it is replaced by its
 * midpoint subdivision (the vertices, faces and order of acvd_subdivide, one
 * level at a time on the context's GPU), and an indicator set with
 * acvd_set_indicator is carried along, each midpoint taking the mean of its
 * two ends (vtkDiscreteRemeshing.h:700-745).  Fixed items keep their ids.
 * *levels receives the number of levels applied (0 when the mesh already has
 * enough vertices).  ACVD_ERR_STATE when initial labels are set (they label
 * the input's vertices), ACVD_ERR_LIMIT past 2^27 vertices.  The subdivided
 * mesh is what acvd_get_input() returns and what the next build clusters. */
int acvd_subdivide_input(acvd_ctx* ctx, int32_t mesh_id, int64_t min_vertices, int32_t* levels);

/* The input mesh mesh_id as the context holds it (vtkDiscreteRemeshing's
 * GetInput(), the subdivided mesh after acvd_subdivide_input): sizes, then
 * xyz[3*nv] and tri[3*nf] (either may be NULL). */
int acvd_get_input_sizes(const acvd_ctx* ctx, int32_t mesh_id, int64_t* nv, int64_t* nf);
int acvd_get_input(const acvd_ctx* ctx, int32_t mesh_id, double* xyz, int32_t* tri);

/* ---- output mesh ---------------------------------------------------------- */
/* BuildDelaunayTriangulation (vtkDiscreteRemeshing.h:1002-1152, vertex
 * clustering branch), i.e. what GetOutput() (:61) returns after Remesh():
 * K vertices at the cluster points (an empty cluster takes the first
 * non-empty cluster's point), one triangle per input face spanning three
 * distinct clusters, the first such face kept for each cluster triple.
 * quadrics_level > 0 then applies ACVD's quadric placement
 * (Examples/ACVD.cxx:217-291, level 1 by default there) to the vertices.
 * ForceManifold / boundary fixing (:937-952) are not part of this call. */
int acvd_build_output(acvd_ctx* ctx, int32_t mesh_id, int32_t quadrics_level);
int acvd_get_output_sizes(acvd_ctx* ctx, int32_t mesh_id, int64_t* nv, int64_t* nf);
/* xyz: 3*nv doubles, tri: 3*nf cluster ids (either may be NULL). */
int acvd_get_output(acvd_ctx* ctx, int32_t mesh_id, double* xyz, int32_t* tri);

/* ---- device inputs of the sweep (for replay and inspection) --------------- */
/* Number of CSR adjacency entries of the mesh (valid after a build). */
int acvd_get_adjacency_size(acvd_ctx* ctx, int32_t mesh_id, int64_t* n_adj);
/* CSR with mesh-local vertex ids: row_ptr[nv+1], col[n_adj], rows sorted. */
int acvd_get_adjacency(acvd_ctx* ctx, int32_t mesh_id, int32_t* row_ptr, int32_t* col);
/* Items: xw[4*nv] = (x, y, z, w) in FP32; q[10*nv] vertex quadrics (may be
 * NULL).  Build the device inputs without clustering with acvd_build(). */
int acvd_get_items(acvd_ctx* ctx, int32_t mesh_id, float* xw, float* q);
/* Fixed-point scales of the mesh: {weight, weighted position, quadric}. */
int acvd_get_scales(acvd_ctx* ctx, int32_t mesh_id, double* scales3);
/* Runs the device build (CSR, items, scales) without clustering. */
int acvd_build(acvd_ctx* ctx);

/* ---- launch-shape tuning -------------------------------------------------------
 * acvd_set_tuning / acvd_get_tuning: the context's launch shape (acvd_tuning);
 * a new segment count rebuilds the device state at the next run.
 * acvd_autotune: picks the launch shape for the context's current meshes,
 * parameters and initial labels by timing full clusterings (the best of
 * `repeats` runs each; coordinate descent over the knobs, a change kept when
 * it saves more than 1 %), checks that every shape gives the same labels
 * (ACVD_ERR_STATE otherwise), leaves the best shape set and the clustering of
 * the last run in place.  best / best_ms (device ms per clustering) may be
 * null.  Runs about 20 clusterings times `repeats`. */
int acvd_set_tuning(acvd_ctx* ctx, const acvd_tuning* t);
int acvd_get_tuning(acvd_ctx* ctx, acvd_tuning* t);
int acvd_autotune(acvd_ctx* ctx, int32_t repeats, acvd_tuning* best, double* best_ms);

/* ---- ForceManifold (Remesh with SetForceManifold(1)) ------------------------
 * acvd_force_manifold: the restart loop of vtkDiscreteRemeshing::Remesh
 * (vtkDiscreteRemeshing.h:937-946) on a context holding one clustered mesh.
 * While the output triangulation, with the wire edges ForceManifold adds
 * between clusters adjacent in the input but joined by no output edge
 * (:1114-1131), has a non-manifold vertex whose cluster holds no non-manifold
 * input vertex (DetectNonManifoldOutputVertices, :165-383): a new cluster takes
 * an item of each such cluster, every cluster not next to one is frozen, and
 * MinimizeEnergy runs again from the current labels (ConnexityConstraint 0).
 * The number of clusters grows by the clusters added (acvd_get_params shows
 * it).  The reference loops without bound; here at most max_restarts restarts
 * run.  *restarts receives the restarts run, *remaining_issues the issues the
 * last detection found (0 when the output passed).  The output mesh of the
 * final clustering is built (acvd_get_output). */
int acvd_force_manifold(acvd_ctx* ctx, int32_t mesh_id, int32_t max_restarts, int32_t* restarts,
                        int64_t* remaining_issues);
/* FixMeshBoundaries (vtkDiscreteRemeshing.h:425-579; Remesh runs it last when
 * SetBoundaryFixing is on, :952) on the output built by acvd_build_output:
 * along the input's boundary, every edge whose ends lie in two clusters adds
 * an output vertex at its midpoint and the triangle (that vertex, the two
 * clusters); each run of boundary edges inside one cluster adds (cluster,
 * the first two such vertices met); a vertex between two cut edges adds
 * (cluster, their two vertices) when the reference's GetNumberOfBoundaries
 * test passes (Common/vtkSurfaceBase.cxx:939-949, which reads the mesh's first
 * edges rather than the vertex's).  The output grows in place
 * (acvd_get_output_sizes / acvd_get_output); *changed = 1 when it did. */
int acvd_fix_mesh_boundaries(acvd_ctx* ctx, int32_t mesh_id, int32_t* changed);
/* InitSamples type 0 on the host (vtkUniformClustering.h:1126-1159): labels
 * (0..K-1, K = null) with one item per cluster drawn with vtkMath::Random
 * seeded 5000 (VTK's minimal standard generator). */
int acvd_reference_seed_sampling(int64_t nv, int32_t n_clusters, int32_t* labels);
/* Host-only form of the same rule on given labels (0..K-1, others: none) of
 * an input mesh: *nv_added vertices (output ids K, K + 1, ...) and *nf_added
 * triangles that FixMeshBoundaries appends; call with null buffers for the
 * counts. */
int acvd_mesh_boundary_fix(int64_t nv, const double* xyz, int64_t nf, const int32_t* tri, const int32_t* labels,
                           int32_t n_clusters, int64_t* nv_added, int64_t* nf_added, double* added_xyz,
                           int32_t* added_tri);
/* IsClusterFreezed per cluster (vtkUniformClustering.h:329, 910-919; no public
 * setter in the reference: DetectNonManifoldOutputVertices sets it).  While
 * set (n = number_of_clusters flags, nonzero = frozen; n = 0 clears), the
 * following runs move items only between two clusters that are not frozen
 * and count only the items of the others for the early stop (:733-736). */
int acvd_set_frozen_clusters(acvd_ctx* ctx, int32_t mesh_id, int64_t n, const int32_t* flags);
/* Host-only helpers of the loop (no device needed):
 * IsVertexManifold (Common/vtkSurfaceBase.cxx:259-317) of every vertex:
 * manifold[v] = 1 when v has at least two edges, each in exactly two faces,
 * and one fan of faces through all of them (a boundary vertex is not). */
int acvd_vertex_manifold(int64_t nv, int64_t nf, const int32_t* tri, uint8_t* manifold);
/* One detection step (DetectNonManifoldOutputVertices) on given labels
 * (0..K-1, other values: none) and output triangles: updates labels and
 * *n_clusters with the added clusters, writes the frozen flags of the
 * resulting clusters (capacity frozen_capacity; may be null) and *issues. */
int acvd_force_manifold_step(int64_t nv, int64_t nf, const int32_t* tri, int64_t n_out, const int32_t* out_tri,
                             int32_t* n_clusters, int32_t* labels, int32_t* frozen, int64_t frozen_capacity,
                             int64_t* issues);

/* ---- fixed-assignment energy and its adjoint -------------------------------
 * No reference counterpart: the reference minimises the clustering energy but
 * never differentiates it.  For frozen labels, per mesh:
 *   isotropic: E = sum_i w_i |x_i - c|^2 with c = S/W of the item's cluster
 *              (the unshifted energy of vtkIsotropicMetricForClustering.h:114-121);
 *   QEM:       E = sum_i w_i |x_i - q|^2 with q the rank-truncated quadric point
 *              of the cluster (vtkQEMetricForClustering.h:194-208,
 *              Common/vtkQuadricTools.cxx:83-177), the cluster quadric being the
 *              sum of the plane quadrics of the faces around its vertices,
 *              recomputed from x so that E is differentiable in x.
 * All buffers of forward/backward are DEVICE pointers: xyz float[3N], w
 * float[N], labels int32[N] (mesh-local ids; K or out of range = no cluster),
 * energy / grad_energy double[M], dxyz float[3N], dw float[N]; meshes are
 * packed in the order they were added.  stream is a hipStream_t (NULL: the
 * null stream); calls are asynchronous on it.  backward uses the per-cluster
 * state of the last forward, which must have run on the same inputs. */
typedef struct acvd_energy acvd_energy;
int acvd_energy_create(acvd_energy** out, int device);
int acvd_energy_destroy(acvd_energy* e);
/* tri: 3*nf mesh-local vertex ids (host memory, copied). */
int acvd_energy_add_mesh(acvd_energy* e, int64_t nv, int64_t nf, const int32_t* tri, int32_t n_clusters,
                         int32_t metric, int32_t quadrics_level, int32_t* mesh_id);
int acvd_energy_num_items(const acvd_energy* e, int64_t* n_items, int32_t* n_meshes);
int acvd_energy_forward(acvd_energy* e, const float* xyz, const float* w, const int32_t* labels, double* energy,
                        void* stream);
int acvd_energy_backward(acvd_energy* e, const float* xyz, const float* w, const int32_t* labels,
                         const double* grad_energy, float* dxyz, float* dw, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): acvd_energy_forward of ONE isotropic (CVT) mesh without a
 * context and without any allocation.  xyz float[3n], w float[n], labels
 * int32[n] (0..n_clusters-1; others = no cluster) and energy double[1] are
 * device pointers on `device`; the per-cluster sums live in a static device
 * array of the library (one per device, held until the process ends), so
 * n_clusters is at most 65536 (ACVD_ERR_LIMIT beyond).  Asynchronous on
 * `stream` (NULL: the null stream).  Calls that overlap on one device race on
 * that array: serialise them (one stream, or events). */
int acvd_energy_forward_cvt_noalloc(int device, int64_t n, int32_t n_clusters, const float* xyz, const float* w,
                                    const int32_t* labels, double* energy, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ACVD_C_H */
