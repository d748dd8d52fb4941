"""Drop-in mirror of the reference's remeshing classes on the MI355X path.

Reference classes (DiscreteRemeshing/vtkIsotropicDiscreteRemeshing.h:42-102):
  vtkIsotropicDiscreteRemeshing  = vtkVerticesProcessing< vtkDiscreteRemeshing< vtkIsotropicMetricForClustering > >
  vtkQIsotropicDiscreteRemeshing = vtkVerticesProcessing< vtkDiscreteRemeshing< vtkQEMetricForClustering > >
The method names, argument meaning and error behaviour follow the reference
(vtkUniformClustering.h:62-134, vtkDiscreteRemeshing.h:61-83,
vtkSurfaceClustering.h:58); the work runs in libacvd_amd.so through the
C-ABI of include/acvd_c.h.  `Context` exposes the batched form (many meshes
clustered by one call).
"""
import ctypes

import numpy as np

from . import _native as N
from ._native import AcvdError, check

_p = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731


class Mesh:
    """Triangle mesh (replaces vtkSurface for the clustering path)."""

    def __init__(self, points, triangles):
        pts = np.asarray(points)
        if pts.dtype not in (np.float32, np.float64):
            pts = pts.astype(np.float64)
        self.points = np.ascontiguousarray(pts).reshape(-1, 3)
        self.triangles = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)

    def GetNumberOfPoints(self):
        return len(self.points)

    def GetNumberOfCells(self):
        return len(self.triangles)

    def GetPoint(self, i):
        return tuple(float(v) for v in self.points[i])


class Context:
    """One library context (one HIP stream) holding one or more meshes."""

    def __init__(self, device=0):
        L = N.lib()
        h = ctypes.c_void_p()
        check(L.acvd_create(ctypes.byref(h), device))
        self._h = h
        self._nv = []
        # held by the instance: module globals may already be gone when a
        # context is collected at interpreter exit
        self._destroy = L.acvd_destroy

    def close(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @staticmethod
    def _mesh_args(points, triangles):
        pts = np.ascontiguousarray(points)
        if pts.dtype not in (np.float32, np.float64):
            pts = pts.astype(np.float64)
        pts = pts.reshape(-1, 3)
        tri = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
        return pts, tri

    def set_input(self, points, triangles):
        pts, tri = self._mesh_args(points, triangles)
        f = N.lib().acvd_set_input_f32 if pts.dtype == np.float32 else N.lib().acvd_set_input_f64
        check(f(self._h, len(pts), _p(pts), len(tri), _p(tri)))
        self._nv = [len(pts)]
        return 0

    def add_input(self, points, triangles):
        pts, tri = self._mesh_args(points, triangles)
        f = N.lib().acvd_add_input_f32 if pts.dtype == np.float32 else N.lib().acvd_add_input_f64
        mid = ctypes.c_int32()
        check(f(self._h, len(pts), _p(pts), len(tri), _p(tri), ctypes.byref(mid)))
        self._nv.append(len(pts))
        return mid.value

    @property
    def num_meshes(self):
        return len(self._nv)

    def _n(self, mesh_id):
        if not 0 <= mesh_id < len(self._nv):
            raise AcvdError(f"mesh id {mesh_id} out of range (context holds {len(self._nv)} meshes)")
        return self._nv[mesh_id]

    def set_params(self, mesh_id=-1, **kw):
        p = N.Params()
        if mesh_id >= 0:
            check(N.lib().acvd_get_params(self._h, mesh_id, ctypes.byref(p)))
        else:
            N.lib().acvd_params_init(ctypes.byref(p))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(f"unknown parameter {k}")
            setattr(p, k, v)
        check(N.lib().acvd_set_params(self._h, mesh_id, ctypes.byref(p)))

    def get_params(self, mesh_id=0):
        p = N.Params()
        check(N.lib().acvd_get_params(self._h, mesh_id, ctypes.byref(p)))
        return {k: getattr(p, k) for k, _ in p._fields_}

    def set_indicator(self, mesh_id, indicator):
        ind = None if indicator is None else np.ascontiguousarray(indicator, np.float64)
        if ind is not None and len(ind) != self._n(mesh_id):
            raise ValueError("indicator length must equal the number of vertices")
        check(N.lib().acvd_set_indicator(self._h, mesh_id, _p(ind)))

    def set_initial_labels(self, mesh_id, labels):
        lab = None if labels is None else np.ascontiguousarray(labels, np.int32)
        if lab is not None and len(lab) != self._n(mesh_id):
            raise ValueError("initial clustering length must equal the number of vertices")
        check(N.lib().acvd_set_initial_labels(self._h, mesh_id, _p(lab)))

    def set_fixed_clusters(self, mesh_id, items):
        """SetFixedClusters + AnchorItem (cluster k fixed at item items[k])."""
        ids = np.zeros(0, np.int32) if items is None else np.ascontiguousarray(items, np.int32).reshape(-1)
        check(N.lib().acvd_set_fixed_clusters(self._h, mesh_id, len(ids), _p(ids) if len(ids) else None))

    def indicator(self, mesh_id=0):
        """The indicator the last build weighted the mesh with: the one set,
        else the computed curvature indicator (params curvature_indicator)."""
        out = np.empty(self._n(mesh_id), np.float64)
        check(N.lib().acvd_get_indicator(self._h, mesh_id, _p(out)))
        return out

    def subdivide_input(self, mesh_id, min_vertices):
        """Midpoint subdivision of the input on the device until it has at
        least min_vertices vertices (CheckSubsamplingRatio,
        vtkDiscreteRemeshing.h:840-875); a set indicator follows, each
        midpoint taking the mean of its ends.  Returns the number of levels."""
        self._n(mesh_id)
        lv = ctypes.c_int32()
        check(N.lib().acvd_subdivide_input(self._h, mesh_id, int(min_vertices), ctypes.byref(lv)))
        nv = ctypes.c_int64()
        check(N.lib().acvd_get_input_sizes(self._h, mesh_id, ctypes.byref(nv), None))
        self._nv[mesh_id] = nv.value
        return lv.value

    def input(self, mesh_id=0):
        """The input mesh as the context holds it (points, triangles)."""
        nv, nf = ctypes.c_int64(), ctypes.c_int64()
        check(N.lib().acvd_get_input_sizes(self._h, mesh_id, ctypes.byref(nv), ctypes.byref(nf)))
        xyz = np.empty((nv.value, 3), np.float64)
        tri = np.empty((nf.value, 3), np.int32)
        check(N.lib().acvd_get_input(self._h, mesh_id, _p(xyz), _p(tri)))
        return xyz, tri

    def build(self):
        check(N.lib().acvd_build(self._h))

    def initial_sampling(self):
        """Device initial clustering of every mesh without user labels."""
        check(N.lib().acvd_initial_sampling(self._h))

    def set_timing(self, enable, counters=True):
        """Per-kernel HIP-event timing of the next runs; counters: also count
        the accept-round work (the counting kernels run slightly slower)."""
        check(N.lib().acvd_set_timing(self._h, (1 if counters else 2) if enable else 0))

    def process(self):
        check(N.lib().acvd_process_clustering(self._h))

    def labels(self, mesh_id=0):
        out = np.empty(self._n(mesh_id), np.int32)
        check(N.lib().acvd_get_labels(self._h, mesh_id, _p(out)))
        return out

    def cluster_sizes(self, mesh_id=0):
        K = self.get_params(mesh_id)["number_of_clusters"]
        out = np.empty(K, np.int32)
        check(N.lib().acvd_get_cluster_sizes(self._h, mesh_id, _p(out)))
        return out

    def cluster_points(self, mesh_id=0):
        K = self.get_params(mesh_id)["number_of_clusters"]
        out = np.empty((K, 3), np.float64)
        check(N.lib().acvd_get_cluster_points(self._h, mesh_id, _p(out)))
        return out

    def energy(self, mesh_id=0):
        e = np.zeros(1, np.float64)
        check(N.lib().acvd_get_energy(self._h, mesh_id, _p(e)))
        return float(e[0])

    def fix_mesh_boundaries(self, mesh_id=0):
        """FixMeshBoundaries on the built output (acvd_fix_mesh_boundaries):
        True when the output changed."""
        changed = ctypes.c_int32()
        check(N.lib().acvd_fix_mesh_boundaries(self._h, mesh_id, ctypes.byref(changed)))
        return bool(changed.value)

    def set_tuning(self, **kw):
        """Launch shape (acvd_set_tuning): sweep_segments, tile_on_divisor,
        chunk_passes, short_chunk_passes; unnamed knobs keep their value."""
        t = N.Tuning()
        check(N.lib().acvd_get_tuning(self._h, ctypes.byref(t)))
        for k, v in kw.items():
            if not hasattr(t, k):
                raise AttributeError(f"unknown tuning knob {k}")
            setattr(t, k, int(v))
        check(N.lib().acvd_set_tuning(self._h, ctypes.byref(t)))

    def get_tuning(self):
        t = N.Tuning()
        check(N.lib().acvd_get_tuning(self._h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def autotune(self, repeats=2):
        """acvd_autotune: (best launch shape, device ms per clustering); the
        shape stays set."""
        t = N.Tuning()
        ms = ctypes.c_double()
        check(N.lib().acvd_autotune(self._h, int(repeats), ctypes.byref(t), ctypes.byref(ms)))
        return {k: getattr(t, k) for k, _ in t._fields_}, ms.value

    def sizes(self):
        """(items, clusters, QEM meshes) over the context's meshes."""
        n = sum(self._nv)
        k = q = 0
        for m in range(len(self._nv)):
            p = self.get_params(m)
            k += p["number_of_clusters"]
            q += p["metric"] == N.METRIC_QEM
        return n, k, q

    def set_frozen_clusters(self, mesh_id, flags):
        """IsClusterFreezed per cluster for the following runs (None clears)."""
        f = np.zeros(0, np.int32) if flags is None else np.ascontiguousarray(flags, np.int32).reshape(-1)
        check(N.lib().acvd_set_frozen_clusters(self._h, mesh_id, len(f), _p(f) if len(f) else None))

    def force_manifold(self, mesh_id=0, max_restarts=100):
        """Remesh's ForceManifold loop (acvd_force_manifold): returns
        (restarts run, issues left)."""
        r = ctypes.c_int32()
        left = ctypes.c_int64()
        check(N.lib().acvd_force_manifold(self._h, mesh_id, int(max_restarts), ctypes.byref(r), ctypes.byref(left)))
        return r.value, left.value

    def energy_log(self, mesh_id=0):
        """(seconds, energy, final_energy) of the last run with
        write_energy_log set: one entry per loop (acvd_get_energy_log)."""
        n = ctypes.c_int64()
        check(N.lib().acvd_get_energy_log(self._h, mesh_id, 0, ctypes.byref(n), None, None, None))
        sec = np.empty(n.value, np.float64)
        en = np.empty(n.value, np.float64)
        fin = np.zeros(1, np.float64)
        check(N.lib().acvd_get_energy_log(self._h, mesh_id, n.value, ctypes.byref(n), _p(sec), _p(en), _p(fin)))
        return sec, en, float(fin[0])

    def stats(self, mesh_id=0):
        st = N.Stats()
        check(N.lib().acvd_get_stats(self._h, mesh_id, ctypes.byref(st)))
        return st.as_dict()

    def adjacency(self, mesh_id=0):
        n = ctypes.c_int64()
        check(N.lib().acvd_get_adjacency_size(self._h, mesh_id, ctypes.byref(n)))
        rp = np.empty(self._n(mesh_id) + 1, np.int32)
        col = np.empty(n.value, np.int32)
        check(N.lib().acvd_get_adjacency(self._h, mesh_id, _p(rp), _p(col)))
        return rp, col

    def items(self, mesh_id=0):
        nv = self._n(mesh_id)
        xw = np.empty((nv, 4), np.float32)
        q = np.empty((nv, 10), np.float32)
        check(N.lib().acvd_get_items(self._h, mesh_id, _p(xw), _p(q)))
        return xw, q

    def output(self, mesh_id=0, quadrics_level=0, build=True):
        """Output mesh (cluster points, triangles) of the last clustering
        (build=False: the output as last built and fixed, without rebuilding)."""
        if build:
            check(N.lib().acvd_build_output(self._h, mesh_id, quadrics_level))
        nv, nf = ctypes.c_int64(), ctypes.c_int64()
        check(N.lib().acvd_get_output_sizes(self._h, mesh_id, ctypes.byref(nv), ctypes.byref(nf)))
        xyz = np.empty((nv.value, 3), np.float64)
        tri = np.empty((nf.value, 3), np.int32)
        check(N.lib().acvd_get_output(self._h, mesh_id, _p(xyz), _p(tri)))
        return xyz, tri

    def scales(self, mesh_id=0):
        s = np.empty(3, np.float64)
        check(N.lib().acvd_get_scales(self._h, mesh_id, _p(s)))
        return s


def reference_initial_sampling(xyz, tri, n_clusters, metric=N.METRIC_ISOTROPIC, indicator=None, gradation=0.0):
    """The reference's default initial sampling (region growing from shuffled
    items, vtkUniformClustering.h:1177-1316) on the host: labels 0..K-1, K for
    unassigned items."""
    x = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(tri, np.int32).reshape(-1, 3)
    ind = None if indicator is None else np.ascontiguousarray(indicator, np.float64)
    out = np.empty(len(x), np.int32)
    check(N.lib().acvd_reference_initial_sampling(len(x), _p(x), len(t), _p(t), int(metric), _p(ind),
                                                  float(gradation), int(n_clusters), _p(out)))
    return out


def vertex_manifold(n_vertices, triangles):
    """IsVertexManifold of every vertex (acvd_vertex_manifold): bool array."""
    tri = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
    out = np.zeros(int(n_vertices), np.uint8)
    check(N.lib().acvd_vertex_manifold(int(n_vertices), len(tri), _p(tri), _p(out)))
    return out.astype(bool)


def force_manifold_step(n_vertices, triangles, labels, n_clusters, out_triangles):
    """One DetectNonManifoldOutputVertices step (acvd_force_manifold_step):
    (issues, labels, n_clusters, frozen flags)."""
    tri = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
    otri = np.ascontiguousarray(out_triangles, np.int32).reshape(-1, 3)
    lab = np.array(labels, np.int32)
    k = ctypes.c_int32(int(n_clusters))
    frozen = np.zeros(2 * int(n_clusters) + 1, np.int32)
    n = ctypes.c_int64()
    check(N.lib().acvd_force_manifold_step(int(n_vertices), len(tri), _p(tri), len(otri), _p(otri), ctypes.byref(k),
                                           _p(lab), _p(frozen), len(frozen), ctypes.byref(n)))
    return n.value, lab, k.value, frozen[:k.value]


def reference_seed_sampling(n_items, n_clusters):
    """InitSamples type 0 (acvd_reference_seed_sampling): labels 0..K-1,
    K = null, one random item per cluster."""
    out = np.empty(int(n_items), np.int32)
    check(N.lib().acvd_reference_seed_sampling(int(n_items), int(n_clusters), _p(out)))
    return out


def mesh_boundary_fix(points, triangles, labels, n_clusters):
    """FixMeshBoundaries on given labels (acvd_mesh_boundary_fix): the output
    vertices it adds (ids n_clusters, ...) and the triangles it appends."""
    x = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    tri = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
    lab = np.ascontiguousarray(labels, np.int32)
    a, b = ctypes.c_int64(), ctypes.c_int64()
    check(N.lib().acvd_mesh_boundary_fix(len(x), _p(x), len(tri), _p(tri), _p(lab), int(n_clusters), ctypes.byref(a),
                                         ctypes.byref(b), None, None))
    ax = np.empty((a.value, 3), np.float64)
    at = np.empty((b.value, 3), np.int32)
    check(N.lib().acvd_mesh_boundary_fix(len(x), _p(x), len(tri), _p(tri), _p(lab), int(n_clusters), ctypes.byref(a),
                                         ctypes.byref(b), _p(ax), _p(at)))
    return ax, at


def subdivide(points, triangles):
    """One level of midpoint subdivision (vtkSurface::Subdivide,
    Common/vtkSurface.cxx:605-677) on the host: (points, triangles, parents)
    with the input vertices first, then one midpoint per edge in the
    reference's edge order; parents[k] = the two ends of new vertex nv + k."""
    x = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
    onv, onf = ctypes.c_int64(), ctypes.c_int64()
    check(N.lib().acvd_subdivide(len(x), _p(x), len(t), _p(t), ctypes.byref(onv), ctypes.byref(onf), None, None,
                                 None))
    ox = np.empty((onv.value, 3), np.float64)
    ot = np.empty((onf.value, 3), np.int32)
    par = np.empty((onv.value - len(x), 2), np.int32)
    check(N.lib().acvd_subdivide(len(x), _p(x), len(t), _p(t), ctypes.byref(onv), ctypes.byref(onf), _p(ox), _p(ot),
                                 _p(par)))
    return ox, ot, par


class _MetricProxy:
    """What GetMetric() returns: the metric-dependent options."""

    def __init__(self, owner):
        self._o = owner

    def SetGradation(self, g):
        self._o._params["gradation"] = float(g)

    def GetGradation(self):
        return self._o._params["gradation"]

    def SetQuadricsOptimizationLevel(self, level):
        if self._o.METRIC != N.METRIC_QEM:
            raise AttributeError("SetQuadricsOptimizationLevel exists on the QEM metric only")
        self._o._params["quadrics_optimization_level"] = int(level)

    def SetCurvatureInfo(self, indicator):
        """Per-vertex indicator, used as weight = area * indicator^gradation."""
        self._o._indicator = None if indicator is None else np.ascontiguousarray(indicator, np.float64)


class _DiscreteRemeshing:
    METRIC = N.METRIC_ISOTROPIC

    def __init__(self):
        self._input = None
        self._params = {"number_of_clusters": 0, "metric": self.METRIC, "quadrics_optimization_level": 3,
                        "unconstrained_initialization": 0, "max_number_of_loops": 5000000,
                        "max_number_of_convergences": 1000000000, "console_output": 0, "seed": 0,
                        "gradation": 0.0}
        self._indicator = None
        self._initial = None
        self._fixed = None
        self._anchors = {}
        self._clustering = None
        self._ctx = None
        self._subsampling_threshold = 10
        self._output = None
        self._metric = _MetricProxy(self)
        self._output_directory = None
        self._force_manifold = False
        self._boundary_fixing = 0

    @classmethod
    def New(cls):
        return cls()

    def Delete(self):
        if self._ctx is not None:
            self._ctx.close()
            self._ctx = None

    # ---- configuration (vtkUniformClustering.h:62-134) ----
    def SetInput(self, mesh):
        if not isinstance(mesh, Mesh):
            mesh = Mesh(*mesh)
        self._input = mesh

    def GetInput(self):
        return self._input

    def SetNumberOfClusters(self, n):
        self._params["number_of_clusters"] = int(n)

    def GetNumberOfClusters(self):
        return self._params["number_of_clusters"]

    def SetConsoleOutput(self, v):
        self._params["console_output"] = int(v)

    def SetUnconstrainedInitialization(self, v):
        self._params["unconstrained_initialization"] = int(v)

    def SetMaxNumberOfLoops(self, v):
        self._params["max_number_of_loops"] = int(v)

    def SetMaxNumberOfConvergences(self, v):
        self._params["max_number_of_convergences"] = int(v)

    def SetInitialSamplingType(self, t):
        """vtkUniformClustering::SetInitialSamplingType (vtkUniformClustering.h:108,
        1123-1175): 1 the region growing (default), 0 one random item per
        cluster (vtkMath::Random seeded 5000), -1 item k for cluster k, 2 the
        labels of SetInitialClustering."""
        t = int(t)
        if t not in (-1, 0, 1, 2):
            raise ValueError("initial sampling type must be -1, 0, 1 or 2")
        self._params["initial_sampling"] = {1: 0, 0: 2, -1: 3, 2: 0}[t]

    def SetInitialClustering(self, labels):
        self._initial = None if labels is None else np.ascontiguousarray(labels, np.int32)

    def SetFixedClusters(self, ids):
        """vtkUniformClustering::SetFixedClusters (vtkUniformClustering.h:134):
        cluster k starts at item ids[k] and keeps it (NumberOfClusters counts
        these clusters).  With the QEM metric the item also becomes the
        cluster's AnchorItem, as ACVDQ sets it (Examples/ACVDQ.cxx:327-334)."""
        self._fixed = None if ids is None else np.ascontiguousarray(ids, np.int32).reshape(-1)
        self._anchors = {}

    def GetCluster(self, i):
        """The cluster record, for its AnchorItem (the fixed item, or -1)."""
        return _ClusterProxy(self, int(i))

    def SetWriteToGlobalEnergyLog(self, v):
        """vtkUniformClustering::SetWriteToGlobalEnergyLog (vtkUniformClustering.h:123):
        ProcessClustering then writes energy.txt, prefixed by the output
        directory as given (no separator is added, :670-675): one line
        "loop seconds energy" per loop and "Final Energy :<energy>" after the
        run (:695-699, 1383-1395)."""
        self._params["write_energy_log"] = int(v)

    def SetOutputDirectory(self, d):
        """vtkUniformClustering::SetOutputDirectory (vtkUniformClustering.h:126)."""
        self._output_directory = None if d is None else str(d)

    def SetBoundaryFixing(self, v):
        """vtkDiscreteRemeshing::SetBoundaryFixing (vtkDiscreteRemeshing.h:77):
        Remesh closes the output along the input's boundaries
        (FixMeshBoundaries, :425-579)."""
        self._boundary_fixing = int(v)

    def SetForceManifold(self, v):
        """vtkDiscreteRemeshing::SetForceManifold (vtkDiscreteRemeshing.h:79):
        Remesh restarts the minimisation with added clusters while the output
        has non-manifold vertices (:937-946), at most 100 times."""
        self._force_manifold = bool(v)

    def SetSubsamplingThreshold(self, t):
        self._subsampling_threshold = int(t)

    def SetDisplay(self, d):
        pass  # no rendering on this path

    def SetNumberOfThreads(self, n):
        """vtkThreadedClustering::SetNumberOfThreads (vtkThreadedClustering.h:58-74,
        ACVDP / ACVDQP): accepted for drop-in use; the device runs the sweep."""

    def SetPoolingRatio(self, r):
        """vtkThreadedClustering::SetPoolingRatio: accepted; see SetNumberOfThreads."""

    def SetSeed(self, s):
        self._params["seed"] = int(s) & 0xFFFFFFFF

    def GetMetric(self):
        return self._metric

    def GetNumberOfItems(self):
        return self._input.GetNumberOfPoints()

    def GetClusteringType(self):
        return 1  # items are vertices

    # ---- run ----
    def ProcessClustering(self, curvature_indicator=0):
        """vtkUniformClustering::ProcessClustering; returns the labels (0..K, K = null).
        Returns None when NumberOfClusters is zero, like the reference (:659-664).
        curvature_indicator = 1 (set by Remesh) weights the items with the
        reference's curvature indicator when the gradation is positive and no
        indicator was set."""
        if self._params["number_of_clusters"] <= 0:
            return None
        if self._input is None:
            raise AcvdError("SetInput must be called first")
        if self._ctx is None:
            self._ctx = Context()
        ctx = self._ctx
        ctx.set_input(self._input.points, self._input.triangles)
        ctx.set_params(0, curvature_indicator=int(curvature_indicator), **self._params)
        ctx.set_indicator(0, self._indicator)
        ctx.set_initial_labels(0, self._initial)
        ctx.set_fixed_clusters(0, self._fixed)
        ctx.process()
        self._clustering = ctx.labels(0)
        if self._params.get("write_energy_log"):
            self._write_energy_log()
        return self._clustering

    def _write_energy_log(self):
        """The file of WriteToGlobalEnergyLog: loop number, seconds since the
        start of the minimisation (6 significant digits), energy (15); then the
        final energy.  Like the reference's ofstream, a file that cannot be
        opened is skipped silently."""
        sec, en, fin = self._ctx.energy_log(0)
        name = (self._output_directory or "") + "energy.txt"
        lines = ["%d %.6g %.15g\n" % (i, t, e) for i, (t, e) in enumerate(zip(sec, en))]
        lines.append("Final Energy :%.15g\n" % fin)
        try:
            with open(name, "w") as f:
                f.writelines(lines)
        except OSError:
            pass

    def GetClustering(self):
        return self._clustering

    def Remesh(self):
        """vtkDiscreteRemeshing::Remesh (vtkDiscreteRemeshing.h:877-954):
        CheckSubsamplingRatio (:840-875: while the input has fewer than
        SubsamplingThreshold * K vertices it is replaced by its midpoint
        subdivision, and GetInput() returns the subdivided mesh),
        SamplingPreProcessing (with a positive gradation and no indicator set,
        the items are weighted by the reference's curvature indicator, computed
        on the device on the original input and interpolated to the midpoints
        as the mean of their two parents, :583-653, 700-745),
        ProcessClustering, the ForceManifold restarts when set (:937-946),
        the output triangulation, then FixMeshBoundaries when boundary
        fixing is set (:952)."""
        K = self._params["number_of_clusters"]
        if K <= 0:
            return
        original = self._input
        curvature = 1
        min_vertices = int(np.ceil(self._subsampling_threshold * K))
        if original.GetNumberOfPoints() < min_vertices:
            if self._initial is not None:
                raise AcvdError("SetInitialClustering labels the input's vertices; the input needs subdivision")
            # subdivision on the device; the indicator (set, or the curvature
            # indicator of the original input) follows to the midpoints
            c = Context()
            c.set_input(original.points, original.triangles)
            ind = self._indicator
            if ind is None and self._params["gradation"] > 0:
                c.set_params(0, number_of_clusters=1, metric=self.METRIC, gradation=self._params["gradation"],
                             curvature_indicator=1)
                c.build()
                ind = c.indicator(0)
            if ind is not None:
                c.set_indicator(0, ind)
            levels = c.subdivide_input(0, min_vertices)
            if self._params["console_output"]:
                for _ in range(levels):
                    print("Subdividing mesh")
            x, t = c.input(0)
            if ind is not None:
                ind = c.indicator(0)
            c.close()
            self._original_input = original
            self._input = Mesh(x, t)
            self._indicator = ind
            curvature = 0
        if self.ProcessClustering(curvature_indicator=curvature) is None:
            return
        if self._force_manifold:
            self._restarts = self._ctx.force_manifold(0)
            self._params["number_of_clusters"] = self._ctx.get_params(0)["number_of_clusters"]
            self._clustering = self._ctx.labels(0)
        xyz, tri = self._ctx.output(0, 0)
        if self._boundary_fixing and self._ctx.fix_mesh_boundaries(0):
            xyz, tri = self._ctx.output(0, build=False)
        self._output = Mesh(xyz, tri)

    def GetOutput(self):
        return self._output

    def QuadricPostProcessing(self, level=1):
        """Quadric placement of the output vertices (Examples/ACVD.cxx:217-291);
        with boundary fixing on, the boundary vertices and triangles are kept
        (the placement moves only the cluster points)."""
        xyz, tri = self._ctx.output(0, level)
        if self._boundary_fixing and self._ctx.fix_mesh_boundaries(0):
            xyz, tri = self._ctx.output(0, build=False)
        self._output = Mesh(xyz, tri)
        return self._output

    def GetClusterSize(self, c):
        return int(self._ctx.cluster_sizes(0)[c])

    def GetClusterPoints(self):
        return self._ctx.cluster_points(0)

    def GetStats(self):
        return self._ctx.stats(0)

    def GetContext(self):
        return self._ctx


class _ClusterProxy:
    """GetCluster(i): its AnchorItem is the fixed item of cluster i (the
    reference sets it to the same item, Examples/ACVDQ.cxx:333-334); another
    value than that item is refused."""

    def __init__(self, owner, i):
        self._owner = owner
        self._i = i

    @property
    def AnchorItem(self):
        f = self._owner._fixed
        return int(f[self._i]) if f is not None and self._i < len(f) else -1

    @AnchorItem.setter
    def AnchorItem(self, item):
        if int(item) != self.AnchorItem:
            raise AcvdError("AnchorItem must be the cluster's fixed item (SetFixedClusters)")


class IsotropicDiscreteRemeshing(_DiscreteRemeshing):
    """vtkIsotropicDiscreteRemeshing (ACVD)."""
    METRIC = N.METRIC_ISOTROPIC


class QIsotropicDiscreteRemeshing(_DiscreteRemeshing):
    """vtkQIsotropicDiscreteRemeshing (ACVDQ)."""
    METRIC = N.METRIC_QEM
