"""TEST INFRASTRUCTURE — ctypes wrapper of the CPU oracles (oracle/_build/liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker / CPU baseline.

Oracles (all restate the reference in C++, see the headers for file:line):
  serial_run    A: vtkUniformClustering + vtkVerticesProcessing + metric (FP64)
  threaded_run  B: vtkThreadedClustering (ACVDP / ACVDQP), the timed CPU baseline
  replay_run    C: the GPU pass schedule replayed on the host (bit-exact target)
  energy        FP64 unshifted energy of a labelling
  RefMesh.curvature_indicator  the reference's per-vertex curvature indicator
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

ISO, QEM = 0, 1


class OrcStats(ctypes.Structure):
    _fields_ = [("loops", ctypes.c_int64), ("convergences", ctypes.c_int64),
                ("total_moves", ctypes.c_int64), ("evaluated", ctypes.c_int64),
                ("seconds", ctypes.c_double), ("shifted_energy", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        L.orc_mesh_create.restype = P
        L.orc_mesh_create.argtypes = [ctypes.c_int64, P, ctypes.c_int64, P]
        L.orc_mesh_destroy.argtypes = [P]
        L.orc_mesh_num_edges.restype = ctypes.c_int64
        L.orc_mesh_num_edges.argtypes = [P]
        L.orc_mesh_edges.argtypes = [P, P, P]
        L.orc_mesh_vertex_areas.argtypes = [P, P]
        L.orc_curvature_indicator.argtypes = [P, P]
        L.orc_curvature_face_list.restype = ctypes.c_int64
        L.orc_curvature_face_list.argtypes = [P, ctypes.c_int64, P, ctypes.c_int64]
        L.orc_build_items.argtypes = [P, ctypes.c_int, P, ctypes.c_double, P, P]
        L.orc_serial_run.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, P, ctypes.c_int, ctypes.c_int,
                                     P, ctypes.c_int, P, P]
        L.orc_set_progress.argtypes = [ctypes.c_int]
        L.orc_set_fixed.argtypes = [P, ctypes.c_int64]
        L.orc_serial_trace.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, P, ctypes.c_int, ctypes.c_int,
                                       P, ctypes.c_int, P, P, ctypes.c_int64, P, P, ctypes.c_int64, P, P]
        L.orc_threaded_run.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, P, ctypes.c_int,
                                       ctypes.c_int, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P]
        L.orc_initial_sampling.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, P, P]
        L.orc_fill_holes.argtypes = [P, ctypes.c_int, P]
        L.orc_clean.argtypes = [P, ctypes.c_int, P]
        L.orc_energy.argtypes = [P, ctypes.c_int, P, ctypes.c_double, ctypes.c_int, ctypes.c_int, P, P]
        L.orc_rep_point.argtypes = [P, P, ctypes.c_int, ctypes.c_double]
        L.orc_sym_eigen.argtypes = [P, P, P]
        L.orc_triangle_quadric.argtypes = [P, P, P, P]
        L.orc_ring_problem.argtypes = [P, P, ctypes.c_int64, ctypes.c_int32]
        L.orc_hdr_rep_point_delta.argtypes = [P, P, ctypes.c_int, P, P, P]
        L.orc_hdr_ring_problem.argtypes = [P, P, P, ctypes.c_int32, ctypes.c_int32]
        L.orc_hdr_ring_problem_mask.argtypes = [P, P, P, ctypes.c_int32, ctypes.c_int32]
        L.orc_hdr_accept_key.restype = ctypes.c_uint64
        L.orc_hdr_accept_key.argtypes = [ctypes.c_uint32, ctypes.c_int32, ctypes.c_int, ctypes.c_uint32]
        L.orc_replay_run.argtypes = [ctypes.c_int64, P, P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P]
        L.orc_replay_run_path.argtypes = [ctypes.c_int] + L.orc_replay_run.argtypes
        L.orc_replay_energy_log.restype = ctypes.c_int64
        L.orc_replay_nonadjacent.restype = ctypes.c_int64
        L.orc_replay_nonadjacent.argtypes = []
        L.orc_replay_energy_log.argtypes = [ctypes.c_int64, P, P]
        L.orc_set_energy_log.argtypes = [ctypes.c_int]
        L.orc_set_frozen.argtypes = [P, ctypes.c_int64]
        L.orc_serial_energy_log.restype = ctypes.c_int64
        L.orc_serial_energy_log.argtypes = [ctypes.c_int64, P, P]
        L.orc_devinputs_create.restype = P
        L.orc_devinputs_create.argtypes = [ctypes.c_int64, P, ctypes.c_int64, P, ctypes.c_int, P, ctypes.c_double]
        L.orc_devinputs_ncol.restype = ctypes.c_int64
        L.orc_devinputs_ncol.argtypes = [P]
        L.orc_devinputs_get.argtypes = [P, P, P, P, P, P]
        L.orc_devinputs_destroy.argtypes = [P]
        L.orc_hdr_terms.argtypes = [P, P, ctypes.c_double, P, ctypes.c_double, P]
        L.orc_hdr_qem_T.restype = ctypes.c_double
        L.orc_hdr_qem_T.argtypes = [P, P, P, ctypes.c_int, P]
        L.orc_hdr_fix_item.argtypes = [P, P, P, P]
        L.orc_mt19937_10000.restype = ctypes.c_uint32
        L.orc_minstd_10000.restype = ctypes.c_int64
        L.orc_seed_sampling.argtypes = [P, ctypes.c_int, P]
        L.orc_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class RefMesh:
    """Reference-ordered mesh (edge table in face order)."""

    def __init__(self, xyz, tri):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
        self.nv = len(self.xyz)
        self.h = lib().orc_mesh_create(self.nv, _p(self.xyz), len(self.tri), _p(self.tri))
        if not self.h:
            raise ValueError(lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_destroy(self.h)
            self.h = None

    def edges(self):
        ne = lib().orc_mesh_num_edges(self.h)
        a = np.empty(ne, np.int32)
        b = np.empty(ne, np.int32)
        lib().orc_mesh_edges(self.h, _p(a), _p(b))
        return a, b

    def curvature_indicator(self):
        """vtkCurvatureMeasure's per-vertex indicator (oracle/curvature.h)"""
        out = np.empty(self.nv, np.float64)
        lib().orc_curvature_indicator(self.h, _p(out))
        return out

    def curvature_face_list(self, v):
        cap = 4096
        out = np.empty(cap, np.int32)
        n = lib().orc_curvature_face_list(self.h, int(v), _p(out), cap)
        return out[:min(n, cap)].copy()

    def vertex_areas(self):
        out = np.empty(self.nv, np.float64)
        lib().orc_mesh_vertex_areas(self.h, _p(out))
        return out

    def items(self, metric, indicator=None, gradation=0.0):
        w = np.empty(self.nv, np.float64)
        q = np.empty((self.nv, 9), np.float64) if metric == QEM else None
        ind = None if indicator is None else np.ascontiguousarray(indicator, np.float64)
        lib().orc_build_items(self.h, metric, _p(ind), gradation, _p(w), _p(q))
        return w, q

    def serial_run(self, metric, K, gradation=0.0, indicator=None, level=3, unconstrained=0, init=None,
                   max_loops=5000000):
        out = np.empty(self.nv, np.int32)
        st = OrcStats()
        ind = None if indicator is None else np.ascontiguousarray(indicator, np.float64)
        ini = None if init is None else np.ascontiguousarray(init, np.int32)
        rc = lib().orc_serial_run(self.h, metric, K, gradation, _p(ind), level, unconstrained, _p(ini), max_loops,
                                  _p(out), ctypes.byref(st))
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        return out, st.as_dict()

    def serial_trace(self, metric, K, gradation=0.0, indicator=None, level=3, unconstrained=0, init=None,
                     max_loops=5000000, edge_cap=100000, loop_cap=100000):
        """serial_run with its trace: (labels, stats, edges, loops) where edges
        is a dict of arrays loop, decision (0 stay, 1 move I1, 2 move I2),
        edge, i1, i2, try1, try2, try3 (one per evaluated cut edge, the first
        edge_cap) and loops a dict of arrays loop, mods, connexity,
        constrained, convergence, disconnected (one per loop)."""
        out = np.empty(self.nv, np.int32)
        st = OrcStats()
        ind = None if indicator is None else np.ascontiguousarray(indicator, np.float64)
        ini = None if init is None else np.ascontiguousarray(init, np.int32)
        ei = np.empty((edge_cap, 5), np.int64)
        ed = np.empty((edge_cap, 3), np.float64)
        li = np.empty((loop_cap, 6), np.int32)
        n = np.zeros(2, np.int64)
        rc = lib().orc_serial_trace(self.h, metric, K, gradation, _p(ind), level, unconstrained, _p(ini), max_loops,
                                    _p(out), ctypes.byref(st), edge_cap, _p(ei), _p(ed), loop_cap, _p(li), _p(n))
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        ne, nl = int(n[0]), int(n[1])
        edges = {k: ei[:ne, j].copy() for j, k in enumerate(("loop", "decision", "edge", "i1", "i2"))}
        edges.update({k: ed[:ne, j].copy() for j, k in enumerate(("try1", "try2", "try3"))})
        loops = {k: li[:nl, j].copy() for j, k in
                 enumerate(("loop", "mods", "connexity", "constrained", "convergence", "disconnected"))}
        return out, st.as_dict(), edges, loops

    def threaded_run(self, metric, K, gradation=0.0, indicator=None, level=3, unconstrained=0, init=None,
                     max_loops=5000000, threads=1, pooling=5):
        out = np.empty(self.nv, np.int32)
        st = OrcStats()
        ind = None if indicator is None else np.ascontiguousarray(indicator, np.float64)
        ini = None if init is None else np.ascontiguousarray(init, np.int32)
        rc = lib().orc_threaded_run(self.h, metric, K, gradation, _p(ind), level, unconstrained, _p(ini), max_loops,
                                    threads, pooling, _p(out), ctypes.byref(st))
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        return out, st.as_dict()

    def initial_sampling(self, metric, K, gradation=0.0, indicator=None):
        out = np.empty(self.nv, np.int32)
        ind = None if indicator is None else np.ascontiguousarray(indicator, np.float64)
        lib().orc_initial_sampling(self.h, metric, K, gradation, _p(ind), _p(out))
        return out

    def seed_sampling(self, K):
        """InitSamples type 0 (vtkMath::Random seeded 5000)."""
        out = np.empty(self.nv, np.int32)
        lib().orc_seed_sampling(self.h, K, _p(out))
        return out

    def fill_holes(self, K, labels):
        lab = np.array(labels, np.int32, copy=True)
        lib().orc_fill_holes(self.h, K, _p(lab))
        return lab

    def clean(self, K, labels):
        lab = np.array(labels, np.int32, copy=True)
        n = lib().orc_clean(self.h, K, _p(lab))
        return lab, n

    def energy(self, metric, K, labels, gradation=0.0, indicator=None, level=3):
        """(unshifted, shifted, sum w|x|^2) in FP64."""
        out = np.empty(3, np.float64)
        ind = None if indicator is None else np.ascontiguousarray(indicator, np.float64)
        lab = np.ascontiguousarray(labels, np.int32)
        lib().orc_energy(self.h, metric, _p(ind), gradation, level, K, _p(lab), _p(out))
        return float(out[0]), float(out[1]), float(out[2])

    def ring_problem(self, labels, item, cluster):
        lab = np.ascontiguousarray(labels, np.int32)
        return lib().orc_ring_problem(self.h, _p(lab), int(item), int(cluster))


def set_fixed(items):
    """Fixed clusters of the following oracle runs, samplings and replays
    (cluster k holds item items[k]; the QEM clusters are anchored there);
    None or [] clears."""
    ids = np.zeros(0, np.int32) if items is None else np.ascontiguousarray(items, np.int32)
    lib().orc_set_fixed(_p(ids) if len(ids) else None, len(ids))


def set_frozen(flags):
    """Frozen clusters (IsClusterFreezed) of the following serial runs and
    replays whose cluster count equals len(flags); None clears."""
    f = np.zeros(0, np.int32) if flags is None else np.ascontiguousarray(flags, np.int32)
    lib().orc_set_frozen(_p(f) if len(f) else None, len(f))


def set_energy_log(on):
    """Serial runs (oracle A) record ComputeGlobalEnergy after each loop and
    after the run (SetWriteToGlobalEnergyLog); read with serial_energy_log()."""
    lib().orc_set_energy_log(1 if on else 0)


def serial_energy_log():
    """(energy per loop, final energy) of the last serial run."""
    n = lib().orc_serial_energy_log(0, None, None)
    e = np.empty(n, np.float64)
    fin = np.zeros(1, np.float64)
    lib().orc_serial_energy_log(n, _p(e), _p(fin))
    return e, float(fin[0])


def set_progress(on):
    """per-loop progress lines of the serial runs on stderr"""
    lib().orc_set_progress(int(bool(on)))


def rep_point(Q9, P, level=3, thr=1e-3):
    Q = np.ascontiguousarray(Q9, np.float64)
    p = np.array(P, np.float64, copy=True)
    rd = lib().orc_rep_point(_p(Q), _p(p), level, thr)
    return p, rd


def sym_eigen(A):
    A = np.ascontiguousarray(A, np.float64).reshape(9)
    lam = np.empty(3)
    V = np.empty(9)
    lib().orc_sym_eigen(_p(A), _p(lam), _p(V))
    return lam, V.reshape(3, 3)


def triangle_quadric(x1, x2, x3):
    out = np.empty(10)
    a, b, c = (np.ascontiguousarray(x, np.float64) for x in (x1, x2, x3))
    lib().orc_triangle_quadric(_p(a), _p(b), _p(c), _p(out))
    return out


def hdr_rep_point_delta(Q9, c, level=3, V0=None, want_V=False):
    """shared-header solve: (D, rank deficiency) or, with want_V, (D, rd, V)
    where V's columns are the eigenvectors; V0 warm-starts the solve"""
    Q = np.ascontiguousarray(Q9, np.float64)
    cc = np.ascontiguousarray(c, np.float64)
    D = np.empty(3)
    V0a = None if V0 is None else np.ascontiguousarray(V0, np.float64).reshape(9)
    V = np.empty(9) if want_V else None
    rd = lib().orc_hdr_rep_point_delta(_p(Q), _p(cc), level, _p(D), _p(V0a), _p(V))
    return (D, rd, V.reshape(3, 3)) if want_V else (D, rd)


def hdr_ring_problem(row_ptr, col, labels, item, cluster):
    return lib().orc_hdr_ring_problem(_p(np.ascontiguousarray(row_ptr, np.int32)),
                                      _p(np.ascontiguousarray(col, np.int32)),
                                      _p(np.ascontiguousarray(labels, np.int32)), int(item), int(cluster))


def hdr_ring_problem_mask(row_ptr, col, labels, item, cluster):
    """ring check through the ring adjacency mask (-1: row too long for it)"""
    return lib().orc_hdr_ring_problem_mask(_p(np.ascontiguousarray(row_ptr, np.int32)),
                                           _p(np.ascontiguousarray(col, np.int32)),
                                           _p(np.ascontiguousarray(labels, np.int32)), int(item), int(cluster))


def hdr_accept_key(seed, loop, item, round=0):
    return lib().orc_hdr_accept_key(seed, loop, round, item)


def replay_run(row_ptr, col, xw, q, scales, K, metric, init, level=3, seed=0, unconstrained=0, max_loops=5000000,
               max_conv=1000000000, generic_qem=False, all_rounds=False, first_rounds=0, energy_log=False):
    """Replay C.  With energy_log the stats also hold "energy_log" (energy
    after each pass) and "final_energy", summed as the library sums them."""
    n = len(row_ptr) - 1
    out = np.empty(n, np.int32)
    st = np.zeros(6, np.int64)
    rp = np.ascontiguousarray(row_ptr, np.int32)
    cl = np.ascontiguousarray(col, np.int32)
    x = np.ascontiguousarray(xw, np.float32)
    qq = None if q is None else np.ascontiguousarray(q, np.float32)
    sc = np.ascontiguousarray(scales, np.float64)
    ini = np.ascontiguousarray(init, np.int32)
    path = int(generic_qem) | (int(all_rounds) << 1) | (int(first_rounds) << 8) | (int(bool(energy_log)) << 16)
    lib().orc_replay_run_path(path, n, _p(rp), _p(cl), _p(x), _p(qq), _p(sc), K, metric, level, seed,
                              unconstrained, max_loops, max_conv, _p(ini), _p(out), _p(st))
    keys = ["loops", "convergences", "total_moves", "evaluated", "proposed", "last_disconnected"]
    stats = dict(zip(keys, (int(v) for v in st)))
    stats["non_adjacent_moves"] = int(lib().orc_replay_nonadjacent())
    if energy_log:
        m = lib().orc_replay_energy_log(0, None, None)
        e = np.empty(m, np.float64)
        fin = np.zeros(1, np.float64)
        lib().orc_replay_energy_log(m, _p(e), _p(fin))
        stats["energy_log"] = e
        stats["final_energy"] = float(fin[0])
    return out, stats


def device_inputs(xyz, tri, metric, indicator=None, gradation=0.0):
    """Host build of the sweep inputs: (row_ptr, col, xw[n,4], q[n,10], scales[3])."""
    x = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(tri, np.int32).reshape(-1, 3)
    ind = None if indicator is None else np.ascontiguousarray(indicator, np.float64)
    h = lib().orc_devinputs_create(len(x), _p(x), len(t), _p(t), metric, _p(ind), gradation)
    try:
        ncol = lib().orc_devinputs_ncol(h)
        rp = np.empty(len(x) + 1, np.int32)
        col = np.empty(ncol, np.int32)
        xw = np.empty((len(x), 4), np.float32)
        q = np.empty((len(x), 10), np.float32)
        sc = np.empty(3, np.float64)
        lib().orc_devinputs_get(h, _p(rp), _p(col), _p(xw), _p(q), _p(sc))
    finally:
        lib().orc_devinputs_destroy(h)
    return rp, col, xw, q, sc


def hdr_terms(xw4, ca, Wa, cb, Wb):
    x = np.ascontiguousarray(xw4, np.float32)
    out = np.empty(2, np.float32)
    lib().orc_hdr_terms(_p(x), _p(np.ascontiguousarray(ca, np.float64)), float(Wa),
                        _p(np.ascontiguousarray(cb, np.float64)), float(Wb), _p(out))
    return float(out[0]), float(out[1])


def hdr_qem_T(ws4, q10, scales, level=3):
    q = np.empty(3)
    T = lib().orc_hdr_qem_T(_p(np.ascontiguousarray(ws4, np.int64)), _p(np.ascontiguousarray(q10, np.int64)),
                            _p(np.ascontiguousarray(scales, np.float64)), level, _p(q))
    return T, q


def hdr_fix_item(xw4, q10, scales):
    out = np.empty(14, np.int64)
    qq = None if q10 is None else np.ascontiguousarray(q10, np.float32)
    lib().orc_hdr_fix_item(_p(np.ascontiguousarray(xw4, np.float32)), _p(qq),
                           _p(np.ascontiguousarray(scales, np.float64)), _p(out))
    return out


def minstd_10000():
    return int(lib().orc_minstd_10000())


def mt19937_10000():
    return int(lib().orc_mt19937_10000())


# ---- output mesh (small meshes only: pure Python loops) -------------------
def output_triangles(tri, labels, K):
    """Triangles of the output mesh, vertex clustering branch of
    BuildDelaunayTriangulation (DiscreteRemeshing/vtkDiscreteRemeshing.h:1041,
    1096-1108): per face in order, the clusters of its vertices with
    duplicates and the null label K removed in first-occurrence order
    (GetDualItemNeighbourClusters :974-988); three of them give a triangle
    unless one on the same three clusters exists already (AddFace :992-1000,
    vtkSurfaceBase::IsFace Common/vtkSurfaceBase.cxx:723-763 matches the
    vertex set regardless of orientation)."""
    seen = set()
    out = []
    for f in np.asarray(tri):
        ids = []
        for v in f:
            c = int(labels[int(v)])
            if 0 <= c < K and c not in ids:
                ids.append(c)
        if len(ids) == 4:  # unreachable for triangles
            raise AssertionError
        if len(ids) == 3:
            key = frozenset(ids)
            if key not in seen:
                seen.add(key)
                out.append(ids)
    return np.asarray(out, np.int32).reshape(-1, 3)


def output_points(points, sizes):
    """Output vertices (vtkDiscreteRemeshing.h:1012-1035): the cluster point,
    an empty cluster taking the point of the first non-empty cluster."""
    pts = np.array(points, np.float64, copy=True)
    nonempty = np.flatnonzero(np.asarray(sizes) > 0)
    valid = int(nonempty[0]) if len(nonempty) else 0
    for i in np.flatnonzero(np.asarray(sizes) == 0):
        pts[i] = pts[valid]
    return pts


def quadric_post_placement(xyz, tri, labels, K, points, level):
    """ACVD quadric post-processing (DiscreteRemeshing/Examples/ACVD.cxx:237-276):
    each cluster sums the plane quadrics of the faces around each of its
    vertices (AddTriangleQuadric, Common/vtkQuadricTools.cxx:56-79), then
    ComputeRepresentativePoint moves its output vertex."""
    xyz = np.asarray(xyz, np.float64)
    tri = np.asarray(tri)
    fq = np.array([triangle_quadric(xyz[a], xyz[b], xyz[c]) for a, b, c in tri])
    vq = np.zeros((len(xyz), 10))
    for j in range(3):
        np.add.at(vq, tri[:, j], fq)
    cq = np.zeros((K, 10))
    lab = np.asarray(labels)
    ok = (lab >= 0) & (lab < K)
    np.add.at(cq, lab[ok], vq[ok])
    out = np.array(points, np.float64, copy=True)
    for c in range(K):
        out[c], _ = rep_point(cq[c, :9], out[c], level)
    return out


# ---- fixed-assignment energy reference (oracle/energy_ref.h)
class EnergyRef:
    """FP64 fixed-assignment energy of one mesh (isotropic CVT or QEM) with
    dual-number derivatives and central differences per input coordinate
    (3*i+k: x_i[k]; 3*nv+i: w_i)."""

    def __init__(self, nv, tri, labels, K, metric, level=3):
        L = lib()
        L.orc_energy_create.restype = ctypes.c_void_p
        L.orc_energy_create.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_energy_destroy.argtypes = [ctypes.c_void_p]
        L.orc_energy_total.restype = ctypes.c_double
        L.orc_energy_total.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_energy_derivatives.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                             ctypes.c_void_p, ctypes.c_void_p]
        L.orc_energy_central_differences.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_int64, ctypes.c_void_p, ctypes.c_double,
                                                     ctypes.c_void_p]
        self.nv = int(nv)
        t = np.ascontiguousarray(tri, np.int32)
        lab = np.ascontiguousarray(labels, np.int32)
        self.h = L.orc_energy_create(self.nv, len(t), _p(t), _p(lab), int(K), int(metric), int(level))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_energy_destroy(self.h)
            self.h = None

    def total(self, x, w):
        x = np.ascontiguousarray(x, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        return lib().orc_energy_total(self.h, _p(x), _p(w))

    def derivatives(self, x, w, coords):
        x = np.ascontiguousarray(x, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        c = np.ascontiguousarray(coords, np.int64)
        out = np.empty(len(c))
        lib().orc_energy_derivatives(self.h, _p(x), _p(w), len(c), _p(c), _p(out))
        return out

    def central_differences(self, x, w, coords, step=1e-6):
        x = np.ascontiguousarray(x, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        c = np.ascontiguousarray(coords, np.int64)
        out = np.empty(len(c))
        lib().orc_energy_central_differences(self.h, _p(x), _p(w), len(c), _p(c), float(step), _p(out))
        return out
