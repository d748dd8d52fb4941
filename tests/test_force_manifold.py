"""ForceManifold (vtkDiscreteRemeshing::Remesh, DiscreteRemeshing/
vtkDiscreteRemeshing.h:937-946, 165-383, 1114-1131) and frozen clusters
(IsClusterFreezed, Common/vtkUniformClustering.h:329, 733-736, 910-919).

CPU: the library's vertex manifold test and its detection step equal a
pure-Python restatement written after the reference's edge table and walk
(oracle/manifold.py), which is itself pinned by what topology fixes (a closed
sphere is manifold everywhere, an open patch is not on its boundary, a bowtie
vertex and an edge of three faces are not); the serial restatement with
frozen clusters never moves an item in or out of a frozen cluster.
GPU: a run with frozen clusters equals the replay bit for bit; one restart of
acvd_force_manifold equals the detection step followed by that run; the loop
ends with an output the detection passes."""
import numpy as np
import pytest

import acvd_amd
from acvd_amd import meshgen
from acvd_amd import remeshing as rem
from oracle import manifold as om
from oracle import pyoracle as po

ISO, QEM = acvd_amd.METRIC_ISOTROPIC, acvd_amd.METRIC_QEM


@pytest.fixture(autouse=True)
def reset_oracle():
    yield
    po.set_frozen(None)


def voronoi(xyz, K, seed):
    rng = np.random.default_rng(seed)
    seeds = xyz[rng.choice(len(xyz), K, replace=False)]
    d = ((xyz[:, None, :] - seeds[None, :, :]) ** 2).sum(-1)
    return d.argmin(1).astype(np.int32)


def connected_voronoi(xyz, tri, K, seed):
    """Voronoi labels cleaned and filled by the restatement (CleanClustering,
    FillHolesInClustering): every cluster connected, so that no later clean
    moves the items of a frozen cluster."""
    m = po.RefMesh(xyz, tri)
    lab, _ = m.clean(K, voronoi(xyz, K, seed))
    lab = m.fill_holes(K, lab)
    assert m.clean(K, lab)[1] == 0
    return lab


def enclose(xyz, tri, lab, K):
    """Relabels a vertex deep inside one cluster and its ring as cluster K:
    a cluster enclosed by another (K + 1 clusters)."""
    s = om.Surface(len(xyz), tri)
    for v in range(len(xyz)):
        ring = s.neighbours(v)
        two = set(ring)
        for u in ring:
            two.update(s.neighbours(u))
        if all(lab[u] == lab[v] for u in two):
            out = lab.copy()
            out[[v] + ring] = K
            return out
    raise AssertionError("no interior vertex")


# ------------------------------------------------------------------ CPU
def grid(n):
    xyz = np.array([[i, j, 0.0] for j in range(n) for i in range(n)])
    tri = []
    for j in range(n - 1):
        for i in range(n - 1):
            a, b, c, d = j * n + i, j * n + i + 1, (j + 1) * n + i + 1, (j + 1) * n + i
            tri += [[a, b, c], [a, c, d]]
    return xyz, np.array(tri, np.int32)


def test_manifold_restatement_pins():
    xyz, tri = meshgen.icosphere(3, 1, 0.0)
    s = om.Surface(len(xyz), tri)
    assert all(s.is_vertex_manifold(v) for v in range(len(xyz)))
    gx, gt = grid(5)
    s = om.Surface(len(gx), gt)
    for v in range(len(gx)):
        i, j = v % 5, v // 5
        assert s.is_vertex_manifold(v) == (0 < i < 4 and 0 < j < 4)
    bowtie = np.array([[0, 1, 2], [0, 3, 4]])
    s = om.Surface(5, bowtie)
    assert not s.is_vertex_manifold(0)
    fin = np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]])  # edge (0,1) in three faces
    s = om.Surface(5, fin)
    assert not s.is_vertex_manifold(0) and not s.is_vertex_manifold(1)


def test_vertex_manifold_equals_restatement():
    xyz, tri = meshgen.icosphere(3, 1, 0.0)
    nv = len(xyz)
    rng = np.random.default_rng(0)
    cases = [(nv, tri), (36, grid(6)[1]), (8, np.array([[0, 1, 2], [0, 3, 4], [5, 6, 7]]))]
    for _ in range(6):
        cases.append((nv, tri[rng.random(len(tri)) > 0.15]))
    for n, t in cases:
        got = rem.vertex_manifold(n, t)
        s = om.Surface(n, t)
        want = np.array([s.is_vertex_manifold(v) for v in range(n)])
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("K,seed,enclosed,noise", [(30, 1, False, 0.0), (120, 3, False, 0.0), (60, 5, True, 0.0),
                                                    (80, 6, False, 0.05), (40, 7, True, 0.02)])
def test_detection_step_equals_restatement(K, seed, enclosed, noise):
    xyz, tri = meshgen.icosphere(4, 2, 0.02)
    xyz = np.asarray(xyz, np.float64)
    lab = voronoi(xyz, K, seed)
    Kc = K
    if enclosed:
        lab = enclose(xyz, tri, lab, K)
        Kc = K + 1
    if noise:
        rng = np.random.default_rng(seed)
        pick = rng.random(len(lab)) < noise
        lab[pick] = rng.integers(0, Kc, pick.sum())
    out = po.output_triangles(tri, lab, Kc)
    n, lab2, K2, fr = rem.force_manifold_step(len(xyz), tri, lab, Kc, out)
    rn, rlab, rK, rfr = om.force_manifold_step(len(xyz), tri, lab, Kc, out)
    assert n == rn and K2 == rK
    np.testing.assert_array_equal(lab2, rlab)
    np.testing.assert_array_equal(fr, rfr)
    if enclosed:
        assert n >= 1


def test_serial_frozen_clusters_keep_their_items():
    xyz, tri = meshgen.icosphere(4, 3, 0.02)
    xyz = np.asarray(xyz, np.float64)
    K = 60
    lab = connected_voronoi(xyz, tri, K, 11)
    frozen = np.ones(K, np.int32)
    frozen[::3] = 0
    po.set_frozen(frozen)
    m = po.RefMesh(xyz, tri)
    out, st = m.serial_run(ISO, K, init=lab)
    was_frozen = frozen[lab] == 1
    np.testing.assert_array_equal(out[was_frozen], lab[was_frozen])
    assert np.all(frozen[out[~was_frozen]] == 0)
    assert st["total_moves"] > 0


# ------------------------------------------------------------------ GPU
def _run(xyz, tri, metric, K, init, frozen, unc=0, seed=2, **kw):
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=K, metric=metric, unconstrained_initialization=unc, seed=seed, **kw)
    ctx.set_initial_labels(0, init)
    ctx.set_frozen_clusters(0, frozen)
    ctx.process()
    return ctx


@pytest.mark.gpu
@pytest.mark.parametrize("metric", [ISO, QEM])
def test_device_frozen_equals_replay(gpu_ready, metric):
    xyz, tri = meshgen.icosphere(5, 4, 0.03)
    xyz = np.asarray(xyz, np.float64)
    K = 300
    lab = connected_voronoi(xyz, tri, K, 12)
    frozen = (np.random.default_rng(3).random(K) < 0.7).astype(np.int32)
    print(f"frozen clusters: metric {metric}, {frozen.sum()} of {K} frozen", flush=True)
    ctx = _run(xyz, tri, metric, K, lab, frozen)
    got = ctx.labels(0)
    rp, col = ctx.adjacency(0)
    xw, q = ctx.items(0)
    po.set_frozen(frozen)
    rlab, rst = po.replay_run(rp, col, xw, q, ctx.scales(0), K, metric, lab, seed=2, unconstrained=0)
    np.testing.assert_array_equal(got, rlab)
    assert ctx.stats(0)["loops"] == rst["loops"]
    was_frozen = frozen[lab] == 1
    np.testing.assert_array_equal(got[was_frozen], lab[was_frozen])


@pytest.mark.gpu
def test_force_manifold_restart_equals_step_then_run(gpu_ready):
    xyz, tri = meshgen.icosphere(4, 5, 0.02)
    xyz = np.asarray(xyz, np.float64)
    K0 = 60
    lab = enclose(xyz, tri, voronoi(xyz, K0, 13), K0)
    K = K0 + 1
    # one pass, then the first convergence ends the run: the enclosed cluster stays
    kw = dict(max_number_of_loops=0)
    ctx = _run(xyz, tri, ISO, K, lab, None, **kw)
    first = ctx.labels(0)
    first[first == K] = -1
    xo, to = ctx.output(0, 0)
    n, lab2, K2, fr = rem.force_manifold_step(len(xyz), tri, first, K, to)
    print(f"force manifold: {n} issues, {K} -> {K2} clusters", flush=True)
    assert n >= 1 and K2 > K
    manual = _run(xyz, tri, ISO, K2, lab2, fr, **kw).labels(0)
    r, left = ctx.force_manifold(0, max_restarts=1)
    assert r == 1
    assert ctx.get_params(0)["number_of_clusters"] == K2
    np.testing.assert_array_equal(ctx.labels(0), manual)
    # a later plain run clears the frozen flags (Init)
    ctx.process()


@pytest.mark.gpu
def test_force_manifold_loop_ends_manifold(gpu_ready):
    xyz, tri = meshgen.icosphere(5, 6, 0.02)
    r = acvd_amd.IsotropicDiscreteRemeshing.New()
    r.SetInput(acvd_amd.Mesh(xyz, tri))
    r.SetNumberOfClusters(400)
    r.SetForceManifold(1)
    r.Remesh()
    restarts, left = r._restarts
    assert left == 0
    out = r.GetOutput()
    K = r.GetNumberOfClusters()
    lab = r.GetClustering().copy()
    n, _, _, _ = rem.force_manifold_step(len(xyz), tri, lab, K, out.triangles)
    assert n == 0
    r.Delete()
