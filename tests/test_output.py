"""Output mesh of a clustering (GetOutput after Remesh, and the ACVD quadric
post-placement).

CPU: the output-triangle restatement against what the definition fixes on
hand-made inputs (identity clustering returns the input faces; faces on the
same three clusters in either orientation give one triangle; null and
repeated labels give none).  GPU: triangles equal the restatement exactly
(same faces, same order, same vertex order); points equal the cluster points
with the empty-cluster rule; quadric placement matches the FP64 restatement
to the precision of the FP32 vertex quadrics the device sums (rtol 1e-5 of
the mesh scale), and the level-3 placement puts corner clusters of a cube on
the corner."""
import numpy as np
import pytest

import acvd_amd
from acvd_amd import meshgen
from oracle import pyoracle as po

ISO, QEM = acvd_amd.METRIC_ISOTROPIC, acvd_amd.METRIC_QEM


def test_identity_clustering_returns_input_faces():
    xyz, tri = meshgen.icosphere(2, 1, 0.0)
    n = len(xyz)
    out = po.output_triangles(tri, np.arange(n), n)
    assert np.array_equal(out, tri)


def test_duplicate_triple_kept_once_first_face_order():
    tri = np.array([[0, 1, 2], [3, 5, 4], [0, 2, 1], [6, 7, 8]])
    lab = np.array([0, 1, 2, 0, 1, 2, 5, 5, 1])  # faces 0,1,2 on {0,1,2}; face 3 two clusters
    out = po.output_triangles(tri, lab, 6)
    assert out.tolist() == [[0, 1, 2]]
    out = po.output_triangles(tri[[1, 0, 2]], lab, 6)
    assert out.tolist() == [[0, 2, 1]]  # vertex order of the first face (3, 5, 4)


def test_null_label_gives_no_triangle():
    tri = np.array([[0, 1, 2], [1, 2, 3]])
    lab = np.array([0, 1, 4, 2])  # K = 4 is null
    assert po.output_triangles(tri, lab, 4).shape == (0, 3)


def test_output_points_empty_cluster_rule():
    pts = np.array([[9.0, 9, 9], [1, 2, 3], [4, 5, 6], [7, 8, 9]])
    out = po.output_points(pts, [0, 3, 0, 2])
    assert out.tolist() == [[1, 2, 3], [1, 2, 3], [1, 2, 3], [7, 8, 9]]


def test_post_placement_level1_projects_on_dominant_plane():
    # single planar cluster: level 1 keeps the point on the plane z = 0.5
    xyz = np.array([[0, 0, 0.5], [1, 0, 0.5], [0, 1, 0.5], [1, 1, 0.5]])
    tri = np.array([[0, 1, 2], [1, 3, 2]])
    out = po.quadric_post_placement(xyz, tri, np.zeros(4, int), 1, np.array([[0.3, 0.4, 0.9]]), 1)
    assert np.allclose(out[0], [0.3, 0.4, 0.5])


def _gpu_run(xyz, tri, metric, K, seed=3, level=3):
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=K, metric=metric, unconstrained_initialization=int(metric == QEM),
                   seed=seed, quadrics_optimization_level=level)
    ctx.set_initial_labels(0, meshgen.seed_labels(len(xyz), K, seed))
    ctx.process()
    return ctx


@pytest.mark.gpu
@pytest.mark.parametrize("metric", [ISO, QEM])
def test_gpu_output_triangles_and_points(gpu_ready, metric):
    xyz, tri = meshgen.icosphere(5, 1, 0.02)
    K = 400
    ctx = _gpu_run(xyz, tri, metric, K)
    labels = ctx.labels(0)
    out_xyz, out_tri = ctx.output(0, 0)
    assert np.array_equal(out_tri, po.output_triangles(tri, labels, K))
    pts = po.output_points(ctx.cluster_points(0), ctx.cluster_sizes(0))
    assert np.array_equal(out_xyz, pts)
    # a clean clustering of a sphere gives a closed 2-manifold-like dual: F ~ 2K
    assert abs(len(out_tri) - 2 * K) < 0.05 * K


@pytest.mark.gpu
def test_gpu_quadric_post_placement_matches_restatement(gpu_ready):
    xyz, tri = meshgen.icosphere(5, 1, 0.02)
    K = 400
    ctx = _gpu_run(xyz, tri, ISO, K)
    labels = ctx.labels(0)
    base, _ = ctx.output(0, 0)
    for level in (1, 2, 3):
        placed, _ = ctx.output(0, level)
        ref = po.quadric_post_placement(xyz, tri, labels, K, base, level)
        assert np.max(np.abs(placed - ref)) < 1e-5, level


@pytest.mark.gpu
def test_gpu_quadric_post_placement_cube_corners(gpu_ready):
    xyz, tri = meshgen.cube(20, 2, 0.0)
    K = 150
    ctx = _gpu_run(xyz, tri, QEM, K)
    labels = ctx.labels(0)
    placed, _ = ctx.output(0, 3)
    corners = np.flatnonzero(np.all(np.abs(np.abs(xyz) - 1.0) < 1e-12, axis=1))
    assert len(corners) == 8
    for v in corners:
        c = labels[v]
        assert np.max(np.abs(placed[c] - xyz[v])) < 1e-5


@pytest.mark.gpu
def test_gpu_output_errors(gpu_ready):
    xyz, tri = meshgen.icosphere(3, 1, 0.0)
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=50)
    with pytest.raises(acvd_amd.AcvdError):
        ctx.output(0, 0)  # no clustering yet
    ctx.process()
    with pytest.raises(acvd_amd.AcvdError):
        ctx.output(0, 4)


@pytest.mark.gpu
def test_gpu_remesh_class_api(gpu_ready):
    xyz, tri = meshgen.icosphere(5, 1, 0.02)
    r = acvd_amd.IsotropicDiscreteRemeshing()
    r.SetInput(acvd_amd.Mesh(xyz, tri))
    r.SetNumberOfClusters(300)
    r.SetConsoleOutput(0)
    r.Remesh()
    out = r.GetOutput()
    assert out.points.shape == (300, 3)
    assert np.array_equal(out.triangles, po.output_triangles(tri, r.GetClustering(), 300))
    r.QuadricPostProcessing(1)
    assert r.GetOutput().points.shape == (300, 3)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["iso", "qem"])
def test_gpu_cpp_dropin_program(gpu_ready, mode):
    """include/acvd/remeshing.hpp used as ACVD.cxx / ACVDQ.cxx use the
    reference classes: icosphere L4 (2562 vertices) -> 200 clusters."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "dropin")
    r = subprocess.run([exe] + (["q"] if mode == "qem" else []), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    tag, nv, nf, nnull = r.stdout.split()
    assert tag == "ok" and int(nv) == 200 and int(nnull) == 0
    assert abs(int(nf) - 396) <= 20
