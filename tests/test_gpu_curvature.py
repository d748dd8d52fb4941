"""GPU: the curvature indicator computed by the build (params
curvature_indicator, vtkCurvatureMeasure restated in k_build.hip) against the
CPU oracle (oracle/curvature.h) on closed, open and irregular meshes; the
item weights it gives; and Remesh() turning it on as the reference's
SamplingPreProcessing does (vtkDiscreteRemeshing.h:583-653).
Tolerance: FP64 on both sides with different summation orders and solvers
(Cholesky of the normal equations here, their LU inverse in the oracle):
rtol 1e-7."""
import numpy as np
import pytest

import acvd_amd
from acvd_amd import meshgen
from oracle import pyoracle as po
from test_curvature_oracle import cylinder, grid
from test_gpu_parity import centroid_split

pytestmark = pytest.mark.gpu

QEM, ISO = acvd_amd.METRIC_QEM, acvd_amd.METRIC_ISOTROPIC


def meshes():
    sph = np.asarray(meshgen.icosphere(5, 1, 0.0)[0], np.float64)
    sph /= np.linalg.norm(sph, axis=1)[:, None]
    P, T, _ = cylinder(0.5, 48, 30)
    return {
        "sphere": (sph, meshgen.icosphere(5, 1, 0.0)[1]),
        "ico4-noisy": meshgen.icosphere(4, 2, 0.05),
        "torus": meshgen.torus(128, 64, 3),
        "cube": meshgen.cube(40, 2, 1e-4),
        "split-ico": centroid_split(*meshgen.icosphere(3, 9, 0.02)),
        "open-grid": grid(40, 30, lambda u, v: np.stack([u * 0.05, v * 0.05, 0.3 * np.sin(u * 0.2) * np.cos(v * 0.15)], 1)),
        "open-cylinder": (P, T),
    }


def device_indicator(xyz, tri, metric=QEM, gradation=1.0):
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=max(1, len(xyz) // 20), metric=metric, gradation=gradation,
                   curvature_indicator=1)
    ctx.build()
    return ctx, ctx.indicator(0)


@pytest.mark.parametrize("name", list(meshes()))
def test_indicator_equals_the_oracle(gpu_ready, name):
    xyz, tri = meshes()[name]
    print("curvature", name, len(xyz), flush=True)
    _, k = device_indicator(xyz, tri)
    ref = po.RefMesh(np.asarray(xyz, np.float64), tri).curvature_indicator()
    np.testing.assert_allclose(k, ref, rtol=1e-7, atol=1e-12)


@pytest.mark.parametrize("metric", [ISO, QEM])
def test_weights_use_the_computed_indicator(gpu_ready, metric):
    xyz, tri = meshgen.torus(128, 64, 3)
    ctx, k = device_indicator(xyz, tri, metric, 1.5)
    ref = po.RefMesh(np.asarray(xyz, np.float64), tri).curvature_indicator()
    _, _, hxw, _, _ = po.device_inputs(np.asarray(xyz, np.float64), tri, metric, ref, 1.5)
    xw, _ = ctx.items(0)
    np.testing.assert_allclose(xw[:, 3], hxw[:, 3], rtol=1e-6)


def test_no_indicator_without_gradation_or_flag(gpu_ready):
    xyz, tri = meshgen.icosphere(4, 2, 0.05)
    for kw in [dict(gradation=0.0, curvature_indicator=1), dict(gradation=1.0, curvature_indicator=0)]:
        ctx = acvd_amd.Context()
        ctx.set_input(xyz, tri)
        ctx.set_params(0, number_of_clusters=100, metric=QEM, **kw)
        ctx.build()
        with pytest.raises(acvd_amd.AcvdError):
            ctx.indicator(0)
    # a set indicator wins over the computed one
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=100, metric=QEM, gradation=1.0, curvature_indicator=1)
    mine = np.linspace(1.0, 2.0, len(xyz))
    ctx.set_indicator(0, mine)
    ctx.build()
    np.testing.assert_array_equal(ctx.indicator(0), mine)


def test_remesh_weights_by_curvature(gpu_ready):
    xyz, tri = meshgen.torus(128, 64, 3)
    K = 300
    rem = acvd_amd.QIsotropicDiscreteRemeshing.New()
    rem.SetInput((xyz, tri))
    rem.SetNumberOfClusters(K)
    rem.GetMetric().SetGradation(2.0)
    rem.SetUnconstrainedInitialization(1)
    rem.Remesh()
    ctx = rem.GetContext()
    k = ctx.indicator(0)
    ref = po.RefMesh(np.asarray(xyz, np.float64), tri).curvature_indicator()
    np.testing.assert_allclose(k, ref, rtol=1e-7, atol=1e-12)
    # the clustering is the replay's on the weights it was given
    rp, col = ctx.adjacency(0)
    xw, q = ctx.items(0)
    lab = ctx.labels(0)
    init = po.RefMesh(xyz, tri).initial_sampling(QEM, K, 2.0, k)
    rlab, _ = po.replay_run(rp, col, xw, q, ctx.scales(0), K, QEM, init, level=3, seed=0, unconstrained=1)
    np.testing.assert_array_equal(lab, rlab)
    assert len(rem.GetOutput().points) == K
    # ProcessClustering alone (no SamplingPreProcessing) leaves the items unweighted
    rem2 = acvd_amd.QIsotropicDiscreteRemeshing.New()
    rem2.SetInput((xyz, tri))
    rem2.SetNumberOfClusters(K)
    rem2.GetMetric().SetGradation(2.0)
    rem2.ProcessClustering()
    with pytest.raises(acvd_amd.AcvdError):
        rem2.GetContext().indicator(0)
    rem.Delete()
    rem2.Delete()


def test_batch_computes_the_indicator_per_mesh(gpu_ready):
    # mesh 0 asks for the curvature indicator, mesh 1 has a set one, mesh 2 none
    m0 = meshgen.torus(96, 48, 5)
    m1 = meshgen.icosphere(4, 2, 0.05)
    m2 = meshgen.icosphere(4, 3, 0.05)
    ctx = acvd_amd.Context()
    ids = [ctx.add_input(*m) for m in (m0, m1, m2)]
    ctx.set_params(ids[0], number_of_clusters=200, metric=QEM, gradation=1.0, curvature_indicator=1)
    ctx.set_params(ids[1], number_of_clusters=150, metric=ISO, gradation=1.0, curvature_indicator=1)
    ctx.set_params(ids[2], number_of_clusters=150, metric=ISO, gradation=0.0, curvature_indicator=1)
    mine = np.linspace(0.5, 1.5, len(m1[0]))
    ctx.set_indicator(ids[1], mine)
    ctx.build()
    ref0 = po.RefMesh(np.asarray(m0[0], np.float64), m0[1]).curvature_indicator()
    np.testing.assert_allclose(ctx.indicator(ids[0]), ref0, rtol=1e-7, atol=1e-12)
    np.testing.assert_array_equal(ctx.indicator(ids[1]), mine)
    with pytest.raises(acvd_amd.AcvdError):
        ctx.indicator(ids[2])
    # each mesh's weights are those of its own inputs
    for mid, (xyz, tri), metric, ind, grad in [(ids[0], m0, QEM, ref0, 1.0), (ids[1], m1, ISO, mine, 1.0),
                                                (ids[2], m2, ISO, None, 0.0)]:
        _, _, hxw, _, _ = po.device_inputs(np.asarray(xyz, np.float64), tri, metric, ind, grad)
        xw, _ = ctx.items(mid)
        np.testing.assert_allclose(xw[:, 3], hxw[:, 3], rtol=1e-6)
