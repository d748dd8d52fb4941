"""The reference's initial sampling (type 1, ComputeInitialRandomSampling,
Common/vtkUniformClustering.h:1177-1316) on a built context: the rings in
edge-creation order and the vertex areas come from the device, the weights
and the region growing from the host.  The labels equal the oracle's
restatement (and the product's host-only pass) item for item: graded and
ungraded, isotropic and QEM, several meshes packed in one context (each mesh
sampled on its own), and meshes with an inactive face (first two vertices
equal) and a face listing a vertex twice."""
import numpy as np
import pytest

import acvd_amd
from acvd_amd import meshgen
from acvd_amd import remeshing as rem
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ISO, QEM = acvd_amd.METRIC_ISOTROPIC, acvd_amd.METRIC_QEM


def sampled(meshes):
    """Context labels after initial_sampling(), before the device fill:
    ctx.labels() holds the region growing's labels for the items it reached."""
    ctx = acvd_amd.Context()
    for xyz, tri, K, metric, ind, g in meshes:
        mid = ctx.add_input(xyz, tri)
        ctx.set_params(mid, number_of_clusters=K, metric=metric, gradation=g, initial_sampling=0)
        if ind is not None:
            ctx.set_indicator(mid, ind)
    ctx.initial_sampling()
    out = [ctx.labels(m) for m in range(len(meshes))]
    ctx.close()
    return out


def cases():
    xi, ti = meshgen.icosphere(5, 2, 0.02)
    xc, tc = meshgen.cube(40, 2, 1e-4)
    xt, tt = meshgen.torus(192, 96, seed=4)
    ht = meshgen.mean_curvature(xt, tt)
    return {
        "ico-iso": (xi, ti, 500, ISO, None, 0.0),
        "cube-qem": (xc, tc, 800, QEM, None, 0.0),
        "torus-qem-graded": (xt, tt, 700, QEM, ht, 1.5),
        "torus-iso-graded": (xt, tt, 900, ISO, ht, 1.0),
    }


@pytest.mark.parametrize("name", list(cases()))
def test_device_topology_sampling_equals_oracle(gpu_ready, name):
    xyz, tri, K, metric, ind, g = cases()[name]
    got = sampled([(xyz, tri, K, metric, ind, g)])[0]
    want = po.RefMesh(xyz, tri).initial_sampling(metric, K, g, ind)
    reached = want < K
    print(f"{name}: {len(xyz)} vertices, K={K}, {reached.sum()} items grown", flush=True)
    np.testing.assert_array_equal(got[reached], want[reached])


def test_packed_meshes_are_sampled_one_by_one(gpu_ready):
    c = cases()
    meshes = [c["torus-qem-graded"], c["ico-iso"], c["cube-qem"]]
    got = sampled(meshes)
    for j, (xyz, tri, K, metric, ind, g) in enumerate(meshes):
        want = po.RefMesh(xyz, tri).initial_sampling(metric, K, g, ind)
        reached = want < K
        np.testing.assert_array_equal(got[j][reached], want[reached], err_msg=f"mesh {j}")


def test_degenerate_faces_equal_the_host_pass(gpu_ready):
    xyz, tri = meshgen.icosphere(4, 3, 0.02)
    n = len(xyz)
    extra = np.array([[4, 4, 9], [10, 11, 10], [20, 21, 22]], np.int32)  # inactive, vertex twice, duplicate face
    tri = np.concatenate([tri, extra, tri[:1]]).astype(np.int32)
    K = 150
    got = sampled([(xyz, tri, K, QEM, None, 0.0)])[0]
    want = rem.reference_initial_sampling(xyz, tri, K, QEM, None, 0.0)
    reached = want < K
    assert reached.sum() > n // 2
    np.testing.assert_array_equal(got[reached], want[reached])


def test_large_mesh_with_fixed_clusters(gpu_ready):
    """2^18 items and more take the speculative parallel growth; fixed
    clusters start at their items and the grown regions take the ids after
    them."""
    xyz, tri = meshgen.torus(640, 480, seed=6)  # 307,200 items
    K = 4000
    fixed = np.random.default_rng(3).choice(len(xyz), 40, replace=False).astype(np.int32)
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=K, metric=QEM, initial_sampling=0)
    ctx.set_fixed_clusters(0, fixed)
    ctx.initial_sampling()
    got = ctx.labels(0)
    po.set_fixed(fixed)
    try:
        want = po.RefMesh(xyz, tri).initial_sampling(QEM, K)
    finally:
        po.set_fixed(None)
    reached = want < K
    np.testing.assert_array_equal(got[reached], want[reached])
    np.testing.assert_array_equal(got[fixed], np.arange(len(fixed)))
    ctx.close()
