"""Fixed clusters on the GPU (acvd_set_fixed_clusters; the class API's
SetFixedClusters with GetCluster(i).AnchorItem; ACVDQ's -fv / -ft):
  - the library's initial sampling equals the reference restatement's
    (cluster k starts at its fixed item, Common/vtkUniformClustering.h:1186-1191);
  - labels equal the CPU replay of the same schedule bit for bit (QEM with
    and without unconstrained initialisation, isotropic, and a batch where
    only one mesh has fixed clusters), every fixed item ends in its cluster;
  - ACVDQ's own check (DiscreteRemeshing/Examples/ACVDQ.cxx:340-366): output
    vertex i equals input vertex fixed[i] coordinate for coordinate, with
    input coordinates that FP32 cannot hold;
  - final energy (anchored points) within 0.2 % of the serial reference
    restatement from the same initial labels (QEM, unconstrained
    initialisation, cube 100 x 100 per face, K = 2,500, 20 fixed items)."""
import numpy as np
import pytest

import acvd_amd
from acvd_amd import meshgen
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ISO, QEM = acvd_amd.METRIC_ISOTROPIC, acvd_amd.METRIC_QEM


@pytest.fixture(autouse=True)
def clear_oracle_fixed():
    yield
    po.set_fixed(None)


def pick_fixed(n_items, n, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n_items, n, replace=False)).astype(np.int32)


def run(xyz, tri, metric, K, unc, fixed, init=None, seed=5):
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=K, metric=metric, unconstrained_initialization=unc, seed=seed)
    ctx.set_fixed_clusters(0, fixed)
    ctx.set_initial_labels(0, init)
    ctx.process()
    return ctx


def test_sampling_with_fixed_clusters_equals_the_restatement(gpu_ready):
    xyz, tri = meshgen.icosphere(5, 3, 0.02)
    K = 400
    fixed = pick_fixed(len(xyz), 12, 1)
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=K, metric=QEM, unconstrained_initialization=1)
    ctx.set_fixed_clusters(0, fixed)
    ctx.initial_sampling()
    got = ctx.labels(0)
    po.set_fixed(fixed)
    want = po.RefMesh(xyz, tri).initial_sampling(QEM, K)
    assigned = want < K
    np.testing.assert_array_equal(got[assigned], want[assigned])
    np.testing.assert_array_equal(got[fixed], np.arange(len(fixed)))


@pytest.mark.parametrize("metric,unc", [(QEM, 1), (QEM, 0), (ISO, 0)])
def test_labels_equal_replay_with_fixed_clusters(gpu_ready, metric, unc):
    xyz, tri = meshgen.icosphere(5, 4, 0.03)
    K = 500
    fixed = pick_fixed(len(xyz), 16, 2)
    po.set_fixed(fixed)
    m = po.RefMesh(xyz, tri)
    init = m.initial_sampling(metric, K)
    print(f"fixed clusters: metric {metric}, unconstrained {unc}", flush=True)
    ctx = run(xyz, tri, metric, K, unc, fixed, init)
    lab = ctx.labels(0)
    rp, col = ctx.adjacency(0)
    xw, q = ctx.items(0)
    rlab, rst = po.replay_run(rp, col, xw, q, ctx.scales(0), K, metric, init, seed=5, unconstrained=unc)
    np.testing.assert_array_equal(lab, rlab)
    assert ctx.stats(0)["loops"] == rst["loops"]
    if metric == QEM:
        np.testing.assert_array_equal(lab[fixed], np.arange(len(fixed)))
        # the cluster point of a fixed cluster is its item's input position
        pts = ctx.cluster_points(0)
        np.testing.assert_array_equal(pts[:len(fixed)], np.asarray(xyz, np.float64)[fixed])
    assert m.clean(K, lab)[1] == 0


def test_batch_with_one_anchored_mesh(gpu_ready):
    specs = [(meshgen.icosphere(4, 6, 0.03), QEM, 170, 1, pick_fixed(2562, 6, 3)),
             (meshgen.star_icosphere(4, 7), QEM, 170, 1, None),
             (meshgen.icosphere(4, 8, 0.02), ISO, 170, 0, pick_fixed(2562, 4, 4))]
    ctx = acvd_amd.Context()
    inits = []
    for (xyz, tri), metric, K, unc, fixed in specs:
        mid = ctx.add_input(xyz, tri)
        ctx.set_params(mid, number_of_clusters=K, metric=metric, unconstrained_initialization=unc, seed=mid)
        ctx.set_fixed_clusters(mid, fixed)
        po.set_fixed(fixed)
        init = po.RefMesh(xyz, tri).initial_sampling(metric, K)
        ctx.set_initial_labels(mid, init)
        inits.append(init)
    ctx.process()
    for mid, ((xyz, tri), metric, K, unc, fixed) in enumerate(specs):
        po.set_fixed(fixed)
        rp, col = ctx.adjacency(mid)
        xw, q = ctx.items(mid)
        rlab, _ = po.replay_run(rp, col, xw, q, ctx.scales(mid), K, metric, inits[mid], seed=mid,
                                unconstrained=unc)
        np.testing.assert_array_equal(ctx.labels(mid), rlab)


def test_acvdq_anchored_vertex_check(gpu_ready):
    xyz, tri = meshgen.icosphere(5, 9, 0.02)
    xyz = np.asarray(xyz, np.float64) * (1.0 + 1e-9)  # not representable in FP32
    fixed = pick_fixed(len(xyz), 24, 5)
    r = acvd_amd.QIsotropicDiscreteRemeshing.New()
    r.SetInput(acvd_amd.Mesh(xyz, tri))
    r.SetFixedClusters(fixed)
    r.SetNumberOfClusters(300 + len(fixed))
    for i in range(len(fixed)):
        r.GetCluster(i).AnchorItem = int(fixed[i])
    r.SetUnconstrainedInitialization(1)
    r.SetSubsamplingThreshold(10)
    r.Remesh()
    out = r.GetOutput()
    for i, v in enumerate(fixed):
        assert tuple(out.GetPoint(i)) == tuple(xyz[v]), f"vertex {v} has been lost"
    r.Delete()


def test_energy_gap_with_fixed_clusters(gpu_ready):
    xyz, tri = meshgen.cube(100, 2, 1e-4)
    K = 2500
    fixed = pick_fixed(len(xyz), 20, 7)
    po.set_fixed(fixed)
    m = po.RefMesh(xyz, tri)
    init = m.fill_holes(K, m.initial_sampling(QEM, K))
    slab, _ = m.serial_run(QEM, K, unconstrained=1, init=init)
    es = m.energy(QEM, K, slab)[0]
    for seed in (0, 1):
        lab = run(xyz, tri, QEM, K, 1, fixed, init, seed).labels(0)
        eg = m.energy(QEM, K, lab)[0]
        print(f"fixed clusters, seed {seed}: gap {(eg - es) / es * 100:+.4f}%", flush=True)
        assert (eg - es) / es <= 2e-3
        np.testing.assert_array_equal(lab[fixed], np.arange(len(fixed)))
