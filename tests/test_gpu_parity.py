"""GPU parity through the C-ABI (libacvd_amd.so on the MI355X):
  - the device build (CSR, FP32 items, scales) equals the host restatement;
  - labels are bit-identical to the CPU replay of the same pass schedule;
  - final energy against the serial reference restatement (oracle A);
  - batching: each packed mesh equals its own replay;
  - determinism, edge cases, error paths, outputs (points, energy), and the
    drop-in class API.
Tolerances: integer/label work is compared exactly; FP64 energies and points
to the precision of FP32 items (rtol 1e-6) unless stated."""
import numpy as np
import pytest

import acvd_amd
from acvd_amd import meshgen
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ISO, QEM = acvd_amd.METRIC_ISOTROPIC, acvd_amd.METRIC_QEM


def centroid_split(xyz, tri):
    """Each triangle split 1-to-3 at its centroid: the original vertices get
    twice their valence (10 and 12 on an icosphere), so a third of the rows
    are longer than 8 neighbours."""
    xyz = np.asarray(xyz)
    tri = np.asarray(tri, np.int64)
    c = xyz[tri].mean(axis=1).astype(xyz.dtype)
    mid = len(xyz) + np.arange(len(tri))
    out = np.concatenate([np.stack([tri[:, 0], tri[:, 1], mid], 1), np.stack([tri[:, 1], tri[:, 2], mid], 1),
                          np.stack([tri[:, 2], tri[:, 0], mid], 1)])
    return np.concatenate([xyz, c]), out.astype(np.int32)


def small_cases():
    return {
        "ico5-iso": (meshgen.icosphere(5, 1, 0.02), ISO, 500, 0, None, 0.0),
        "ico5-qem": (meshgen.icosphere(5, 1, 0.02), QEM, 500, 1, None, 0.0),
        "cube40-qem": (meshgen.cube(40, 2, 1e-4), QEM, 400, 1, None, 0.0),
        "cube40-qem-constrained": (meshgen.cube(40, 2, 1e-4), QEM, 300, 0, None, 0.0),
        "C1-iso": (meshgen.icosphere(6, 1, 0.02), ISO, 2000, 0, None, 0.0),
        "torus-graded-qem": (meshgen.torus(128, 64, 3), QEM, 600, 1, "H", 1.5),
        "ico4-qem-level1": (meshgen.icosphere(4, 5, 0.05), QEM, 100, 1, None, 0.0),
        "split-ico4-qem": (centroid_split(*meshgen.icosphere(4, 9, 0.02)), QEM, 300, 1, None, 0.0),
        "split-ico4-qem-constrained": (centroid_split(*meshgen.icosphere(4, 9, 0.02)), QEM, 250, 0, None, 0.0),
    }


CASES = small_cases()


def run_gpu(xyz, tri, metric, K, unconstrained, init, ind=None, gradation=0.0, seed=7, level=3):
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=K, metric=metric, unconstrained_initialization=unconstrained,
                   max_number_of_loops=200000, seed=seed, gradation=gradation, quadrics_optimization_level=level)
    ctx.set_indicator(0, ind)
    ctx.set_initial_labels(0, init)
    ctx.process()
    return ctx


def case_inputs(name):
    (xyz, tri), metric, K, unc, ind, grad = CASES[name]
    if ind == "H":
        ind = meshgen.mean_curvature(xyz, tri)
    level = 1 if name.endswith("level1") else 3
    return xyz, tri, metric, K, unc, ind, grad, level


@pytest.mark.parametrize("name", list(CASES))
def test_device_build_equals_host_restatement(gpu_ready, name):
    xyz, tri, metric, K, unc, ind, grad, level = case_inputs(name)
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=K, metric=metric, gradation=grad)
    ctx.set_indicator(0, ind)
    ctx.build()
    rp, col = ctx.adjacency(0)
    xw, q = ctx.items(0)
    hrp, hcol, hxw, hq, hsc = po.device_inputs(np.asarray(xyz, np.float64), tri, metric, ind, grad)
    np.testing.assert_array_equal(rp, hrp)
    np.testing.assert_array_equal(col, hcol)
    np.testing.assert_array_equal(xw[:, :3], hxw[:, :3])
    if ind is None:
        np.testing.assert_array_equal(xw[:, 3], hxw[:, 3])
        np.testing.assert_array_equal(ctx.scales(0), hsc)
    else:  # device pow vs host pow: at most one FP32 ulp apart
        np.testing.assert_allclose(xw[:, 3], hxw[:, 3], rtol=2.5e-7)
    if metric == QEM:
        np.testing.assert_array_equal(q, hq)


@pytest.mark.parametrize("name", list(CASES))
def test_labels_bit_exact_against_replay(gpu_ready, name):
    xyz, tri, metric, K, unc, ind, grad, level = case_inputs(name)
    m = po.RefMesh(xyz, tri)
    init = m.initial_sampling(metric, K, grad, ind)
    ctx = run_gpu(xyz, tri, metric, K, unc, init, ind, grad, level=level)
    lab = ctx.labels(0)
    rp, col = ctx.adjacency(0)
    xw, q = ctx.items(0)
    rlab, rst = po.replay_run(rp, col, xw, q, ctx.scales(0), K, metric, init, level=level, seed=7,
                              unconstrained=unc, max_loops=200000)
    st = ctx.stats(0)
    print(name, st)
    assert st["loops"] == rst["loops"] and st["total_moves"] == rst["total_moves"]
    np.testing.assert_array_equal(lab, rlab)
    assert st["finished"] == 1
    assert lab.min() >= 0 and lab.max() < K
    # cluster sizes reported by the device match the labels
    np.testing.assert_array_equal(ctx.cluster_sizes(0), np.bincount(lab, minlength=K)[:K])


@pytest.mark.parametrize("name", ["C1-iso", "ico5-qem", "cube40-qem", "torus-graded-qem"])
def test_energy_against_serial_reference(gpu_ready, name):
    """Final unshifted FP64 energy of the GPU labels against oracle A run from
    the same initial labels.  Distinct local minima of meshes this small
    scatter by about 1 % (graded: the serial reference itself lands 5-45 %
    apart across initial labellings, and the GPU gap over twelve accept seeds
    spans 0.4-1.8 % on this 8,192-vertex torus), so the statistic is the median
    gap over five accept seeds, bounded by 0.5 % (1.5 % graded); the 0.1 %
    target is measured at the benchmark sizes."""
    xyz, tri, metric, K, unc, ind, grad, level = case_inputs(name)
    m = po.RefMesh(xyz, tri)
    init = m.initial_sampling(metric, K, grad, ind)
    slab, _ = m.serial_run(metric, K, grad, ind, 3, unc, init)
    es = m.energy(metric, K, slab, grad, ind)[0]
    gaps = []
    for seed in (7, 8, 9, 10, 11):
        ctx = run_gpu(xyz, tri, metric, K, unc, init, ind, grad, seed=seed)
        lab = ctx.labels(0)
        eg = m.energy(metric, K, lab, grad, ind)[0]
        gaps.append((eg - es) / es)
        # the device's own FP64 energy (FP32 items) agrees with the oracle's
        assert np.isclose(ctx.energy(0), eg, rtol=1e-5)
    gap = float(np.median(gaps))
    print(f"{name}: serial {es:.9e} gaps {np.round(np.array(gaps) * 100, 4)} % median {gap * 100:+.4f}%")
    assert gap < (1.5e-2 if grad > 0 else 5e-3)


@pytest.mark.parametrize("metric,unc", [(ISO, 0), (QEM, 1)])
def test_energy_gap_at_scale(gpu_ready, metric, unc):
    # larger meshes: the GPU schedule should match or beat the serial minimum
    if metric == ISO:
        xyz, tri = meshgen.icosphere(7, 1, 0.02)
        K = 8000
    else:
        xyz, tri = meshgen.cube(100, 2, 1e-4)
        K = 2500
    m = po.RefMesh(xyz, tri)
    init = m.initial_sampling(metric, K)
    lab = run_gpu(xyz, tri, metric, K, unc, init).labels(0)
    slab, _ = m.serial_run(metric, K, unconstrained=unc, init=init)
    eg, es = m.energy(metric, K, lab)[0], m.energy(metric, K, slab)[0]
    print(f"gap {(eg - es) / es * 100:+.4f}%")
    assert (eg - es) / es < 3e-3


def test_determinism(gpu_ready):
    (xyz, tri), metric, K, unc, _, _ = CASES["cube40-qem"]
    init = po.RefMesh(xyz, tri).initial_sampling(metric, K)
    a = run_gpu(xyz, tri, metric, K, unc, init).labels(0)
    b = run_gpu(xyz, tri, metric, K, unc, init).labels(0)
    np.testing.assert_array_equal(a, b)


def test_batch_equals_each_mesh_alone(gpu_ready):
    specs = [(meshgen.icosphere(4, 11, 0.03), ISO, 170, 0, 1),
             (meshgen.star_icosphere(4, 12), QEM, 170, 1, 2),
             (meshgen.icosphere(5, 13, 0.01), QEM, 680, 1, 3),
             (meshgen.cube(20, 14, 1e-4), ISO, 160, 0, 4),
             (meshgen.star_icosphere(5, 15), ISO, 680, 0, 5)]
    ctx = acvd_amd.Context()
    inits = []
    for (xyz, tri), metric, K, unc, seed in specs:
        mid = ctx.add_input(xyz, tri)
        ctx.set_params(mid, number_of_clusters=K, metric=metric, unconstrained_initialization=unc, seed=seed)
        init = po.RefMesh(xyz, tri).initial_sampling(metric, K)
        ctx.set_initial_labels(mid, init)
        inits.append(init)
    ctx.process()
    for mid, ((xyz, tri), metric, K, unc, seed) in enumerate(specs):
        rp, col = ctx.adjacency(mid)
        xw, q = ctx.items(mid)
        rlab, _ = po.replay_run(rp, col, xw, q, ctx.scales(mid), K, metric, inits[mid], seed=seed,
                                unconstrained=unc)
        np.testing.assert_array_equal(ctx.labels(mid), rlab)
        assert ctx.stats(mid)["finished"] == 1


@pytest.mark.parametrize("sampling", [0, 1])
def test_initial_sampling_without_initial_labels(gpu_ready, sampling):
    """0: the reference's region growing on the host; 1: weighted seeds grown
    on the device.  Either way every cluster ends non-empty and connected."""
    xyz, tri = meshgen.icosphere(6, 1, 0.02)
    K = 2000
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=K, seed=3, initial_sampling=sampling)
    if sampling == 0:
        ctx.initial_sampling()
        want = po.RefMesh(xyz, tri).initial_sampling(ISO, K)
        got = ctx.labels(0)
        assigned = want < K
        np.testing.assert_array_equal(got[assigned], want[assigned])  # the fill only touches unassigned items
    ctx.process()
    lab = ctx.labels(0)
    assert lab.min() >= 0 and lab.max() < K
    assert np.all(np.bincount(lab, minlength=K) > 0)
    m = po.RefMesh(xyz, tri)
    assert m.clean(K, lab)[1] == 0
    ctx.process()  # deterministic for a fixed seed
    np.testing.assert_array_equal(ctx.labels(0), lab)


def test_edge_cases(gpu_ready):
    xyz, tri = meshgen.icosphere(3)
    n = len(xyz)
    # K = 1: one cluster, nothing to move
    ctx = run_gpu(xyz, tri, ISO, 1, 0, None)
    assert np.all(ctx.labels(0) == 0)
    # K = N: every vertex is its own cluster (size-1 clusters never move)
    init = np.arange(n, dtype=np.int32)
    ctx = run_gpu(xyz, tri, ISO, n, 0, init)
    np.testing.assert_array_equal(ctx.labels(0), init)
    # two disconnected components in one mesh
    x2 = np.concatenate([xyz, xyz + [3.0, 0, 0]])
    t2 = np.concatenate([tri, tri + n])
    ctx = run_gpu(x2, t2, QEM, 60, 1, None)
    lab = ctx.labels(0)
    assert lab.max() < 60 and np.all(np.bincount(lab, minlength=60) > 0)
    # an all-null initial labelling of one component stays null (no seed there)
    init = np.full(2 * n, 60, np.int32)
    init[:60] = np.arange(60)
    ctx = run_gpu(x2, t2, ISO, 60, 0, init)
    lab = ctx.labels(0)
    assert np.all(lab[n:] == 60) and np.all(lab[:n] < 60)


def test_error_paths(gpu_ready):
    xyz, tri = meshgen.icosphere(2)
    ctx = acvd_amd.Context()
    with pytest.raises(acvd_amd.AcvdError, match="references vertex"):
        ctx.set_input(xyz, np.array([[0, 1, 9999]], np.int32))
    ctx.set_input(xyz, tri)
    with pytest.raises(acvd_amd.AcvdError, match="more than zero"):
        ctx.process()
    ctx.set_params(0, number_of_clusters=len(xyz) + 1)
    with pytest.raises(acvd_amd.AcvdError, match="more clusters than vertices"):
        ctx.process()
    with pytest.raises(acvd_amd.AcvdError, match="level"):
        ctx.set_params(0, number_of_clusters=10, quadrics_optimization_level=7)
    with pytest.raises(acvd_amd.AcvdError, match="out of range"):
        ctx.labels(5)


def test_cluster_points(gpu_ready):
    for name in ("ico5-iso", "cube40-qem"):
        xyz, tri, metric, K, unc, ind, grad, level = case_inputs(name)
        m = po.RefMesh(xyz, tri)
        init = m.initial_sampling(metric, K)
        ctx = run_gpu(xyz, tri, metric, K, unc, init)
        lab = ctx.labels(0)
        pts = ctx.cluster_points(0)
        xw, q = ctx.items(0)
        w = xw[:, 3].astype(np.float64)
        W = np.bincount(lab, weights=w, minlength=K)
        cen = np.stack([np.bincount(lab, weights=w * xw[:, k], minlength=K) for k in range(3)], 1) / W[:, None]
        if metric == ISO:
            np.testing.assert_allclose(pts, cen, rtol=1e-9, atol=1e-12)
        else:
            Qc = np.zeros((K, 9))
            np.add.at(Qc, lab, q[:, :9].astype(np.float64))
            for c in range(K):
                p, _ = po.rep_point(Qc[c], cen[c], 3)
                np.testing.assert_allclose(pts[c], p, atol=2e-5)


def test_dropin_class_api(gpu_ready):
    xyz, tri = meshgen.icosphere(5, 1, 0.02)
    m = po.RefMesh(xyz, tri)
    init = m.initial_sampling(QEM, 400)
    r = acvd_amd.QIsotropicDiscreteRemeshing.New()
    r.SetInput(acvd_amd.Mesh(xyz, tri))
    r.SetConsoleOutput(0)
    r.SetNumberOfClusters(400)
    r.GetMetric().SetGradation(0.0)
    r.SetUnconstrainedInitialization(1)
    r.SetInitialClustering(init)
    r.SetSeed(7)
    lab = r.ProcessClustering()
    assert lab is r.GetClustering()
    ref = run_gpu(xyz, tri, QEM, 400, 1, init).labels(0)
    np.testing.assert_array_equal(lab, ref)
    assert sum(r.GetClusterSize(c) for c in range(400)) == len(xyz)
    r.Delete()
