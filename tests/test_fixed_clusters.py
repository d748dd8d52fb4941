"""Fixed clusters (SetFixedClusters with the QEM metric's AnchorItem, as
ACVDQ's -fv / -ft options set them, DiscreteRemeshing/Examples/ACVDQ.cxx:
327-366) in the CPU oracles: the reference's initial sampling starts cluster
k at its fixed item (Common/vtkUniformClustering.h:1186-1191), the clean keeps
the component holding the fixed item (:436-447), and with the QEM metric the
item never leaves its cluster, whose point is the item's position
(DiscreteRemeshing/vtkQEMetricForClustering.h:196-199, 259, 270-274).
Checked: the sampling and clean rules on crafted labellings; runs of the
serial restatement and of the replay end with every fixed item in its
cluster, at a local minimum of the anchored energy (tests/certify.py), and
the anchored cluster's energy is the closed form W|a|^2 - 2 a.S."""
import numpy as np
import pytest

from acvd_amd import meshgen
from oracle import pyoracle as po
from tests import certify

ISO, QEM = 0, 1


@pytest.fixture
def fixed_ids():
    yield np.array([1711, 40, 2301, 977, 5, 1200], np.int32)
    po.set_fixed(None)


def test_sampling_starts_fixed_clusters_at_their_items(fixed_ids):
    xyz, tri = meshgen.icosphere(4, 5, 0.05)
    K = 100
    m = po.RefMesh(xyz, tri)
    free = m.initial_sampling(QEM, K)
    po.set_fixed(fixed_ids)
    lab = m.initial_sampling(QEM, K)
    n = len(fixed_ids)
    np.testing.assert_array_equal(lab[fixed_ids], np.arange(n))
    # the grown regions take the ids n..K-1, each item once
    grown = np.flatnonzero(~np.isin(np.arange(len(xyz)), fixed_ids) & (lab < K))
    assert lab[grown].min() >= n
    assert set(np.unique(lab[lab < K])) == set(range(K))
    assert not np.array_equal(lab, free)


def test_clean_keeps_the_component_of_the_fixed_item(fixed_ids):
    xyz, tri = meshgen.icosphere(3, 5, 0.0)
    m = po.RefMesh(xyz, tri)
    a, b = m.edges()
    nb = certify.neighbour_sets(len(xyz), a, b)
    # cluster 0: a large blob and, apart from it, a single item holding the
    # fixed item; everything else cluster 1 (item 0 stays in cluster 1: the
    # reference never records a second component of a cluster whose first
    # one starts at item 0, vtkUniformClustering.h:463-466)
    c = 100
    lab = np.ones(len(xyz), np.int32)
    blob = [c] + sorted(nb[c])
    assert 0 not in blob
    lab[blob] = 0
    far = int(np.argmax(np.linalg.norm(xyz - xyz[c], axis=1)))
    lab[far] = 0
    free, nd = m.clean(2, lab)
    assert nd == 1 and free[far] == 2 and np.all(free[blob] == 0)  # largest component kept
    po.set_fixed(np.array([far], np.int32))
    kept, nd = m.clean(2, lab)
    assert nd == 1 and kept[far] == 0 and np.all(kept[blob] == 2)  # the fixed item's component kept


@pytest.mark.parametrize("unc", [0, 1])
def test_runs_keep_anchors_and_end_at_an_anchored_local_minimum(fixed_ids, unc):
    xyz, tri = meshgen.icosphere(4, 5, 0.05)
    K = 100
    m = po.RefMesh(xyz, tri)
    po.set_fixed(fixed_ids)
    init = m.initial_sampling(QEM, K)
    lab, st, _, L = m.serial_trace(QEM, K, unconstrained=unc, init=init)
    np.testing.assert_array_equal(lab[fixed_ids], np.arange(len(fixed_ids)))
    w, q9 = m.items(QEM)
    a, b = m.edges()
    bad = certify.violations(QEM, xyz, a, b, lab, K, w, q9, 3, connexity=bool(L["connexity"][-1]),
                             constrained=bool(L["constrained"][-1]), fixed=fixed_ids)
    assert not bad, bad[:5]
    # the replay of the device schedule on the same input
    rp, col, xw, q, sc = po.device_inputs(np.asarray(xyz, np.float64), tri, QEM)
    rlab, rst = po.replay_run(rp, col, xw, q, sc, K, QEM, init, seed=3, unconstrained=unc)
    np.testing.assert_array_equal(rlab[fixed_ids], np.arange(len(fixed_ids)))
    bad = certify.violations(QEM, xw[:, :3].astype(np.float64), a, b, rlab, K, xw[:, 3].astype(np.float64),
                             q[:, :9].astype(np.float64), 3, connexity=True, fixed=fixed_ids,
                             accept_eps=2.0 ** -18 + 1e-6)
    assert not bad, bad[:5]
    assert m.clean(K, rlab)[1] == 0


def test_anchored_energy_is_the_closed_form(fixed_ids):
    xyz, tri = meshgen.icosphere(3, 5, 0.05)
    K = 30
    m = po.RefMesh(xyz, tri)
    lab = m.fill_holes(K, meshgen.seed_labels(len(xyz), K, 2))
    w, _ = m.items(QEM)
    ids = np.array([int(np.flatnonzero(lab == k)[0]) for k in range(3)], np.int32)
    po.set_fixed(ids)
    e_fixed = m.energy(QEM, K, lab)[0]
    po.set_fixed(None)
    e_free = m.energy(QEM, K, lab)[0]
    # the unshifted energy of cluster k changes from sum w|x - q|^2 to sum w|x - a|^2
    x = np.asarray(xyz, np.float64)
    want = e_free
    for k in range(3):
        mem = lab == k
        pts = m.energy(QEM, K, np.where(mem, lab, K))  # cluster k alone, free point
        want += np.sum(w[mem] * np.sum((x[mem] - x[ids[k]]) ** 2, 1)) - pts[0]
    np.testing.assert_allclose(e_fixed, want, rtol=1e-12)
