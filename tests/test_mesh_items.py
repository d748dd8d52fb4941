"""Mesh topology and item build: the reference-ordered edge table
(vtkSurfaceBase.cxx:1166-1221, 1407-1468), vertex areas (vtkSurface.cxx:1342-1359),
weights with clamp (vtkIsotropicMetricForClustering.h:249-291,
vtkQEMetricForClustering.h:324-362), and the host build of the device inputs
(CSR, FP32 items, fixed-point scales) against the FP64 oracle items."""
import numpy as np
import pytest

from acvd_amd import meshgen
from oracle import pyoracle as po


def test_generator_sizes_and_closed_topology():
    for L in range(0, 5):
        xyz, tri = meshgen.icosphere(L)
        assert len(xyz) == 10 * 4 ** L + 2 and len(tri) == 20 * 4 ** L
        np.testing.assert_allclose(np.linalg.norm(xyz, axis=1), 1.0, rtol=1e-14)
    xyz, tri = meshgen.cube(20)
    assert len(xyz) == 6 * 21 ** 2 - 12 * 21 + 8 and len(tri) == 12 * 400
    xyz, tri = meshgen.torus(64, 32)
    assert len(xyz) == 64 * 32 and len(tri) == 2 * 64 * 32
    for (xyz, tri, chi) in [(*meshgen.icosphere(3), 2), (*meshgen.cube(10), 2), (*meshgen.torus(40, 20), 0)]:
        m = po.RefMesh(xyz, tri)
        ne = len(m.edges()[0])
        assert len(xyz) - ne + len(tri) == chi  # Euler characteristic of a closed surface


def test_cube_geometry_is_exact_on_features():
    xyz, tri = meshgen.cube(50, seed=2, jitter=1e-4)
    ext = np.abs(xyz) == 1.0
    assert np.all(ext.any(axis=1))  # every vertex lies on a face
    corners = ext.sum(axis=1) == 3
    assert corners.sum() == 8
    edges = ext.sum(axis=1) == 2
    assert edges.sum() == 12 * 49


def test_edge_table_follows_face_order():
    xyz, tri = meshgen.icosphere(1)
    m = po.RefMesh(xyz, tri)
    a, b = m.edges()
    # edges appear with the direction and order of their first face occurrence
    seen = {}
    order = []
    for f in tri:
        for j in range(3):
            v1, v2 = int(f[j]), int(f[(j + 1) % 3])
            key = (min(v1, v2), max(v1, v2))
            if key not in seen:
                seen[key] = (v1, v2)
                order.append((v1, v2))
    assert list(zip(a.tolist(), b.tolist())) == order


def test_vertex_areas_sum_to_surface_area():
    xyz, tri = meshgen.cube(30, seed=2, jitter=0.0)
    m = po.RefMesh(xyz, tri)
    assert np.isclose(m.vertex_areas().sum(), 24.0, rtol=1e-12)
    xyz, tri = meshgen.icosphere(5)
    a = po.RefMesh(xyz, tri).vertex_areas().sum()
    assert abs(a - 4 * np.pi) / (4 * np.pi) < 2e-3


def test_weight_clamp_bounds():
    # clamp to [avg/ratio, avg*ratio] (ratio 1e5 isotropic, 1e4 QEM); the mesh
    # has more vertices than either ratio so one outlier can hit the top bound
    xyz, tri = meshgen.icosphere(7)
    ind = np.ones(len(xyz))
    ind[0] = 1e8    # large density on one vertex (above avg * ratio for both ratios)
    ind[1] = 1e-12  # tiny density on another
    m = po.RefMesh(xyz, tri)
    area = m.vertex_areas()
    raw = area * ind
    avg = raw.mean()
    for metric, ratio in ((po.ISO, 1e5), (po.QEM, 1e4)):
        w, _ = m.items(metric, ind, 1.0)
        assert np.isclose(w[0], avg * ratio, rtol=1e-12)
        assert np.isclose(w[1], avg / ratio, rtol=1e-12)
        np.testing.assert_allclose(w[2:], raw[2:], rtol=1e-12)


@pytest.mark.parametrize("metric", [po.ISO, po.QEM])
def test_device_inputs_match_oracle_items(metric):
    xyz, tri = meshgen.icosphere(4, seed=1, sigma=0.02)
    m = po.RefMesh(xyz, tri)
    rp, col, xw, q, sc = po.device_inputs(xyz, tri, metric)
    n = len(xyz)
    # CSR: sorted unique rows, symmetric, 2E entries
    ne = len(m.edges()[0])
    assert rp[-1] == 2 * ne
    for i in range(n):
        row = col[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0) and i not in row
    pairs = set()
    for i in range(n):
        for j in col[rp[i]:rp[i + 1]]:
            pairs.add((i, int(j)))
    assert all((j, i) in pairs for (i, j) in pairs)
    w, q9 = m.items(metric)
    np.testing.assert_allclose(xw[:, 3], w, rtol=2e-7)
    np.testing.assert_allclose(xw[:, :3], xyz, rtol=1e-7, atol=1e-7)
    if metric == po.QEM:
        np.testing.assert_allclose(q[:, :9], q9, rtol=1e-5, atol=1e-7 * np.abs(q9).max())
    # scales: powers of two with sums below 2^60
    for s in sc:
        assert s > 0 and np.log2(s) == np.round(np.log2(s))
    assert xw[:, 3].astype(np.float64).sum() * sc[0] < 2.0 ** 60
    assert (xw[:, 3].astype(np.float64) * np.abs(xw[:, :3]).max(axis=1)).sum() * sc[1] < 2.0 ** 60


def test_mean_curvature_of_a_sphere():
    xyz, tri = meshgen.icosphere(5)
    H = meshgen.mean_curvature(xyz, tri)
    # unit sphere: |H| = 1; the 12 valence-5 vertices keep the O(1) bias of the
    # barycentric area (about +15 %), every other vertex converges
    assert abs(np.median(H) - 1.0) < 1e-3
    assert np.sum(np.abs(H - 1.0) > 0.01) <= 12
    assert np.all(np.abs(H - 1.0) < 0.2)


def test_mt19937_known_answer():
    # the reference's initial sampling shuffles with std::mt19937
    # (vtkUniformClustering.h:1202-1207); the C++ standard fixes this value
    assert po.mt19937_10000() == 4123659995
