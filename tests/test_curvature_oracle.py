"""The curvature-indicator oracle (oracle/curvature.h, restating
Common/vtkCurvatureMeasure.cxx:188-620 and
Common/vtkNeighbourhoodComputation.cxx:38-110), pinned by closed forms and by
an independent numpy fit of the same face data."""
import numpy as np
import pytest

from acvd_amd import meshgen
from oracle import pyoracle as po


def grid(nu, nv, f):
    u, v = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    P = f(u.ravel().astype(np.float64), v.ravel().astype(np.float64))
    T = []
    for i in range(nu - 1):
        for j in range(nv - 1):
            a, b, c, d = i * nv + j, (i + 1) * nv + j, (i + 1) * nv + j + 1, i * nv + j + 1
            T += [[a, b, c], [a, c, d]]
    return P, np.array(T, np.int32)


def cylinder(r, n_around=96, n_len=60, dz=0.02):
    P, T = grid(n_len, n_around + 1, lambda u, v: np.stack(
        [u * dz, r * np.cos(2 * np.pi * v / n_around), r * np.sin(2 * np.pi * v / n_around)], 1))
    remap = np.arange(len(P))
    for i in range(n_len):  # close the seam
        remap[i * (n_around + 1) + n_around] = i * (n_around + 1)
    return P, remap[T].astype(np.int32), [i * (n_around + 1) + j for i in range(10, n_len - 10)
                                         for j in range(n_around)]


def test_face_list_is_the_distance_two_star():
    # interior vertex of a regular valence-6 grid: every face touching a vertex
    # within two edges, i.e. the hexagon of radius 3: 6 + 18 + 30 faces
    P, T = grid(30, 30, lambda u, v: np.stack([u * 0.1 + 0.05 * v, v * 0.1, 0 * u], 1))
    m = po.RefMesh(P, T)
    v = 15 * 30 + 15
    fl = m.curvature_face_list(v)
    assert len(fl) == 54 and len(set(fl.tolist())) == 54
    # brute force: faces with a vertex at graph distance <= 2
    adj = {i: set() for i in range(len(P))}
    for a, b, c in T:
        adj[a] |= {b, c}; adj[b] |= {a, c}; adj[c] |= {a, b}
    near = {v} | adj[v] | set().union(*(adj[u] for u in adj[v]))
    expect = {f for f, t in enumerate(T) if near & set(t.tolist())}
    assert set(fl.tolist()) == expect
    # the faces around the vertex come first (GetVertexNeighbourFaces)
    assert set(fl[:6].tolist()) == {f for f, t in enumerate(T) if v in t}


def test_plane_is_flat():
    P, T = grid(30, 30, lambda u, v: np.stack([u * 0.1, v * 0.1 + 0.03 * u, 0 * u], 1))
    k = po.RefMesh(P, T).curvature_indicator()
    assert np.abs(k).max() < 1e-9


@pytest.mark.parametrize("R", [1.0, 2.0, 0.25])
def test_sphere_tends_to_sqrt2_over_R(R):
    xyz, tri = meshgen.icosphere(5, 1, 0.0)
    xyz = np.asarray(xyz, np.float64)
    xyz = R * xyz / np.linalg.norm(xyz, axis=1)[:, None]
    k = po.RefMesh(xyz, tri).curvature_indicator()
    np.testing.assert_allclose(k, np.sqrt(2) / R, rtol=0.01)


def test_cylinder_follows_the_references_mixed_term():
    # z = (kappa/2)(x cos t + y sin t)^2 in the fitted frame gives, with the
    # reference's f = 2 a4 / sqrt(..) (vtkCurvatureMeasure.cxx:344-347), the
    # indicator kappa * sqrt(cos^4 t + sin^4 t + 8 cos^2 t sin^2 t), which lies
    # in [kappa, kappa * sqrt(2.5)] (the textbook f would give kappa for all t)
    r = 0.5
    P, T, inner = cylinder(r)
    k = po.RefMesh(P, T).curvature_indicator()[inner]
    kappa = 1.0 / r
    assert k.min() > 0.99 * kappa and k.max() < 1.01 * kappa * np.sqrt(2.5)
    assert k.min() < 1.03 * kappa and k.max() > 0.97 * kappa * np.sqrt(2.5)


def numpy_indicator(xyz, tri, faces):
    """The fit of vtkCurvatureMeasure.cxx:188-509 written independently:
    numpy lstsq (SVD) instead of the normal equations, eigvalsh for the
    upper-symmetrised shape operator."""
    p = xyz[tri[faces]]  # (n, 3, 3)
    cr = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    area = 0.5 * np.linalg.norm(cr, axis=1)
    nrm = cr / np.linalg.norm(cr, axis=1)[:, None]
    bary = p.mean(axis=1)
    origin = (area[:, None] * bary).sum(0) / area.sum()
    n0 = (area[:, None] * nrm).sum(0)
    n0 /= np.linalg.norm(n0)
    f1 = np.array([n0[2], n0[0], n0[1]])
    t0 = np.cross(f1, n0)
    t0 /= np.linalg.norm(t0)
    t1 = np.cross(n0, t0)
    d = bary - origin
    x, y, z = d @ t0, d @ t1, d @ n0
    X = np.stack([np.ones_like(x), x, y, x * x, x * y, y * y], 1)
    q = np.linalg.lstsq(X, z, rcond=None)[0]
    s = np.sqrt(1 + q[1] ** 2 + q[2] ** 2)
    I = np.array([[1 + q[1] ** 2, q[1] * q[2]], [q[1] * q[2], 1 + q[2] ** 2]])
    II = np.array([[2 * q[3], 2 * q[4]], [2 * q[4], 2 * q[5]]]) / s
    A = -(II @ np.linalg.inv(I)).T  # A[r][c] as the reference lays it out
    S = np.array([[A[0, 0], A[0, 1]], [A[0, 1], A[1, 1]]])
    return np.sqrt((np.linalg.eigvalsh(S) ** 2).sum())


def test_fit_equals_an_independent_least_squares():
    rng = np.random.default_rng(3)
    xyz, tri = meshgen.icosphere(4, 2, 0.05)
    xyz = np.asarray(xyz, np.float64) * (1 + 0.3 * np.asarray(xyz, np.float64)[:, :1])  # anisotropic
    m = po.RefMesh(xyz, tri)
    k = m.curvature_indicator()
    for v in rng.choice(len(xyz), 40, replace=False):
        ref = numpy_indicator(xyz, np.asarray(tri), m.curvature_face_list(v))
        np.testing.assert_allclose(k[v], ref, rtol=1e-8)


def test_small_neighbourhood_gives_zero():
    # a lone triangle pair: at most 2 faces around any vertex
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.3]], np.float64)
    T = np.array([[0, 1, 2], [1, 3, 2]], np.int32)
    assert np.all(po.RefMesh(P, T).curvature_indicator() == 0.0)
