"""Pins the quadric tools (Common/vtkQuadricTools.cxx:56-177) of the oracle and
of the shared device header by closed forms and known answers:
cube known-answer tests (inside a face / across an edge / at a corner), level-1
projection, Q = 0, full-rank solve, plane invariants of the triangle quadric,
and the eigen-decomposition against numpy."""
import numpy as np
import pytest

from oracle import pyoracle as po


def tri_q(a, b, c):
    return po.triangle_quadric(np.array(a, float), np.array(b, float), np.array(c, float))


def face_quadric(tris):
    Q = np.zeros(10)
    for t in tris:
        Q += tri_q(*t)
    return Q


def grid_tris(origin, u, v, n, h):
    """triangles of an n x n grid of step h spanned by u, v at origin"""
    o, u, v = (np.array(x, float) for x in (origin, u, v))
    out = []
    for j in range(n):
        for i in range(n):
            p00 = o + h * (i * u + j * v)
            p10 = p00 + h * u
            p01 = p00 + h * v
            p11 = p00 + h * (u + v)
            out += [(p00, p10, p11), (p00, p11, p01)]
    return out


def test_triangle_quadric_plane_invariants():
    rng = np.random.default_rng(1)
    for _ in range(50):
        x = rng.standard_normal((3, 3))
        Q = po.triangle_quadric(*x)
        A = np.array([[Q[0], Q[1], Q[2], Q[3]], [Q[1], Q[4], Q[5], Q[6]],
                      [Q[2], Q[5], Q[7], Q[8]], [Q[3], Q[6], Q[8], Q[9]]])
        n = np.cross(x[1] - x[0], x[2] - x[0])
        area = 0.5 * np.linalg.norm(n)
        # rank one, trace of the 3x3 block = (2 area)^2
        assert np.linalg.matrix_rank(A, tol=1e-9 * np.abs(A).max()) == 1
        assert np.isclose(Q[0] + Q[4] + Q[7], (2 * area) ** 2, rtol=1e-12)
        # any point of the plane has zero error
        for t in rng.uniform(-2, 2, (5, 2)):
            p = x[0] + t[0] * (x[1] - x[0]) + t[1] * (x[2] - x[0])
            ph = np.append(p, 1.0)
            assert abs(ph @ A @ ph) <= 1e-10 * np.abs(A).max() * (1 + p @ p)
        # evaluation form of vtkQuadricTools::Evaluate: Q(p) = (n.p + d)^2
        p = rng.standard_normal(3)
        d = -n @ x[0]
        ph = np.append(p, 1.0)
        assert np.isclose(ph @ A @ ph, (n @ p + d) ** 2, rtol=1e-9, atol=1e-12)


def test_kat_cluster_inside_one_face():
    # a cluster of triangles on the face z = 1: rank-1 quadric, q = centroid
    tris = grid_tris((0.1, 0.2, 1.0), (1, 0, 0), (0, 1, 0), 4, 0.05)
    Q = face_quadric(tris)
    c = np.array([0.2, 0.31, 1.0])
    p, rd = po.rep_point(Q[:9], c, 3)
    assert rd == 2
    np.testing.assert_allclose(p, c, atol=1e-14)
    D, rd2 = po.hdr_rep_point_delta(Q[:9], c, 3)
    assert rd2 == 2
    np.testing.assert_allclose(c + D, c, atol=1e-14)


def test_kat_cluster_across_an_edge():
    # triangles on x = 1 and y = 1 around the cube edge x = y = 1:
    # rank deficiency 1 and q = projection of the centroid onto the edge line
    tris = grid_tris((1.0, 0.8, 0.3), (0, 1, 0), (0, 0, 1), 4, 0.05)  # face x = 1, y in [0.8,1]
    tris += grid_tris((0.8, 1.0, 0.3), (1, 0, 0), (0, 0, 1), 4, 0.05)  # face y = 1
    Q = face_quadric(tris)
    c = np.array([0.93, 0.91, 0.37])
    p, rd = po.rep_point(Q[:9], c, 3)
    assert rd == 1
    np.testing.assert_allclose(p, [1.0, 1.0, 0.37], atol=1e-12)
    D, rd2 = po.hdr_rep_point_delta(Q[:9], c, 3)
    assert rd2 == 1
    np.testing.assert_allclose(c + D, [1.0, 1.0, 0.37], atol=1e-12)


def test_kat_cluster_at_a_corner():
    tris = grid_tris((1.0, 0.8, 0.8), (0, 1, 0), (0, 0, 1), 4, 0.05)
    tris += grid_tris((0.8, 1.0, 0.8), (1, 0, 0), (0, 0, 1), 4, 0.05)
    tris += grid_tris((0.8, 0.8, 1.0), (1, 0, 0), (0, 1, 0), 4, 0.05)
    Q = face_quadric(tris)
    c = np.array([0.9, 0.93, 0.91])
    p, rd = po.rep_point(Q[:9], c, 3)
    assert rd == 0
    np.testing.assert_allclose(p, [1.0, 1.0, 1.0], atol=1e-12)
    D, rd2 = po.hdr_rep_point_delta(Q[:9], c, 3)
    assert rd2 == 0
    np.testing.assert_allclose(c + D, [1.0, 1.0, 1.0], atol=1e-12)


def test_level_one_projects_on_the_dominant_plane():
    # a corner cluster whose x = 1 face dominates: level 1 only moves along x
    tris = grid_tris((1.0, 0.6, 0.6), (0, 1, 0), (0, 0, 1), 8, 0.05)
    tris += grid_tris((0.9, 1.0, 0.9), (1, 0, 0), (0, 0, 1), 2, 0.05)
    Q = face_quadric(tris)
    c = np.array([0.95, 0.97, 0.93])
    p, rd = po.rep_point(Q[:9], c, 1)
    assert rd == 2
    assert abs(p[0] - 1.0) < 0.02 and abs(p[1] - c[1]) < 0.02 and abs(p[2] - c[2]) < 1e-12


def test_zero_quadric_keeps_the_point():
    c = np.array([0.3, -0.2, 0.7])
    p, rd = po.rep_point(np.zeros(9), c, 3)
    assert rd == 3
    np.testing.assert_array_equal(p, c)
    D, rd2 = po.hdr_rep_point_delta(np.zeros(9), c, 3)
    assert rd2 == 3 and np.all(D == 0)


def test_full_rank_solve_matches_linear_algebra():
    rng = np.random.default_rng(2)
    for _ in range(200):
        M = rng.standard_normal((3, 3))
        A = M @ M.T + 0.5 * np.eye(3)
        b = rng.standard_normal(3)
        Q = np.array([A[0, 0], A[0, 1], A[0, 2], -b[0], A[1, 1], A[1, 2], -b[1], A[2, 2], -b[2]])
        c = rng.standard_normal(3)
        x = np.linalg.solve(A, b)
        if np.linalg.eigvalsh(A).min() / np.linalg.eigvalsh(A).max() <= 1.1e-3:
            continue
        p, rd = po.rep_point(Q, c, 3)
        assert rd == 0
        np.testing.assert_allclose(p, x, rtol=1e-9, atol=1e-11)
        D, _ = po.hdr_rep_point_delta(Q, c, 3)
        np.testing.assert_allclose(c + D, x, rtol=1e-9, atol=1e-11)


def test_truncation_threshold():
    # eigenvalues 1, 2e-3, 5e-4: the last one is dropped at threshold 1e-3
    R = np.linalg.qr(np.random.default_rng(3).standard_normal((3, 3)))[0]
    lam = np.array([1.0, 2e-3, 5e-4])
    A = R @ np.diag(lam) @ R.T
    b = np.array([0.3, -0.1, 0.2])
    Q = np.array([A[0, 0], A[0, 1], A[0, 2], -b[0], A[1, 1], A[1, 2], -b[1], A[2, 2], -b[2]])
    c = np.zeros(3)
    p, rd = po.rep_point(Q, c, 3)
    assert rd == 1
    expect = sum(R[:, k] * (R[:, k] @ b) / lam[k] for k in range(2))
    np.testing.assert_allclose(p, expect, rtol=1e-9)
    D, rd2 = po.hdr_rep_point_delta(Q, c, 3)
    assert rd2 == 1
    np.testing.assert_allclose(D, expect, rtol=1e-9)


def test_eigen_decomposition_matches_numpy():
    rng = np.random.default_rng(4)
    for _ in range(100):
        M = rng.standard_normal((3, 3))
        A = (M + M.T) * 10.0 ** rng.uniform(-12, 3)
        lam, V = po.sym_eigen(A)
        np.testing.assert_allclose(np.sort(lam), np.linalg.eigvalsh(A), rtol=1e-10, atol=1e-13 * np.abs(A).max())
        np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-12)
        np.testing.assert_allclose(A @ V, V * lam, atol=1e-10 * np.abs(A).max())


def test_header_solve_agrees_with_oracle_on_random_quadrics():
    """Independent implementations (oracle: full Jacobi to convergence, device
    header: unrolled Jacobi) agree on sums of plane quadrics of all scales."""
    rng = np.random.default_rng(5)
    for _ in range(500):
        Q = np.zeros(9)
        for _ in range(rng.integers(1, 6)):
            Q += po.triangle_quadric(*rng.standard_normal((3, 3)))[:9]
        Q *= 10.0 ** rng.uniform(-14, 2)
        c = rng.standard_normal(3)
        for level in (1, 2, 3):
            p, rd = po.rep_point(Q, c, level)
            D, rd2 = po.hdr_rep_point_delta(Q, c, level)
            assert rd == rd2
            np.testing.assert_allclose(c + D, p, rtol=1e-7, atol=1e-9 * (1 + np.abs(p).max()))


def _random_rotation(rng):
    Qm, R = np.linalg.qr(rng.standard_normal((3, 3)))
    return Qm * np.sign(np.diag(R))


def _q9(A, b):
    return np.array([A[0, 0], A[0, 1], A[0, 2], -b[0], A[1, 1], A[1, 2], -b[1], A[2, 2], -b[2]])


def test_warm_started_solve_equals_the_cold_solve():
    # the sweep starts a trial solve from the eigenvectors of the cluster it
    # perturbs: the result must not depend on the start basis
    rng = np.random.default_rng(5)
    for it in range(300):
        # eigenvalue ratios stay away from the 1e-3 truncation threshold
        w = rng.choice([1.0, 1e-1, 1e-5], 3) if it % 3 else np.array([1.0, 0.3, 0.05])
        U = _random_rotation(rng)
        A = U @ np.diag(w) @ U.T
        b = rng.standard_normal(3)
        c = rng.standard_normal(3)
        Q = _q9(A, b)
        D0, rd0, V = po.hdr_rep_point_delta(Q, c, 3, want_V=True)
        # the returned basis diagonalises A and is orthonormal
        np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-13)
        off = V.T @ A @ V - np.diag(np.diag(V.T @ A @ V))
        assert np.abs(off).max() <= 1e-13 * np.abs(A).max()
        E = rng.standard_normal((3, 3)) * 1e-3 * w.max()
        _, V1 = np.linalg.eigh(A + E + E.T)
        for V0 in (V, V1, _random_rotation(rng)):
            D, rd = po.hdr_rep_point_delta(Q, c, 3, V0=V0)
            assert rd == rd0
            np.testing.assert_allclose(D, D0, rtol=1e-8, atol=1e-9 * (np.abs(D0).max() + np.abs(c).max()))
        # the identity start is the cold solve, bit for bit
        D, rd = po.hdr_rep_point_delta(Q, c, 3, V0=np.eye(3))
        assert rd == rd0 and np.array_equal(D, D0)


def test_warm_started_solve_kat_edge():
    tris = grid_tris((1.0, 0.8, 0.3), (0, 1, 0), (0, 0, 1), 4, 0.05)
    tris += grid_tris((0.8, 1.0, 0.3), (1, 0, 0), (0, 0, 1), 4, 0.05)
    Q = face_quadric(tris)
    c = np.array([0.93, 0.91, 0.37])
    rng = np.random.default_rng(6)
    for _ in range(20):
        D, rd, V = po.hdr_rep_point_delta(Q[:9], c, 3, V0=_random_rotation(rng), want_V=True)
        assert rd == 1
        np.testing.assert_allclose(c + D, [1.0, 1.0, 0.37], atol=1e-12)
        # the dropped direction is the edge line
        lam = np.diag(V.T @ np.array([[Q[0], Q[1], Q[2]], [Q[1], Q[4], Q[5]], [Q[2], Q[5], Q[7]]]) @ V)
        assert abs(abs(V[:, np.argmin(np.abs(lam))] @ [0, 0, 1]) - 1) < 1e-12
