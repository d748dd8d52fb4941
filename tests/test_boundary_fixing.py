"""Boundary fixing (SetBoundaryFixing, FixMeshBoundaries,
DiscreteRemeshing/vtkDiscreteRemeshing.h:425-579, run last by Remesh :952).

CPU: the restatement (oracle/manifold.py) on a hand-made case whose answer
the rule fixes (an open grid cut in two clusters: two midpoints, the two cut
triangles and one triangle per cluster run), nothing on a closed mesh; the
library's host rule equals the restatement on open meshes with many clusters.
GPU: Remesh with boundary fixing returns the dual triangulation followed by
exactly what the rule adds for the final labels, and the quadric
post-placement keeps the added vertices and triangles."""
import numpy as np
import pytest

import acvd_amd
from acvd_amd import meshgen
from acvd_amd import remeshing as rem
from oracle import manifold as om


def grid(n):
    xyz = np.array([[i, j, 0.0] for j in range(n) for i in range(n)])
    tri = []
    for j in range(n - 1):
        for i in range(n - 1):
            a, b, c, d = j * n + i, j * n + i + 1, (j + 1) * n + i + 1, (j + 1) * n + i
            tri += [[a, b, c], [a, c, d]]
    return xyz, np.array(tri, np.int32)


def voronoi(xyz, K, seed):
    rng = np.random.default_rng(seed)
    seeds = xyz[rng.choice(len(xyz), K, replace=False)]
    d = ((xyz[:, None, :] - seeds[None, :, :]) ** 2).sum(-1)
    return d.argmin(1).astype(np.int32)


def test_restatement_on_a_grid_cut_in_two():
    xyz, tri = grid(6)
    lab = (xyz[:, 0] >= 3).astype(np.int32)
    add_xyz, add_tri = om.fix_mesh_boundaries(len(xyz), xyz, tri, lab, 2)
    assert sorted(map(tuple, add_xyz)) == [(2.5, 0.0, 0.0), (2.5, 5.0, 0.0)]
    assert len(add_tri) == 4
    sets = sorted(tuple(sorted(t)) for t in add_tri)
    assert sets == [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]


def test_restatement_adds_nothing_on_a_closed_mesh():
    xyz, tri = meshgen.icosphere(3, 1, 0.0)
    lab = voronoi(np.asarray(xyz, np.float64), 20, 1)
    assert om.fix_mesh_boundaries(len(xyz), xyz, tri, lab, 20) == ([], [])


@pytest.mark.parametrize("case", range(5))
def test_host_rule_equals_restatement(case):
    if case < 2:
        xyz, tri = grid(14)
        K = (8, 25)[case]
    else:
        xyz, tri = meshgen.icosphere(4, 2, 0.02)
        xyz = np.asarray(xyz, np.float64)
        tri = tri[np.random.default_rng(case).random(len(tri)) > 0.06]
        K = (20, 60, 150)[case - 2]
    lab = voronoi(np.asarray(xyz, np.float64), K, case)
    ax, at = rem.mesh_boundary_fix(xyz, tri, lab, K)
    bx, bt = om.fix_mesh_boundaries(len(xyz), xyz, tri, lab, K)
    assert len(ax) == len(bx) and len(at) == len(bt)
    if len(bx):
        np.testing.assert_array_equal(ax, np.array(bx))
        np.testing.assert_array_equal(at, np.array(bt))


@pytest.mark.gpu
def test_remesh_with_boundary_fixing(gpu_ready):
    xyz, tri = meshgen.icosphere(5, 7, 0.02)
    xyz = np.asarray(xyz, np.float64)
    tri = tri[xyz[tri[:, 0], 2] < 0.6]  # a sphere with its cap cut off: one boundary loop
    used = np.unique(tri)
    remap = np.full(len(xyz), -1, np.int32)
    remap[used] = np.arange(len(used))
    xyz, tri = xyz[used], remap[tri]
    K = 300
    r = acvd_amd.IsotropicDiscreteRemeshing.New()
    r.SetInput(acvd_amd.Mesh(xyz, tri))
    r.SetNumberOfClusters(K)
    r.SetBoundaryFixing(1)
    r.Remesh()
    out = r.GetOutput()
    lab = r.GetClustering()
    base_xyz, base_tri = r.GetContext().output(0, 0)
    ax, at = rem.mesh_boundary_fix(xyz, tri, lab, K)
    print(f"boundary fixing: {len(ax)} vertices, {len(at)} triangles added", flush=True)
    assert len(ax) > 0
    np.testing.assert_array_equal(out.points, np.vstack([base_xyz, ax]))
    np.testing.assert_array_equal(out.triangles, np.vstack([base_tri, at]))
    placed = r.QuadricPostProcessing(1)
    assert len(placed.points) == K + len(ax)
    np.testing.assert_array_equal(placed.points[K:], ax)
    np.testing.assert_array_equal(placed.triangles, np.vstack([base_tri, at]))
    r.Delete()
