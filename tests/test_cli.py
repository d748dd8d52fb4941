"""The command-line caller acvd_amd/acvd_run (the argument grammar of
DiscreteRemeshing/Examples/ACVD.cxx:60-131 and ACVDQ.cxx:319-338): mesh
reading (OBJ, ASCII and binary PLY), argument handling, and, on the GPU, its
output meshes against the Python API's on the same input and parameters."""
import os
import struct
import subprocess

import numpy as np
import pytest

from acvd_amd import meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "acvd_amd", "acvd_run")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="acvd_run not built (run __graft_entry__.build())")


def write_obj(path, xyz, tri):
    with open(path, "w") as f:
        for p in xyz:
            f.write("v %r %r %r\n" % tuple(float(v) for v in p))
        for t in tri:
            f.write("f %d %d %d\n" % tuple(int(v) + 1 for v in t))


def write_ply(path, xyz, tri, binary):
    head = ("ply\nformat %s 1.0\nelement vertex %d\nproperty double x\nproperty double y\nproperty double z\n"
            "element face %d\nproperty list uchar int vertex_indices\nend_header\n"
            % ("binary_little_endian" if binary else "ascii", len(xyz), len(tri)))
    with open(path, "wb") as f:
        f.write(head.encode())
        if binary:
            for p in xyz:
                f.write(struct.pack("<3d", *p))
            for t in tri:
                f.write(struct.pack("<B3i", 3, *(int(v) for v in t)))
        else:
            for p in xyz:
                f.write(("%r %r %r\n" % tuple(float(v) for v in p)).encode())
            for t in tri:
                f.write(("3 %d %d %d\n" % tuple(int(v) for v in t)).encode())


def read_ply_ascii(path):
    with open(path) as f:
        lines = f.read().splitlines()
    nv = int(next(l for l in lines if l.startswith("element vertex")).split()[2])
    nf = int(next(l for l in lines if l.startswith("element face")).split()[2])
    body = lines[lines.index("end_header") + 1:]
    xyz = np.array([[float(v) for v in l.split()] for l in body[:nv]], np.float64).reshape(-1, 3)
    tri = np.array([[int(v) for v in l.split()[1:]] for l in body[nv:nv + nf]], np.int64).reshape(-1, 3)
    return xyz, tri


def run_cli(*args, cwd=None):
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=300)


def test_usage():
    r = run_cli()
    assert r.returncode == 0 and "Usage : acvd_run file nvertices gradation" in r.stdout


@pytest.mark.parametrize("fmt", ["obj", "ply-ascii", "ply-binary"])
def test_reads_meshes(tmp_path, fmt):
    xyz, tri = meshgen.icosphere(2, 1, 0.02)
    path = tmp_path / ("m.obj" if fmt == "obj" else "m.ply")
    if fmt == "obj":
        write_obj(path, xyz, tri)
    else:
        write_ply(path, xyz, tri, binary=fmt == "ply-binary")
    # K = 0 stops before any device work (ProcessClustering returns nothing,
    # vtkUniformClustering.h:659-664), after the mesh was read
    r = run_cli(path, 0, 0, cwd=tmp_path)
    assert f"load : {path} ({len(xyz)} vertices, {len(tri)} triangles)" in r.stdout
    assert r.returncode == 1 and "NumberOfClusters is 0" in r.stderr + r.stdout


def test_polygons_are_fanned(tmp_path):
    # one quad and one pentagon: 2 + 3 triangles
    path = tmp_path / "p.obj"
    path.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 2 0 0\nv 2 1 0\nv 1.5 2 0\nf 1 2 3 4\nf 2 5 6 7 3\n")
    r = run_cli(path, 0, 0, cwd=tmp_path)
    assert "(7 vertices, 5 triangles)" in r.stdout


def test_bad_input_is_an_error(tmp_path):
    path = tmp_path / "bad.obj"
    path.write_text("v 0 0 0\nf 1 2 3\n")
    r = run_cli(path, 10, 0, cwd=tmp_path)
    assert r.returncode == 1 and "out of range" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("qem,gradation", [(0, 0.0), (1, 0.0), (1, 1.5)])
def test_cli_output_equals_the_python_api(tmp_path, qem, gradation):
    import acvd_amd
    xyz, tri = meshgen.icosphere(4, 1, 0.02)
    K = 200
    write_obj(tmp_path / "in.obj", xyz, tri)
    print(f"acvd_run qem={qem} K={K} gradation={gradation}", flush=True)
    # with a positive gradation both sides weight the items by the curvature
    # indicator that Remesh computes on the device
    r = run_cli(tmp_path / "in.obj", K, gradation, "-qem", qem, "-of", "out.ply", cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    cx, ct = read_ply_ascii(tmp_path / "out.ply")
    cls = acvd_amd.QIsotropicDiscreteRemeshing if qem else acvd_amd.IsotropicDiscreteRemeshing
    rem = cls.New()
    rem.SetInput((xyz, tri))
    rem.SetNumberOfClusters(K)
    rem.SetSubsamplingThreshold(10)
    rem.GetMetric().SetGradation(gradation)
    if qem:
        rem.GetMetric().SetQuadricsOptimizationLevel(3)
        rem.SetUnconstrainedInitialization(1)
    rem.Remesh()
    if not qem:
        # ACVD.cxx writes the clustering's mesh as smooth_<file>, then places
        # the vertices with the level-1 quadric rule
        sx, st = read_ply_ascii(tmp_path / "smooth_out.ply")
        out = rem.GetOutput()
        np.testing.assert_array_equal(sx, out.points)
        np.testing.assert_array_equal(st, out.triangles)
        out = rem.QuadricPostProcessing(1)
    else:
        out = rem.GetOutput()
    rem.Delete()
    np.testing.assert_array_equal(cx, out.points)
    np.testing.assert_array_equal(ct, out.triangles)
    assert len(cx) == K and len(ct) == 2 * K - 4  # closed genus-0 output


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["-fv", "-ft"])
def test_cli_fixed_vertices_are_kept(gpu_ready, tmp_path, kind):
    """ACVDQ's -fv / -ft (Examples/ACVDQ.cxx:246-281, 327-366): the fixed
    vertices become anchored clusters and the program's own check (output
    vertex i == input vertex fixed[i], coordinate for coordinate) passes."""
    xyz, tri = meshgen.icosphere(5, 4, 0.02)
    src = tmp_path / "in.ply"
    write_ply(src, np.asarray(xyz, np.float64), tri, binary=False)
    lst = tmp_path / "fixed.txt"
    if kind == "-fv":
        fixed = [17, 1200, 9000, 5]
        lst.write_text("\n".join(map(str, fixed)) + "\n")
    else:
        faces = [3, 400, 2047]
        lst.write_text(" ".join(map(str, faces)) + "\n")
        fixed = sorted({int(v) for f in faces for v in tri[f]})
    r = run_cli(src, 300, 0, "-qem", 1, kind, lst, "-o", tmp_path, "-of", "out.ply")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Constraints on vertices have been checked" in r.stdout
    out, _ = read_ply_ascii(tmp_path / "out.ply")
    assert len(out) == 300 + len(fixed)
    np.testing.assert_array_equal(out[:len(fixed)], np.asarray(xyz, np.float64)[fixed])
