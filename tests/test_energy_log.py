"""Energy log (SetWriteToGlobalEnergyLog, Common/vtkUniformClustering.h:123,
670-699, 1318-1346, 1383-1395): after each loop the sum of the cluster
energies, and after the run the energy of the final statistics.

CPU (oracle): the serial restatement's ComputeGlobalEnergy, with its bins,
equals the energy computed independently in numpy from the labels
(-|S|^2/W per cluster); within a phase the logged energy never rises (the
serial descent only accepts improving moves); the replay of the device
schedule logs one entry per pass, its last entry is the final energy when the
run ended on a pass without moves and without disconnected clusters, and its
final energy equals the numpy energy of its labels.
GPU: the library's log equals the replay's bit for bit (isotropic, QEM with and
without unconstrained initialisation, QEM with fixed clusters, a batch), its
clock never runs backwards, and the class API and acvd_run write energy.txt in
the reference's format ("loop seconds energy", then "Final Energy :<e>")."""
import os
import subprocess

import numpy as np
import pytest

import acvd_amd
from acvd_amd import meshgen
from oracle import pyoracle as po

ISO, QEM = acvd_amd.METRIC_ISOTROPIC, acvd_amd.METRIC_QEM
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(autouse=True)
def reset_oracle():
    yield
    po.set_energy_log(False)
    po.set_fixed(None)


def iso_energy(x, w, labels, K):
    """sum over clusters of -|S|^2 / W (vtkIsotropicMetricForClustering.h:114-121)."""
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    ok = (labels >= 0) & (labels < K)
    W = np.bincount(labels[ok], weights=w[ok], minlength=K)
    S = np.stack([np.bincount(labels[ok], weights=w[ok] * x[ok, d], minlength=K) for d in range(3)], 1)
    nz = W > 0
    return float(np.sum(-(S[nz] ** 2).sum(1) / W[nz]))


# ------------------------------------------------------------------ CPU
def test_serial_global_energy_equals_numpy():
    xyz, tri = meshgen.icosphere(4, 3, 0.02)
    m = po.RefMesh(xyz, tri)
    po.set_energy_log(True)
    K = 200
    lab, st = m.serial_run(ISO, K)
    log, fin = po.serial_energy_log()
    assert len(log) == st["loops"]
    w, _ = m.items(ISO)
    want = iso_energy(xyz, w, lab, K)
    assert abs(fin - want) <= 1e-12 * abs(want)
    # the serial descent accepts improving moves only: no rise inside a phase
    assert np.all(np.diff(log) <= 1e-13 * np.abs(log[1:]))


def test_serial_qem_final_energy_is_the_shifted_quadric_energy():
    """QEM: the final line sums W|p|^2 - 2 p.S with p the solved points, i.e.
    the unshifted energy minus sum w|x|^2 (a second restatement, energy_ref.h)."""
    xyz, tri = meshgen.icosphere(4, 5, 0.03)
    m = po.RefMesh(xyz, tri)
    po.set_energy_log(True)
    K = 150
    lab, st = m.serial_run(QEM, K, unconstrained=0)
    log, fin = po.serial_energy_log()
    unshifted, shifted, wx2 = m.energy(QEM, K, lab)
    assert abs(fin - shifted) <= 1e-10 * abs(shifted)
    assert abs((fin + wx2) - unshifted) <= 1e-8 * abs(wx2)
    assert len(log) == st["loops"]


def _replay(xyz, tri, metric, K, unc, init_seed=0):
    rp, col, xw, q, sc = po.device_inputs(xyz, tri, metric)
    init = po.RefMesh(xyz, tri).initial_sampling(metric, K)
    lab, st = po.replay_run(rp, col, xw, q if metric == QEM else None, sc, K, metric, init, seed=init_seed,
                            unconstrained=unc, energy_log=True)
    return lab, st, xw


@pytest.mark.parametrize("metric,unc", [(ISO, 0), (QEM, 1), (QEM, 0)])
def test_replay_log_properties(metric, unc):
    xyz, tri = meshgen.icosphere(4, 11, 0.03)
    K = 180
    lab, st, xw = _replay(xyz, tri, metric, K, unc)
    log, fin = st["energy_log"], st["final_energy"]
    assert len(log) == st["loops"]
    if st["last_disconnected"] == 0:
        # the last pass moved nothing and the clean changed nothing: the final
        # statistics are the last pass's, summed the same way
        assert log[-1] == fin
    if metric == ISO:
        e = iso_energy(xw[:, :3].astype(np.float64), xw[:, 3].astype(np.float64), lab, K)
        assert abs(fin - e) <= 1e-9 * abs(e)
    # the accepted moves lower the energy: no rise beyond rounding in a phase
    d = np.diff(log)
    assert np.sum(d > 1e-9 * np.abs(log[1:])) <= st["convergences"]


# ------------------------------------------------------------------ GPU
def _device_log(xyz, tri, metric, K, unc, fixed=None, seed=3):
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=K, metric=metric, unconstrained_initialization=unc, seed=seed,
                   write_energy_log=1)
    ctx.set_fixed_clusters(0, fixed)
    po.set_fixed(fixed)
    init = po.RefMesh(xyz, tri).initial_sampling(metric, K)
    ctx.set_initial_labels(0, init)
    ctx.process()
    return ctx, init


@pytest.mark.gpu
@pytest.mark.parametrize("metric,unc,nfix", [(ISO, 0, 0), (QEM, 1, 0), (QEM, 0, 0), (QEM, 1, 12)])
def test_device_log_equals_replay(gpu_ready, metric, unc, nfix):
    xyz, tri = meshgen.icosphere(5, 12, 0.03)
    K = 400
    fixed = None
    if nfix:
        fixed = np.sort(np.random.default_rng(4).choice(len(xyz), nfix, replace=False)).astype(np.int32)
    print(f"energy log: metric {metric}, unconstrained {unc}, fixed {nfix}", flush=True)
    ctx, init = _device_log(xyz, tri, metric, K, unc, fixed)
    sec, log, fin = ctx.energy_log(0)
    rp, col = ctx.adjacency(0)
    xw, q = ctx.items(0)
    rlab, rst = po.replay_run(rp, col, xw, q, ctx.scales(0), K, metric, init, seed=3, unconstrained=unc,
                              energy_log=True)
    np.testing.assert_array_equal(ctx.labels(0), rlab)
    assert len(log) == ctx.stats(0)["loops"] == rst["loops"]
    np.testing.assert_array_equal(log, rst["energy_log"])
    assert fin == rst["final_energy"]
    assert np.all(sec >= 0) and np.all(np.diff(sec) >= 0)
    assert sec[-1] < 60.0


@pytest.mark.gpu
def test_batch_logs_only_the_meshes_that_ask(gpu_ready):
    a = meshgen.icosphere(4, 1, 0.02)
    b = meshgen.icosphere(4, 2, 0.02)
    ctx = acvd_amd.Context()
    for i, (xyz, tri) in enumerate((a, b)):
        mid = ctx.add_input(xyz, tri)
        ctx.set_params(mid, number_of_clusters=150, metric=ISO, seed=i, write_energy_log=1 if i == 0 else 0)
    ctx.process()
    sec, log, fin = ctx.energy_log(0)
    assert len(log) == ctx.stats(0)["loops"]
    with pytest.raises(acvd_amd.AcvdError):
        ctx.energy_log(1)
    # the logged mesh's energy equals the numpy energy of its labels
    xw, _ = ctx.items(0)
    e = iso_energy(xw[:, :3].astype(np.float64), xw[:, 3].astype(np.float64), ctx.labels(0), 150)
    assert abs(fin - e) <= 1e-9 * abs(e)


def _parse(path):
    with open(path) as f:
        lines = f.read().splitlines()
    assert lines[-1].startswith("Final Energy :")
    rows = [ln.split(" ") for ln in lines[:-1]]
    return rows, float(lines[-1][len("Final Energy :"):])


@pytest.mark.gpu
def test_class_api_writes_energy_txt(gpu_ready, tmp_path):
    xyz, tri = meshgen.icosphere(5, 13, 0.02)
    r = acvd_amd.QIsotropicDiscreteRemeshing.New()
    r.SetInput(acvd_amd.Mesh(xyz, tri))
    r.SetNumberOfClusters(500)
    r.SetUnconstrainedInitialization(1)
    r.SetWriteToGlobalEnergyLog(1)
    r.SetOutputDirectory(str(tmp_path) + "/")
    r.Remesh()
    rows, fin = _parse(tmp_path / "energy.txt")
    sec, log, efin = r.GetContext().energy_log(0)
    assert len(rows) == len(log) == r.GetStats()["loops"]
    for i, row in enumerate(rows):
        assert len(row) == 3 and int(row[0]) == i
        assert row[2] == "%.15g" % log[i]
        assert float(row[1]) == pytest.approx(sec[i], rel=1e-5, abs=1e-6)
    assert fin == float("%.15g" % efin)
    r.Delete()


@pytest.mark.gpu
def test_acvd_run_w_writes_energy_txt(gpu_ready, tmp_path):
    exe = os.path.join(ROOT, "acvd_amd", "acvd_run")
    xyz, tri = meshgen.icosphere(5, 14, 0.02)
    src = tmp_path / "in.ply"
    with open(src, "w") as f:
        f.write("ply\nformat ascii 1.0\nelement vertex %d\nproperty float x\nproperty float y\nproperty float z\n"
                "element face %d\nproperty list uchar int vertex_indices\nend_header\n" % (len(xyz), len(tri)))
        for p in xyz:
            f.write("%.9g %.9g %.9g\n" % tuple(p))
        for t in tri:
            f.write("3 %d %d %d\n" % tuple(t))
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([exe, str(src), "300", "0", "-w", "1", "-o", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows, fin = _parse(out / "energy.txt")
    assert len(rows) > 0 and all(int(row[0]) == i for i, row in enumerate(rows))
    assert fin < 0
