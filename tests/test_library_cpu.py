"""The C-ABI library loads without a GPU, exports every entry point declared
in include/acvd_c.h, and fails loudly (status + message) when no device is
visible.  No compute calls here."""
import ctypes
import os
import re

import numpy as np
import pytest

import acvd_amd
from acvd_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "acvd_c.h")).read()
    return sorted(set(re.findall(r"\b(acvd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(N.EXPORTS) == syms


def test_version_and_defaults():
    L = N.lib()
    assert b"gfx950" in L.acvd_version()
    p = N.Params()
    L.acvd_params_init(ctypes.byref(p))
    assert p.number_of_clusters == 0 and p.metric == 0 and p.quadrics_optimization_level == 3
    assert p.max_number_of_loops == 5000000 and p.max_number_of_convergences == 1000000000
    assert p.gradation == 0.0 and p.unconstrained_initialization == 0


@pytest.mark.skipif(N.lib().acvd_device_count() > 0, reason="a device is visible")
def test_no_device_fails_loudly():
    with pytest.raises(N.AcvdError, match="no HIP device"):
        acvd_amd.Context()
    r = acvd_amd.IsotropicDiscreteRemeshing.New()
    xyz, tri = acvd_amd.meshgen.icosphere(2)
    r.SetInput(acvd_amd.Mesh(xyz, tri))
    r.SetNumberOfClusters(10)
    with pytest.raises(N.AcvdError):
        r.ProcessClustering()


def test_zero_clusters_returns_none_like_the_reference():
    r = acvd_amd.QIsotropicDiscreteRemeshing.New()
    xyz, tri = acvd_amd.meshgen.icosphere(1)
    r.SetInput(acvd_amd.Mesh(xyz, tri))
    assert r.ProcessClustering() is None  # vtkUniformClustering.h:659-664


def test_null_context_is_rejected():
    L = N.lib()
    assert L.acvd_process_clustering(None) == -1
    assert b"null context" in L.acvd_last_error()
    assert L.acvd_destroy(None) == 0


DROPIN = os.path.join(ROOT, "tests", "cpp", "_build", "dropin")


@pytest.mark.skipif(N.lib().acvd_device_count() > 0, reason="a device is visible")
def test_cpp_dropin_compiles_links_and_fails_loudly_without_device():
    import subprocess
    assert os.path.exists(DROPIN), "build() compiles tests/cpp"
    r = subprocess.run([DROPIN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert r.stdout.startswith("fail") and "no HIP device" in r.stdout


@pytest.mark.parametrize("case", ["ico-iso", "cube-qem", "torus-iso-graded", "torus-qem-graded", "small-K-N",
                                  "large-K3000", "large-K300-graded", "large-K-half"])
def test_reference_initial_sampling_equals_oracle(case):
    """The product's host restatement of the reference's default initial
    sampling (no device needed) gives the oracle's labels exactly: same
    topology order, same FP64 weights, same std::mt19937 shuffle.  The
    meshes of 2^18 items and more ("large-*", 294,912 items) take the
    speculative parallel growth, whose regions must be the serial loop's."""
    from acvd_amd.remeshing import reference_initial_sampling
    from oracle import pyoracle as po
    ind, grad = None, 0.0
    if case == "ico-iso":
        (xyz, tri), metric, K = acvd_amd.meshgen.icosphere(4, 1, 0.02), acvd_amd.METRIC_ISOTROPIC, 150
    elif case == "cube-qem":
        (xyz, tri), metric, K = acvd_amd.meshgen.cube(20, 2, 1e-4), acvd_amd.METRIC_QEM, 120
    elif case == "small-K-N":
        (xyz, tri), metric, K = acvd_amd.meshgen.icosphere(2, 1, 0.0), acvd_amd.METRIC_ISOTROPIC, 100
    elif case.startswith("large"):
        xyz, tri = acvd_amd.meshgen.torus(768, 384, 5)
        metric = acvd_amd.METRIC_QEM
        K = {"large-K3000": 3000, "large-K-half": len(xyz) // 2}.get(case, 300)
        if case.endswith("graded"):
            ind, grad = acvd_amd.meshgen.mean_curvature(xyz, tri), 1.5
    else:
        xyz, tri = acvd_amd.meshgen.torus(96, 48, 3)
        ind, grad = acvd_amd.meshgen.mean_curvature(xyz, tri), 1.5
        metric = acvd_amd.METRIC_ISOTROPIC if "iso" in case else acvd_amd.METRIC_QEM
        K = 300
    got = reference_initial_sampling(xyz, tri, K, metric, ind, grad)
    want = po.RefMesh(xyz, tri).initial_sampling(metric, K, grad, ind)
    np.testing.assert_array_equal(got, want)
    assert got.min() >= 0 and got.max() <= K


def test_one_hip_runtime_per_process():
    """The library binds to the HIP runtime torch loaded (same soname), so
    device pointers and streams of torch tensors are valid in it."""
    N.lib()
    import torch  # noqa: F401
    maps = open("/proc/self/maps").read().splitlines()
    hip = {line.split()[-1] for line in maps if "libamdhip64" in line}
    hsa = {line.split()[-1] for line in maps if "libhsa-runtime64" in line}
    assert len(hip) == 1 and len(hsa) == 1, (hip, hsa)
