"""acvd_energy_forward_cvt_noalloc (the entry point that allocates nothing)
against acvd_energy_forward (the context entry point) through the C-ABI, on
two calls in a row, the second with a smaller input (fewer items and
clusters), plus the rejection of a cluster count beyond its static scratch.
Tolerance: both sum FP64 terms with FP64 atomics in an order that varies from
run to run, so the energies agree to rtol 1e-12, not bit for bit."""
import ctypes

import numpy as np
import pytest

import acvd_amd
from acvd_amd import _native as N
from acvd_amd import meshgen
from acvd_amd.energy import ClusterEnergy

pytestmark = pytest.mark.gpu


def problem(level, K, seed):
    xyz, tri = meshgen.icosphere(level, seed, 0.05)
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 2.0, len(xyz)).astype(np.float32)
    lab = meshgen.seed_labels(len(xyz), K, seed).astype(np.int32)
    lab[lab == K] = np.arange(int(np.sum(lab == K)), dtype=np.int32) % K  # every item labelled
    return np.asarray(xyz, np.float32), tri, w, lab


def noalloc(x, w, lab, K, stream):
    import torch
    e = torch.full((1,), -1.0, dtype=torch.float64, device=x.device)
    N.check(N.lib().acvd_energy_forward_cvt_noalloc(0, x.shape[0], K, ctypes.c_void_p(x.data_ptr()),
                                                    ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(lab.data_ptr()),
                                                    ctypes.c_void_p(e.data_ptr()), ctypes.c_void_p(stream)))
    return e


def test_noalloc_forward_equals_context_forward(gpu_ready):
    import torch
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev).cuda_stream
    for level, K, seed in ((5, 3000, 1), (3, 70, 2)):  # the second call is smaller
        xyz, tri, w, lab = problem(level, K, seed)
        print(f"noalloc forward: icosphere level {level}, {len(xyz)} items, K={K}", flush=True)
        x = torch.tensor(xyz, device=dev).contiguous()
        wt = torch.tensor(w, device=dev).contiguous()
        lt = torch.tensor(lab, device=dev).contiguous()
        e_na = noalloc(x, wt, lt, K, stream)
        ce = ClusterEnergy(0)
        ce.add_mesh(len(xyz), tri, K, acvd_amd.METRIC_ISOTROPIC)
        e_ctx = ce.forward_raw(x, wt, lt)
        torch.cuda.synchronize()
        a, b = float(e_na.item()), float(e_ctx.item())
        print(f"  noalloc {a:.17e} context {b:.17e}", flush=True)
        assert a > 0.0
        assert np.isclose(a, b, rtol=1e-12, atol=0.0)
        ce.close()


def test_noalloc_rejects_too_many_clusters(gpu_ready):
    import torch
    dev = torch.device("cuda", 0)
    x = torch.zeros((4, 3), dtype=torch.float32, device=dev)
    w = torch.ones(4, dtype=torch.float32, device=dev)
    lab = torch.zeros(4, dtype=torch.int32, device=dev)
    with pytest.raises(acvd_amd.AcvdError, match="65536"):
        noalloc(x, w, lab, 65537, 0)
