"""Fixed-assignment energy and its adjoint (no reference counterpart: the
reference never differentiates its energy).

Pins (CPU): the dual-number derivatives of the oracle (oracle/energy_ref.h)
equal the closed form of the isotropic gradient (dE/dx_i = 2 w_i (x_i - c),
dE/dw_i = |x_i - c|^2) and FP64 central differences of the QEM energy.
GPU: the forward equals the oracle's FP64 energy (rtol 1e-9; FP32 inputs read
as FP64 on both sides); the adjoint equals the oracle's derivatives at every
coordinate (tolerance 1e-5 of the largest component: FP32 outputs) and FP64
central differences (step 1e-6) at 1,000 seeded coordinates with relative
error <= 1e-4 (relative to max(|fd|, 1e-3 max|fd|)); autograd plumbing and
batches of mixed metrics."""
import numpy as np
import pytest

import acvd_amd
from acvd_amd import meshgen
from oracle import pyoracle as po

ISO, QEM = acvd_amd.METRIC_ISOTROPIC, acvd_amd.METRIC_QEM


def small_problem(metric, level=3, seed=4):
    xyz, tri = meshgen.icosphere(3, seed, 0.05)
    rng = np.random.default_rng(seed)
    x = np.asarray(xyz, np.float32).astype(np.float64)
    w = rng.uniform(0.5, 2.0, len(x)).astype(np.float32).astype(np.float64)
    K = 40
    lab = po.RefMesh(xyz, tri).fill_holes(K, meshgen.seed_labels(len(x), K, seed))
    return x, w, tri, lab, K


def test_iso_dual_derivatives_match_closed_form():
    x, w, tri, lab, K = small_problem(ISO)
    ref = po.EnergyRef(len(x), tri, lab, K, ISO)
    n = len(x)
    coords = np.arange(4 * n)
    d = ref.derivatives(x, w, coords)
    W = np.bincount(lab, weights=w, minlength=K + 1)[:K]
    S = np.stack([np.bincount(lab, weights=w * x[:, k], minlength=K + 1)[:K] for k in range(3)], 1)
    c = S / W[:, None]
    diff = x - c[lab]
    gx = 2.0 * w[:, None] * diff
    gw = (diff ** 2).sum(1)
    np.testing.assert_allclose(d[:3 * n], gx.reshape(-1), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(d[3 * n:], gw, rtol=1e-9, atol=1e-12)
    # and the energy itself
    assert np.isclose(ref.total(x, w), float((w * gw).sum()), rtol=1e-12)


@pytest.mark.parametrize("level", [1, 2, 3])
def test_qem_dual_derivatives_match_central_differences(level):
    x, w, tri, lab, K = small_problem(QEM, level)
    ref = po.EnergyRef(len(x), tri, lab, K, QEM, level)
    rng = np.random.default_rng(level)
    coords = rng.choice(4 * len(x), 300, replace=False)
    d = ref.derivatives(x, w, coords)
    fd = ref.central_differences(x, w, coords, 1e-6)
    scale = np.abs(fd).max()
    err = np.abs(d - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    assert np.quantile(err, 0.99) < 1e-5 and err.max() < 1e-3, (np.quantile(err, 0.99), err.max())


def test_qem_point_is_quadric_minimiser_for_full_rank_clusters():
    # with level 3 and full-rank cluster quadrics the QEM energy does not
    # depend on the centroid: perturbing a weight only moves the direct term
    x, w, tri, lab, K = small_problem(QEM)
    ref = po.EnergyRef(len(x), tri, lab, K, QEM, 3)
    n = len(x)
    coords = 3 * n + np.arange(0, n, 7)
    d = ref.derivatives(x, w, coords)
    fd = ref.central_differences(x, w, coords, 1e-6)
    np.testing.assert_allclose(d, fd, rtol=1e-5, atol=1e-9)


# ------------------------------------------------------------------- GPU
def gpu_inputs(x, w, lab, K):
    import torch
    dev = torch.device("cuda", 0)
    lab_gpu = np.where((lab >= 0) & (lab < K), lab, K).astype(np.int32)
    return (torch.tensor(x, dtype=torch.float32, device=dev).contiguous(),
            torch.tensor(w, dtype=torch.float32, device=dev).contiguous(),
            torch.tensor(lab_gpu, dtype=torch.int32, device=dev).contiguous())


@pytest.mark.gpu
@pytest.mark.parametrize("metric,level", [(ISO, 3), (QEM, 1), (QEM, 2), (QEM, 3)])
def test_gpu_energy_forward_and_adjoint_match_oracle(gpu_ready, metric, level):
    from acvd_amd.energy import ClusterEnergy
    x, w, tri, lab, K = small_problem(metric, level)
    ref = po.EnergyRef(len(x), tri, lab, K, metric, level)
    ce = ClusterEnergy(0)
    ce.add_mesh(len(x), tri, K, metric, level)
    tx, tw, tl = gpu_inputs(x, w, lab, K)
    e = ce.forward_raw(tx, tw, tl).cpu().numpy()
    assert np.isclose(e[0], ref.total(x, w), rtol=1e-9), (e[0], ref.total(x, w))
    import torch
    dx, dw = ce.backward_raw(tx, tw, tl, torch.ones(1, dtype=torch.float64, device=tx.device))
    g = np.concatenate([dx.cpu().numpy().reshape(-1), dw.cpu().numpy()]).astype(np.float64)
    d = ref.derivatives(x, w, np.arange(4 * len(x)))
    scale = np.abs(d).max()
    assert np.abs(g - d).max() < 1e-5 * scale, np.abs(g - d).max() / scale


@pytest.mark.gpu
def test_gpu_adjoint_against_central_differences_torus(gpu_ready):
    """1,000 seeded coordinates (positions and weights) of a QEM torus with
    frozen labels from a GPU clustering."""
    import torch
    from acvd_amd.energy import ClusterEnergy
    xyz, tri = meshgen.torus(256, 128, 5)
    H = meshgen.mean_curvature(xyz, tri)
    K = 1500
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=K, metric=QEM, unconstrained_initialization=1, gradation=1.5, seed=5)
    ctx.set_indicator(0, H)
    ctx.process()
    lab = ctx.labels(0)
    xw, _ = ctx.items(0)
    x = np.asarray(xyz, np.float32).astype(np.float64)
    w = xw[:, 3].astype(np.float64)
    ce = ClusterEnergy(0)
    ce.add_mesh(len(x), tri, K, QEM, 3)
    tx, tw, tl = gpu_inputs(x, w, lab, K)
    E = ce(tx.requires_grad_(), tw.requires_grad_(), tl)
    E.sum().backward()
    g = np.concatenate([tx.grad.cpu().numpy().reshape(-1), tw.grad.cpu().numpy()]).astype(np.float64)
    ref = po.EnergyRef(len(x), tri, lab, K, QEM, 3)
    assert np.isclose(E.item(), ref.total(x, w), rtol=1e-9)
    rng = np.random.default_rng(5)
    coords = rng.choice(4 * len(x), 1000, replace=False)
    fd = ref.central_differences(x, w, coords, 1e-6)
    fd7 = ref.central_differences(x, w, coords, 1e-7)
    dual = ref.derivatives(x, w, coords)
    scale = np.abs(dual).max()
    # the QEM energy is not differentiable where a cluster's kept eigen-set
    # changes; such coordinates show difference quotients that depend on the step
    smooth = np.abs(fd - fd7) <= 1e-3 * np.maximum(np.abs(fd), 1e-3 * scale)
    assert smooth.sum() >= 990
    err = np.abs(g[coords] - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    # fourth-order differences (Richardson over steps 1e-6 and 2e-6): the
    # second-order quotient's h^2 f''' term dominates its error where the
    # third derivative is large
    fd2 = ref.central_differences(x, w, coords, 2e-6)
    fd4 = (4.0 * fd - fd2) / 3.0
    err4 = np.abs(g[coords] - fd4) / np.maximum(np.abs(fd4), 1e-3 * scale)
    smooth4 = smooth & (np.abs(fd2 - fd) <= 1e-3 * np.maximum(np.abs(fd), 1e-3 * scale))
    print(f"FD check: {int(smooth.sum())} smooth coordinates, median rel err {np.median(err[smooth]):.2e}, "
          f"max {err[smooth].max():.2e}; fourth order max {err4[smooth4].max():.2e}")
    assert np.quantile(err[smooth], 0.99) <= 1e-4 and err[smooth].max() <= 1e-3
    assert smooth4.sum() >= 990 and err4[smooth4].max() <= 1e-4
    # the dual-number derivative is exact at the point, smooth or not
    err_dual = np.abs(g[coords] - dual) / np.maximum(np.abs(dual), 1e-3 * scale)
    print(f"dual check: max rel err {err_dual.max():.2e}")
    assert err_dual.max() <= 1e-6, err_dual.max()


@pytest.mark.gpu
def test_gpu_energy_batch_equals_separate(gpu_ready):
    import torch
    from acvd_amd.energy import ClusterEnergy
    probs = [small_problem(ISO, 3, 4), small_problem(QEM, 3, 6)]
    ce = ClusterEnergy(0)
    xs, ws, ls = [], [], []
    singles = []
    for (x, w, tri, lab, K), metric in zip(probs, (ISO, QEM)):
        ce.add_mesh(len(x), tri, K, metric, 3)
        tx, tw, tl = gpu_inputs(x, w, lab, K)
        xs.append(tx), ws.append(tw), ls.append(tl)
        one = ClusterEnergy(0)
        one.add_mesh(len(x), tri, K, metric, 3)
        e1 = one.forward_raw(tx, tw, tl)
        d1 = one.backward_raw(tx, tw, tl, torch.ones(1, dtype=torch.float64, device=tx.device))
        singles.append((e1.item(), d1))
    tx, tw, tl = torch.cat(xs).contiguous(), torch.cat(ws).contiguous(), torch.cat(ls).contiguous()
    e = ce.forward_raw(tx, tw, tl).cpu().numpy()
    dx, dw = ce.backward_raw(tx, tw, tl, torch.ones(2, dtype=torch.float64, device=tx.device))
    n0 = len(probs[0][0])
    for k, (e1, (dx1, dw1)) in enumerate(singles):
        assert np.isclose(e[k], e1, rtol=1e-12)
        sl = slice(0, n0) if k == 0 else slice(n0, None)
        np.testing.assert_allclose(dx[sl].cpu().numpy(), dx1.cpu().numpy(), rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(dw[sl].cpu().numpy(), dw1.cpu().numpy(), rtol=1e-6, atol=1e-7)


@pytest.mark.gpu
def test_gpu_energy_errors(gpu_ready):
    import torch
    from acvd_amd.energy import ClusterEnergy
    x, w, tri, lab, K = small_problem(ISO)
    ce = ClusterEnergy(0)
    with pytest.raises(acvd_amd.AcvdError):
        ce.add_mesh(len(x), tri, 0, ISO)
    ce.add_mesh(len(x), tri, K, ISO)
    tx, tw, tl = gpu_inputs(x, w, lab, K)
    with pytest.raises(acvd_amd.AcvdError):
        ce.backward_raw(tx, tw, tl, torch.ones(1, dtype=torch.float64, device=tx.device))  # no forward yet
    with pytest.raises(acvd_amd.AcvdError):
        ce.forward_raw(tx.double(), tw, tl)
