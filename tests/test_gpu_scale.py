"""Batched sweeps and the differentiable energy at larger sizes (GPU).

- 48 star-shaped icospheres (levels 4-7, 2,562-163,842 vertices, K = V/15,
  isotropic and QEM alternating, the batch family of the benchmark): the
  packed context gives every mesh the labels it gets alone, and each packed
  mesh equals the replay of its own schedule.
- 2 QEM tori of 1,048,576 vertices each (K = 25,000, |H|^1.5 weights, the
  energy benchmark's family) with labels frozen from one batched clustering:
  per-mesh forward energy = the FP64 oracle's total to 1e-9, the adjoint =
  the oracle's dual-number derivative to 1e-6 at 300 seeded coordinates."""
import numpy as np
import pytest

import acvd_amd
from acvd_amd import meshgen
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ISO, QEM = acvd_amd.METRIC_ISOTROPIC, acvd_amd.METRIC_QEM


def test_batch_48_star_icospheres(gpu_ready):
    levels = meshgen.batch_levels(48, 4, 7, seed=9)
    meshes = []
    for j, L in enumerate(levels):
        xyz, tri = meshgen.star_icosphere(int(L), seed=7000 + j)
        meshes.append((xyz, tri, len(xyz) // 15, ISO if j % 2 == 0 else QEM))
    print(f"batch: {len(meshes)} meshes, {sum(len(m[0]) for m in meshes)} vertices", flush=True)
    ctx = acvd_amd.Context()
    for j, (xyz, tri, K, metric) in enumerate(meshes):
        mid = ctx.add_input(xyz, tri)
        ctx.set_params(mid, number_of_clusters=K, metric=metric, unconstrained_initialization=metric, seed=j)
    ctx.process()
    for j, (xyz, tri, K, metric) in enumerate(meshes):
        alone = acvd_amd.Context()
        alone.set_input(xyz, tri)
        alone.set_params(0, number_of_clusters=K, metric=metric, unconstrained_initialization=metric, seed=j)
        alone.process()
        np.testing.assert_array_equal(ctx.labels(j), alone.labels(0), err_msg=f"mesh {j}")
        alone.close()
    # every eighth mesh against the replay of its own schedule
    for j in range(0, len(meshes), 8):
        xyz, tri, K, metric = meshes[j]
        init = po.RefMesh(xyz, tri).initial_sampling(metric, K)
        rp, col = ctx.adjacency(j)
        xw, q = ctx.items(j)
        rlab, _ = po.replay_run(rp, col, xw, q, ctx.scales(j), K, metric, init, seed=j, unconstrained=metric)
        np.testing.assert_array_equal(ctx.labels(j), rlab, err_msg=f"mesh {j}")


def test_energy_two_megavertex_tori(gpu_ready):
    import torch
    from acvd_amd.energy import ClusterEnergy
    K = 25000
    ctx = acvd_amd.Context()
    meshes = []
    for j in range(2):
        xyz, tri = meshgen.torus(1024, 1024, seed=900 + j)
        mid = ctx.add_input(xyz, tri)
        ctx.set_params(mid, number_of_clusters=K, metric=QEM, unconstrained_initialization=1, gradation=1.5, seed=j)
        ctx.set_indicator(mid, meshgen.mean_curvature(xyz, tri))
        meshes.append((xyz, tri))
    ctx.process()
    ce = ClusterEnergy(0)
    xs, ws, ls = [], [], []
    for j, (xyz, tri) in enumerate(meshes):
        ce.add_mesh(len(xyz), tri, K, QEM, 3)
        xw, _ = ctx.items(j)
        xs.append(np.asarray(xyz, np.float32).astype(np.float64))
        ws.append(xw[:, 3].astype(np.float64))
        ls.append(ctx.labels(j))
    dev = torch.device("cuda", 0)
    x = torch.tensor(np.concatenate(xs), dtype=torch.float32, device=dev).requires_grad_()
    w = torch.tensor(np.concatenate(ws), dtype=torch.float32, device=dev).requires_grad_()
    lab = torch.tensor(np.concatenate(ls).astype(np.int32), device=dev)
    E = ce(x, w, lab)
    E.sum().backward()
    n = len(xs[0])
    for j in range(2):
        ref = po.EnergyRef(n, meshes[j][1], ls[j], K, QEM, 3)
        assert np.isclose(E[j].item(), ref.total(xs[j], ws[j]), rtol=1e-9), (j, E[j].item())
    gx = x.grad.detach().cpu().numpy()[n:].astype(np.float64)
    gw = w.grad.detach().cpu().numpy()[n:].astype(np.float64)
    g = np.concatenate([gx.reshape(-1), gw])
    ref = po.EnergyRef(n, meshes[1][1], ls[1], K, QEM, 3)
    coords = np.random.default_rng(11).choice(4 * n, 300, replace=False)
    dual = ref.derivatives(xs[1], ws[1], coords)
    scale = np.abs(dual).max()
    err = np.abs(g[coords] - dual) / np.maximum(np.abs(dual), 1e-3 * scale)
    print(f"two 1M-vertex tori: dual check max rel err {err.max():.2e}", flush=True)
    assert err.max() <= 1e-6
