"""Initial sampling type 0 (InitSamples, Common/vtkUniformClustering.h:1123-1159):
one item per cluster drawn with vtkMath::Random seeded 5000.

CPU: the oracle's generator meets Park and Miller's published check (from
x = 1 the 10000th value of the minimal standard generator is 1043618065),
the library's host sampling (Schrage's method) equals the oracle's (plain
64-bit products) label for label, every cluster gets exactly one item, and
K = N labels every item.  GPU: a run with initial_sampling = 2 (type 0) or
3 (type -1) equals the replay started from those labels."""
import numpy as np
import pytest

import acvd_amd
from acvd_amd import meshgen
from acvd_amd import remeshing as rem
from oracle import pyoracle as po


def test_minimal_standard_known_answer():
    assert po.minstd_10000() == 1043618065


@pytest.mark.parametrize("level,K", [(3, 1), (4, 10), (4, 700), (4, 2562), (5, 4000)])
def test_seed_sampling_equals_oracle(level, K):
    xyz, tri = meshgen.icosphere(level, 1, 0.0)
    n = len(xyz)
    got = rem.reference_seed_sampling(n, K)
    want = po.RefMesh(xyz, tri).seed_sampling(K)
    np.testing.assert_array_equal(got, want)
    lab = got[got < K]
    assert len(lab) == K and len(np.unique(lab)) == K
    assert np.sum(got == K) == n - K


def test_seed_sampling_rejects_too_many_clusters():
    with pytest.raises(acvd_amd.AcvdError):
        rem.reference_seed_sampling(10, 11)


@pytest.mark.gpu
@pytest.mark.parametrize("metric,mode", [(acvd_amd.METRIC_ISOTROPIC, 2), (acvd_amd.METRIC_QEM, 2),
                                         (acvd_amd.METRIC_ISOTROPIC, 3)])
def test_type0_run_equals_replay(gpu_ready, metric, mode):
    """mode 2: the reference's type 0; mode 3: its type -1 (item k starts
    cluster k)."""
    xyz, tri = meshgen.icosphere(5, 22, 0.03)
    K = 500
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=K, metric=metric, initial_sampling=mode, seed=4,
                   unconstrained_initialization=int(metric == acvd_amd.METRIC_QEM))
    ctx.process()
    if mode == 2:
        init = rem.reference_seed_sampling(len(xyz), K)
    else:
        init = np.where(np.arange(len(xyz)) < K, np.arange(len(xyz)), K).astype(np.int32)
    rp, col = ctx.adjacency(0)
    xw, q = ctx.items(0)
    rlab, _ = po.replay_run(rp, col, xw, q, ctx.scales(0), K, metric, init, seed=4,
                            unconstrained=int(metric == acvd_amd.METRIC_QEM))
    np.testing.assert_array_equal(ctx.labels(0), rlab)
