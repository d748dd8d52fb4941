"""Launch-shape tuning (acvd_set_tuning, acvd_autotune, acvd_amd.autotune).

The knobs change how the passes are launched, never what they compute: every
shape gives the labels of the default one (the autotuner refuses a shape that
does not), and they equal the replay.  The size class of the cache is a pure
function of the context's sizes (CPU)."""
import json

import numpy as np
import pytest

import acvd_amd
from acvd_amd import autotune, meshgen
from oracle import pyoracle as po

ISO, QEM = acvd_amd.METRIC_ISOTROPIC, acvd_amd.METRIC_QEM


class _Sizes:
    def __init__(self, n, k, q, m):
        self._s, self.num_meshes = (n, k, q), m

    def sizes(self):
        return self._s


def test_size_class_keys():
    assert autotune.size_class(_Sizes(8388608, 100000, 1, 1)) == "n23-k16-qem"
    assert autotune.size_class(_Sizes(40962, 2000, 0, 1)) == "n15-k10-iso"
    assert autotune.size_class(_Sizes(62162688, 4144170, 192, 384)) == "n25-k21-mix"


def test_committed_cache_is_well_formed():
    cache = autotune.load()
    for key, e in cache.items():
        t = e["tuning"]
        assert set(t) == {"sweep_segments", "tile_on_divisor", "chunk_passes", "short_chunk_passes"}
        assert 1 <= t["sweep_segments"] <= 2048 and 1 <= t["short_chunk_passes"] <= t["chunk_passes"] <= 16
        json.dumps(e)


def _ctx(metric, K, seed=1):
    xyz, tri = meshgen.icosphere(5, 21, 0.03)
    ctx = acvd_amd.Context()
    ctx.set_input(xyz, tri)
    ctx.set_params(0, number_of_clusters=K, metric=metric, unconstrained_initialization=int(metric == QEM), seed=seed)
    init = po.RefMesh(xyz, tri).initial_sampling(metric, K)
    ctx.set_initial_labels(0, init)
    return ctx, init


@pytest.mark.gpu
def test_shapes_give_the_same_labels(gpu_ready):
    ctx, init = _ctx(QEM, 400)
    ctx.process()
    ref = ctx.labels(0)
    for t in ({"sweep_segments": 7}, {"sweep_segments": 64, "tile_on_divisor": 2},
              {"chunk_passes": 3, "short_chunk_passes": 1}, {"sweep_segments": 2048, "chunk_passes": 16}):
        print(f"tuning {t}", flush=True)
        ctx.set_tuning(**t)
        ctx.process()
        np.testing.assert_array_equal(ctx.labels(0), ref)
    rp, col = ctx.adjacency(0)
    xw, q = ctx.items(0)
    rlab, _ = po.replay_run(rp, col, xw, q, ctx.scales(0), 400, QEM, init, seed=1, unconstrained=1)
    np.testing.assert_array_equal(ref, rlab)


@pytest.mark.gpu
def test_tuning_arguments_are_checked(gpu_ready):
    ctx, _ = _ctx(ISO, 100)
    for bad in ({"sweep_segments": 0}, {"sweep_segments": 4096}, {"chunk_passes": 17},
                {"chunk_passes": 4, "short_chunk_passes": 5}, {"tile_on_divisor": 0}):
        with pytest.raises(acvd_amd.AcvdError):
            ctx.set_tuning(**bad)


@pytest.mark.gpu
def test_autotune_keeps_the_labels(gpu_ready):
    ctx, _ = _ctx(ISO, 300)
    ctx.process()
    ref = ctx.labels(0)
    best, ms = ctx.autotune(1)
    print(f"autotune: {best}, {ms:.2f} ms", flush=True)
    assert ctx.get_tuning() == best and ms > 0
    np.testing.assert_array_equal(ctx.labels(0), ref)
