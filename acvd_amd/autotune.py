"""Launch-shape autotuner and its cache.

The tuner (acvd_autotune in the library) times full clusterings of a
context under candidate launch shapes — the evaluate / accept-round grid,
the threshold that turns the scan's tile sets on, the passes per captured
graph launch — checks that all of them give the same labels, and keeps the
fastest.  Results are stored in tuning.json beside this file, keyed by the
size class of a context: log2 of its items, log2 of its clusters, and
whether its meshes are isotropic, QEM or both.  apply_cached() sets the
stored shape on a context of a known class; the library's defaults apply
otherwise.

    python -m acvd_amd.autotune --workload torus4096 [--repeats 2]

tunes the named workload on the local GPU and records it.
"""
import argparse
import json
import math
import os
import time

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def size_class(ctx):
    n, k, q = ctx.sizes()
    m = ctx.num_meshes
    kind = "qem" if q == m else ("iso" if q == 0 else "mix")
    return f"n{int(math.log2(max(n, 1)))}-k{int(math.log2(max(k, 1)))}-{kind}"


def load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(ctx):
    return load().get(size_class(ctx), {}).get("tuning")


def apply_cached(ctx):
    """Sets the cached launch shape of the context's size class; returns it
    (None: no entry, the library's defaults stay)."""
    t = lookup(ctx)
    if t:
        ctx.set_tuning(**t)
    return t


def tune(ctx, repeats=2, note="", save=True):
    """Runs the library's autotuner on ctx and records the result."""
    t0 = time.time()
    start = ctx.get_tuning()
    runs = []
    for _ in range(repeats):
        ctx.process()
        runs.append(ctx.stats(0)["converge_ms"])
    best, ms = ctx.autotune(repeats)
    entry = {"tuning": best, "ms": round(ms, 2), "start": start, "start_ms": round(min(runs), 2),
             "tuned_s": round(time.time() - t0, 1), "workload": note}
    if save:
        cache = load()
        cache[size_class(ctx)] = entry
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
            f.write("\n")
    return entry


def _workload(name):
    import numpy as np

    import acvd_amd
    from acvd_amd import meshgen
    ctx = acvd_amd.Context(0)
    if name == "torus4096":  # 8.39 M vertices -> 100,000 clusters, QEM, gradation 1.5 on |H|
        xyz, tri = meshgen.torus(4096, 2048, seed=3)
        ctx.set_input(xyz, tri)
        ctx.set_params(0, number_of_clusters=100000, metric=acvd_amd.METRIC_QEM, unconstrained_initialization=1,
                       gradation=1.5)
        ctx.set_indicator(0, meshgen.mean_curvature(xyz, tri))
    elif name == "icosphere6":
        xyz, tri = meshgen.icosphere(6, 1, 0.02)
        ctx.set_input(xyz, tri)
        ctx.set_params(0, number_of_clusters=2000, metric=acvd_amd.METRIC_ISOTROPIC)
    elif name == "cube400":
        xyz, tri = meshgen.cube(400, 2, 1e-4)
        ctx.set_input(xyz, tri)
        ctx.set_params(0, number_of_clusters=20000, metric=acvd_amd.METRIC_QEM, unconstrained_initialization=1)
    else:
        raise SystemExit(f"unknown workload {name}")
    ctx.build()
    ctx.initial_sampling()
    ctx.set_initial_labels(0, ctx.labels(0))
    return ctx, np


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", required=True, choices=["torus4096", "icosphere6", "cube400"])
    ap.add_argument("--repeats", type=int, default=2)
    a = ap.parse_args()
    ctx, _ = _workload(a.workload)
    before = ctx.get_tuning()
    e = tune(ctx, a.repeats, note=a.workload)
    print(json.dumps({"workload": a.workload, "class": size_class(ctx), "default": before, **e}), flush=True)


if __name__ == "__main__":
    main()
