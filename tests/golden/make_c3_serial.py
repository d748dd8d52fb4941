"""Writes tests/golden/c3_serial.json: the serial reference restatement
(oracle A, oracle/serial.h) run to convergence on the 8,388,608-vertex QEM
torus (4096 x 2048 grid, seed 3, |H|^1.5 weights, K = 100,000, unconstrained
initialisation) from the reference's initial sampling filled by the
reference's FillHolesInClustering.  The run takes about two minutes on one
core, too long for a test; the GPU test compares against the energy stored
here after checking that its initial labels hash to the same value.

Run from the repository root:  python tests/golden/make_c3_serial.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from acvd_amd import meshgen  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

NU, NV, K, GRAD = 4096, 2048, 100000, 1.5


def c3_inputs():
    """Mesh, indicator and the filled initial labels of the C3 workload."""
    xyz, tri = meshgen.torus(NU, NV, seed=3)
    H = meshgen.mean_curvature(xyz, tri)
    m = po.RefMesh(xyz, tri)
    init = m.fill_holes(K, m.initial_sampling(po.QEM, K, GRAD, H))
    return xyz, tri, H, m, init


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.int32).tobytes()).hexdigest()


def main():
    t = time.time()
    xyz, tri, H, m, init = c3_inputs()
    print(f"inputs {time.time() - t:.1f} s", flush=True)
    po.set_progress(True)
    lab, st = m.serial_run(po.QEM, K, GRAD, H, 3, 1, init)
    e = m.energy(po.QEM, K, lab, GRAD, H)
    out = {"workload": f"torus {NU}x{NV} seed 3, QEM, gradation {GRAD} on |H|, K={K}, unconstrained initialisation",
           "init_sha256": digest(init), "labels_sha256": digest(lab), "energy": e[0], "shifted_energy": e[1],
           "loops": st["loops"], "convergences": st["convergences"], "modifications": st["total_moves"],
           "serial_seconds": st["seconds"]}
    with open(os.path.join(ROOT, "tests", "golden", "c3_serial.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
