"""bench.py's command line on the CPU: more than one GPU is refused with a
message (the build runs on one GPU), and the byte models it divides by the
measured times are the documented per-unit formulas."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_byte_models():
    import bench
    N, E2, K = 1000, 6000, 50
    # scan: 12 B per item, 4 B per CSR entry and per cluster stamp, 4 B per work item
    assert bench.scan_bytes_per_launch(N, E2, K, 100) == 12.0 * N + 4.0 * E2 + 4.0 * K + 400.0
    # evaluate: per work item 72 B + 4 B per CSR entry, 156 B per touched cluster, 28 B per proposal
    assert bench.eval_bytes_per_launch(E2, N, K, 10, 3) == 10 * (72.0 + 24.0) + 156.0 * 30 + 28.0 * 3
    st = {"ctx_walked": 7, "ctx_proposed": 5, "ctx_reeval_live": 3, "ctx_reevaluated": 2, "ctx_rechecked": 4,
          "ctx_accepted": 1, "ctx_accepted_constrained": 1}
    want = (4.0 * 7 + 24.0 * (5 + 3) + 36.0 * 3 + (72.0 + 48.0 + 468.0) * 2 + (88.0 + 48.0) * 4 + 252.0 + 480.0)
    assert bench.accept_bytes(st, N, E2) == want
