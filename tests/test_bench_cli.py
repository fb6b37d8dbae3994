"""bench.py's command-line contract (CPU only: no device is touched on these paths)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_pass_byte_model_and_committed_traffic_agree():
    sys.path.insert(0, ROOT)
    import bench
    arr32 = (256 // 2 + 1) * 256 * 256 * 8
    assert sum(bench.pass_bytes(k, 256, 256, 256, 4) for k in
               ("rows_fused", "pencils_fwd", "pencil_update", "pencils_inv")) == 10 * arr32
    layout = bench.stored_layout("cahn_hilliard_3d_256", "fp32")
    assert layout is not None
    for k in ("rows_fused", "pencils_fwd", "pencil_update", "pencils_inv"):
        t, why = bench.pmc_traffic("cahn_hilliard_3d_256_fp32", k, layout)
        assert t is not None, why
        # measured HBM bytes within 10 % of the algorithmic bytes
        assert abs(t / bench.pass_bytes(k, 256, 256, 256, 4) - 1.0) < 0.10, k
    # counters measured under another launch layout are refused
    other = dict(layout, rows_fused=layout["rows_fused"] * 2)
    t, why = bench.pmc_traffic("cahn_hilliard_3d_256_fp32", "rows_fused", other)
    assert t is None and "layout" in why
