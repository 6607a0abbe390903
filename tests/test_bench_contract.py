"""bench.py's host-side helpers (no GPU): the roofline's traffic comes from
the committed PMC summary only for the tile configuration it was measured
in, and the JSON line refuses more than one GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, ROOT)
    import bench as b

    return b


def test_traffic_keyed_by_tile_configuration(bench):
    lines = [l.split() for l in open(bench.TRAFFIC_FILE) if l.startswith("traffic ")]
    assert lines, "no traffic lines in the PMC summary"
    for p in lines:
        assert len(p) == 5 and p[3].startswith("cfg")
        kernel, prec, cfg = p[1], p[2], int(p[3][3:])
        assert bench.measured_traffic(kernel, prec, cfg) == float(p[4])
        assert bench.measured_traffic(kernel, prec, cfg + 17) is None
    prof = {"K1": (8.7, 1), "K2": (1.4, 1), "K3": (10.8, 1)}
    fl = {"K1": 4.8e12, "K2": 2e11, "K3": 4.8e12}
    k3cfgs = {int(p[3][3:]) for p in lines if p[1] == "K3"}
    k3cfg = min(k3cfgs)
    r = bench.roofline_from_profile(prof, fl, "f16", True, [0, k3cfg])
    assert r["kernel"] == "K3" and r["traffic"] == bench.measured_traffic("K3", "f16", k3cfg)
    assert abs(r["frac"] - 4.8e12 / 10.8e-3 / bench.PEAK["f16"]) < 1e-12
    unmeasured = next(c for c in range(64) if c not in k3cfgs)
    assert bench.roofline_from_profile(prof, fl, "f16", True, [0, unmeasured])["traffic"] is None
    assert bench.roofline_from_profile(prof, fl, "f16", False, [0, k3cfg])["traffic"] is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)
