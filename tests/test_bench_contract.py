"""bench.py's host-side helpers (no GPU): the QP-solve plan of the driver's
`--steps 20 --warmup 5` stays inside the wall budget, the roofline's traffic
comes from a PMC summary only for the kernel, precision and tile
configuration it was measured in (scaled by vectors per launch), and the
JSON line refuses more than one GPU."""
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


@pytest.mark.parametrize("before_s, step_s_8", [(120.0, 1.15), (200.0, 1.15), (300.0, 2.0), (60.0, 0.4),
                                                (400.0, 1.15)])
def test_driver_arguments_fit_the_budget(bench, before_s, step_s_8):
    """--steps 20 --warmup 5 on the c3 workload (8 QP states, niter 300): the
    plan times at most the 8 distinct states, at least one, and the projected
    end of the timed region stays within the default budget whenever one solve
    fits at all."""
    niter, n_states, budget = 300, 8, bench.DEFAULT_BUDGET_S
    warm = 5 * step_s_8
    remaining = budget - (before_s + warm) - 15.0
    k, reason = bench.plan_solves(20, n_states, niter, step_s_8, n_states, remaining)
    assert 1 <= k <= n_states
    per_solve = niter * step_s_8 / n_states
    projected = before_s + warm + k * per_solve
    if remaining >= per_solve * 1.1:
        assert projected <= budget, (k, reason, projected)
    else:
        assert k == 1
    assert budget < bench.DRIVER_LIMIT_S
    if k < n_states:
        assert "budget" in reason
    else:
        assert "8 QP states" in reason


def test_plan_keeps_requests_below_the_state_count(bench):
    assert bench.plan_solves(2, 8, 300, 0.3, 2, 1e9) == (2, "requested")
    assert bench.plan_solves(20, 8, 300, 0.3, 8, 1e9)[0] == 8


def test_gap_order(bench):
    assert bench.gap_order(100, 8) == [99, 100, 98, 101, 97, 102, 96, 103]


def test_traffic_keyed_by_kernel_precision_tiles_and_vectors(bench, tmp_path):
    f = tmp_path / "pmc_traffic_test.md"
    f.write_text("header\n```\ntraffic K3 f16 cfg2 nvec3200 5.0e10\ntraffic K3 f16 cfg2 nvec12800 1.9e11\n"
                 "traffic K1 f16 cfg0 nvec3200 9.7e9\ntraffic K3 bf16 cfg2 nvec3200 6.0e10\n```\n")
    files = [str(f)]
    v, src = bench.measured_traffic("K3", "f16", 2, 3200, files)
    assert v == 5.0e10 and src.startswith("measured")
    v, src = bench.measured_traffic("K3", "f16", 2, 12800, files)
    assert v == 1.9e11 and src.startswith("measured")
    v, src = bench.measured_traffic("K3", "f16", 2, 4800, files)
    assert abs(v - 5.0e10 * 1.5) < 1.0 and src.startswith("scaled")
    v, src = bench.measured_traffic("K3", "f16", 2, 11200, files)   # nearest is 12800
    assert abs(v - 1.9e11 * 11200 / 12800) < 1.0
    assert bench.measured_traffic("K3", "f16", 3, 3200, files) == (None, None)
    assert bench.measured_traffic("K3", "bf16x3", 2, 3200, files) == (None, None)
    assert bench.measured_traffic("K2", "f16", 0, 3200, files) == (None, None)


def test_committed_traffic_lines_parse(bench):
    for path in bench.traffic_files():
        for line in open(path):
            if line.startswith("traffic "):
                p = line.split()
                assert len(p) == 6 and p[3].startswith("cfg") and p[4].startswith("nvec"), line
                v, src = bench.measured_traffic(p[1], p[2], int(p[3][3:]), int(p[4][4:]))
                assert v == float(p[5]) and src.startswith("measured")


def test_roofline_fraction(bench):
    prof = {"K1": (8.7, 1), "K2": (1.4, 1), "K3": (10.8, 1)}
    fl = {"K1": 4.8e12, "K2": 2e11, "K3": 4.8e12}
    r = bench.roofline_from_profile(prof, fl, "f16", [0, 2], 3200)
    assert r["kernel"] == "K3" and r["tile_config"] == 2 and r["vectors_per_launch"] == 3200
    assert abs(r["frac"] - 4.8e12 / 10.8e-3 / bench.PEAK["f16"]) < 1e-12
    assert bench.roofline_from_profile(prof, fl, "f16", None, 3200)["traffic"] is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)
