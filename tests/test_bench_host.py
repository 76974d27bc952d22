"""Host-side arithmetic of bench.py: the algorithmic work model and the
FER-matched comparison with the CPU baseline (no GPU needed)."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def test_work_model_matches_the_degree_table():
    # 239 lane-ops per degree-6 check and 86 per degree-3 check (18d - 1 + 11 min* evaluations)
    assert bench.ops_per_check(6) == 239 and bench.ops_per_check(3) == 86
    assert bench.ops_per_cwi(1024, "1/2") == bench.OPS_PER_CWI == 1024 * 239 + 512 * 86
    # every code: M checks of degree 3 plus 2M of the rate's degree
    assert bench.ops_per_cwi(16384, "4/5") == 16 * 128 * (86 + 2 * bench.ops_per_check(18))


def test_profile_r_work_model_on_the_chinn_out():
    """23 lane-ops per edge plus 13 per fold step, d - 1 fold steps per row of
    degree d: 203 for degree 6, 95 for degree 3, 256,512 for the .chinn.out's
    1024 degree-6 and 512 degree-3 rows."""
    from ldpc_mi355x import codes
    assert bench.ops_per_row_r(6) == 203 and bench.ops_per_row_r(3) == 95
    _, _, rows = codes.read_chinn_out(bench.SCALAR_MATRIX)
    degs = sorted(len(r) for r in rows)
    assert degs.count(3) == 512 and degs.count(6) == 1024 and len(degs) == 1536
    assert sum(bench.ops_per_row_r(len(r)) for r in rows) == 256_512


def test_fer_match_is_an_exact_two_sample_poisson_test():
    scipy = pytest.importorskip("scipy")  # noqa: F841
    same = bench.fer_match(57, 20_971_520, {"blocks": 368_640, "frame_errors": 1})
    assert same["consistent_95"] and same["p_value"] > 0.05
    far = bench.fer_match(57, 20_971_520, {"blocks": 368_640, "frame_errors": 30})
    assert not far["consistent_95"] and far["p_value"] < 1e-20
    assert bench.fer_match(0, 100, {"blocks": 10, "frame_errors": 0})["p_value"] == 1.0
    assert bench.fer_match(3, 100, None) is None
    assert bench.fer_match(3, 100, {"blocks": None, "frame_errors": None}) is None
    # symmetric in which side holds the errors when the sample sizes are equal
    a = bench.fer_match(10, 1000, {"blocks": 1000, "frame_errors": 2})["p_value"]
    b = bench.fer_match(2, 1000, {"blocks": 1000, "frame_errors": 10})["p_value"]
    assert abs(a - b) < 1e-12


def test_gpus_above_one_exit_with_a_message():
    import subprocess
    p = subprocess.run([sys.executable, bench.__file__, "--gpus", "2"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "one GPU only" in p.stderr and p.stdout == ""


def test_floor_comparisons_quoted_against_the_reference():
    """The error-floor comparisons the profiles quote, with the same exact test:
    34 frame errors in 4e10 (Profile T) against the silicon run's 61 in
    40,000,004,320 (README.md:120) -> p = 0.0073, and 17 in the first 2e10 ->
    p = 0.031; 24 in 1e10 (Profile R) against the scalar engine's 11 in
    4,667,520,000 (README.md:121) -> p = 1.0."""
    pytest.importorskip("scipy")
    t = bench.fer_match(34, 40_000_000_000, {"blocks": 40_000_004_320, "frame_errors": 61})
    assert round(t["p_value"], 4) == 0.0073 and t["consistent_95"] is False
    t1 = bench.fer_match(17, 20_000_000_000, {"blocks": 40_000_004_320, "frame_errors": 61})
    assert round(t1["p_value"], 3) == 0.031
    r = bench.fer_match(24, 10_000_000_000, {"blocks": 4_667_520_000, "frame_errors": 11})
    assert r["p_value"] > 0.99 and r["consistent_95"]


def test_floor_profile_option_is_validated():
    import subprocess
    p = subprocess.run([sys.executable, bench.__file__, "--workload", "floor", "--floor-profile", "X"],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "--floor-profile" in p.stderr and p.stdout == ""
