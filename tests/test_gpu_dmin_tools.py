"""The learned decoder's training and evaluation tools, end to end on tiny sizes:
train_dmin.py writes parameters; a second run resumes from them
(--init-from, checked with lr 0 so the output must equal the input) and a
parameter file of another shape is refused; eval_dmin.py evaluates them through
Profile D with early exit, and through the training forward at an iteration
count other than the trained one (--forward-T)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOLS = os.path.join(ROOT, "tt-ldpc-qb2-sim_amd")
TINY = ["--batch", "256", "--micro-batch", "256", "--val-ebn0", "2.0", "--val-size", "256", "--correction"]


def run(args, ok=True):
    print("[gpu-case]", " ".join(os.path.basename(a) for a in args[:1]), " ".join(args[1:]), flush=True)
    p = subprocess.run([sys.executable, *args], cwd=ROOT, capture_output=True, text=True, timeout=100)
    if ok:
        assert p.returncode == 0, p.stderr[-3000:]
    return p, [json.loads(line) for line in p.stdout.splitlines() if line.startswith("{")]


@pytest.fixture(scope="module")
def trained(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("dmin") / "a.npz")
    _, lines = run([os.path.join(TOOLS, "train_dmin.py"), "--steps", "2", "--val-every", "2", "--lr", "1e-2", *TINY,
                    "--out", out])
    assert lines[-1]["saved"] == out and [ln["step"] for ln in lines[:-1]] == [0, 2]
    return out


def test_train_writes_parameters(trained):
    z = np.load(trained, allow_pickle=False)
    assert z["alpha"].shape == (1, 60) and z["beta"].shape == (1, 60)
    assert bool(z["correction"]) and int(z["T"]) == 20 and int(z["k_info"]) == 1024 and str(z["rate"]) == "1/2"
    assert np.all(np.isfinite(z["alpha"])) and not np.all(z["alpha"] == 1.0)  # two Adam steps moved them


def test_init_from_resumes_exactly(trained, tmp_path):
    out = str(tmp_path / "b.npz")
    run([os.path.join(TOOLS, "train_dmin.py"), "--steps", "1", "--val-every", "1", "--lr", "0", *TINY,
         "--init-from", trained, "--out", out])
    a, b = np.load(trained, allow_pickle=False), np.load(out, allow_pickle=False)
    np.testing.assert_array_equal(a["alpha"], b["alpha"])
    np.testing.assert_array_equal(a["beta"], b["beta"])


def test_init_from_refuses_another_shape(trained, tmp_path):
    p, _ = run([os.path.join(TOOLS, "train_dmin.py"), "--steps", "1", "--per-iteration", *TINY, "--init-from",
                trained, "--out", str(tmp_path / "c.npz")], ok=False)
    assert p.returncode != 0 and "does not match" in p.stderr


def test_eval_profile_d_and_forward(trained):
    _, lines = run([os.path.join(TOOLS, "eval_dmin.py"), trained, "--ebn0", "2.0", "--blocks", "1024", "--iters",
                    "5", "--ref-iters", "20"])
    rows = {ln["decoder"]: ln for ln in lines if "decoder" in ln}
    assert set(rows) == {"approx_min_star_20", "trained_T5"} and "summary" in lines[-1]
    for r in rows.values():
        assert r["blocks"] == 1024 and 0 <= r["frame_errors"] <= 1024 and r["fer_ci95"][0] <= r["fer"]
    assert 1.0 <= rows["trained_T5"]["mean_iterations"] <= 5.0
    _, lines = run([os.path.join(TOOLS, "eval_dmin.py"), trained, "--forward", "--forward-T", "7", "--ebn0", "2.0",
                    "--blocks", "512", "--ref-iters", "20"])
    rows = {ln["decoder"]: ln for ln in lines if "decoder" in ln}
    assert set(rows) == {"trained_forward_T7", "approx_min_star_20"}
    assert rows["trained_forward_T7"]["max_iter"] == 7 and rows["trained_forward_T7"]["blocks"] == 512


def test_profile_tool_reports_hbm_rooflines():
    _, lines = run([os.path.join(TOOLS, "profile_dmin.py"), "--T", "3", "--batch", "512", "--ckpt-every", "5",
                    "--reps", "1"])
    r = lines[-1]
    assert r["layer_steps"] == 36 and r["batch"] == 512
    for phase in ("forward", "forward_with_checkpoints", "backward"):
        assert r[phase]["seconds"] > 0 and 0 < r[phase]["frac_hbm"] < 1.5
