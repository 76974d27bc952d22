"""Trapping-set capture: the ids a Monte Carlo run records are exactly its
failing codewords (checked against the oracle on the same channel), and the
rebuilt Packed4CW dumps replay to the same final state."""
import os
import subprocess

import numpy as np
import pytest

from conftest import channel_params
from ldpc_mi355x import codes, decoder, dumps

pytestmark = pytest.mark.gpu


def failing_ids_t(oracle, code, seed, first, n, ebn0, max_iter):
    mu, sg = channel_params(ebn0, code.k / code.n_tx)
    llr = oracle.gen_channel(code, seed, first, n, mu, sg)
    _, _, ok = oracle.decode_t(code, llr, max_iter)
    return first + np.nonzero(ok == 0)[0]


@pytest.mark.parametrize("k,rate,variant", [(1024, "1/2", 0), (1024, "2/3", 0), (1024, "1/2", 0x101)])
def test_profile_t_captures_exactly_the_failures(oracle, k, rate, variant):
    code = codes.ar4ja(k, rate)
    ebn0 = {"1/2": 1.2, "2/3": 2.0}[rate]
    with decoder.Engine(k, rate, max_iter=30, variant=variant) as eng:
        eng.set_capture(4096)
        r = eng.run_point(ebn0, 1500, seed=77, first_cw=1000, chunk_cw=700)
        ids, seen = eng.captures()
    exp = failing_ids_t(oracle, code, 77, 1000, 1500, ebn0, 30)
    assert seen == r.frame_errors == len(exp) > 0
    np.testing.assert_array_equal(ids, exp)


def test_capture_keeps_at_most_max_frames():
    with decoder.Engine(1024, "1/2", max_iter=20) as eng:
        eng.set_capture(5)
        r = eng.run_point(1.0, 2000, seed=3)
        ids, seen = eng.captures()
        assert seen == r.frame_errors > 5 and len(ids) == 5 and np.all(np.diff(ids.astype(np.int64)) > 0)
        eng.set_capture(0)
        ids0, seen0 = eng.captures()
        assert len(ids0) == 0 and seen0 == 0


def test_profile_r_and_d_capture(oracle):
    import dmin_oracle as DO
    m, n, rows = codes.read_chinn_out(os.path.join(os.path.dirname(__file__), "golden",
                                                   "AR4JA_r45_4c_128c_r12.chinn.out"))
    mu, sg = channel_params(1.3)
    with decoder.Engine(profile="R", matrix=(rows, n), n_punct=512, max_iter=12) as eng:
        eng.set_capture(4096)
        r = eng.run_point(1.3, 640, seed=9)
        ids, seen = eng.captures()
    llr = oracle.gen_channel_clamped(n, 2048, 9, 0, 640, mu, sg)
    out, _, _ = oracle.decode_r(n, rows, llr, 12)
    exp = np.nonzero(((out[:, :2048] >> 15) & 1).sum(1) > 0)[0]
    assert seen == r.frame_errors == len(exp) > 0
    np.testing.assert_array_equal(ids, exp)
    code = codes.ar4ja(1024, "1/2")
    P = DO.n_params(code)
    a, b = np.full((1, P), 0.8, np.float32), np.zeros((1, P), np.float32)
    with decoder.Engine(1024, "1/2", max_iter=15, profile="D") as eng:
        eng.set_dmin_params(a, b)
        eng.set_capture(4096)
        r = eng.run_point(1.3, 700, seed=11)
        ids, seen = eng.captures()
    _, _, ok = DO.decode_d(code, oracle.gen_channel(code, 11, 0, 700, mu, sg), a, b, 15)
    exp = np.nonzero(ok == 0)[0]
    assert seen == r.frame_errors == len(exp) > 0
    np.testing.assert_array_equal(ids, exp)


def test_profile_d_k16384_capture():
    """The K=16384 Profile D kernel records the ids of its failing codewords:
    over 320 codewords (more than one per workgroup) they equal those that
    decode_llrs leaves unconverged on the same channel (decode arithmetic the
    Profile D tests pin to the oracle)."""
    import dmin_oracle as DO
    code = codes.ar4ja(16384, "1/2")
    P = DO.n_params(code)
    a, b = np.full((1, P), 0.85, np.float32), np.full((1, P), 0.05, np.float32)
    print("[gpu-case] Profile D K=16384 capture, 320 codewords at 1.0 dB", flush=True)
    with decoder.Engine(16384, "1/2", max_iter=20, profile="D") as eng:
        eng.set_dmin_params(a, b, True)
        eng.set_capture(4096)
        r = eng.run_point(1.0, 320, seed=13, chunk_cw=320)
        ids, seen = eng.captures()
        _, _, ok = eng.decode_llrs(eng.gen_channel(13, 0, 320, 1.0))
    exp = np.nonzero(ok == 0)[0]
    assert seen == r.frame_errors == len(exp) > 0
    np.testing.assert_array_equal(ids, exp)


def test_trap_dumps_replay(tmp_path):
    with decoder.Engine(1024, "1/2", max_iter=20) as eng:
        eng.set_capture(16)
        eng.run_point(1.0, 800, seed=5)
        ids, _ = eng.captures()
        rep = dumps.trap_dumps(eng, 5, 1.0, int(ids[0]), str(tmp_path))
        assert rep["active_mask"] & (1 << rep["failing_cw"])
        init = dumps.read_packed4(str(tmp_path / dumps.INITIAL))
        fin = dumps.read_packed4(str(tmp_path / dumps.FINAL))
        np.testing.assert_array_equal(init, eng.gen_channel(5, 4 * rep["group"], 4, 1.0))
        np.testing.assert_array_equal(fin, dumps.replay(eng, init)["final"])


def test_cli_trap_writes_dumps(tmp_path):
    from ldpc_mi355x import build as B
    exe = os.path.join(os.path.dirname(B.LIB), "ldpc_sim_circulant")
    p = subprocess.run([exe, "1.0", "4096", "20", "--single-core", "--chunk-batches", "256", "--seed", "5", "--trap"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "FIRST BLOCK ERROR TRAPPED" in p.stdout
    for f in (dumps.INITIAL, dumps.FINAL):
        assert os.path.getsize(tmp_path / f) == 20480
    init = dumps.read_packed4(str(tmp_path / dumps.INITIAL))
    with decoder.Engine(1024, "1/2", max_iter=20) as eng:
        np.testing.assert_array_equal(dumps.read_packed4(str(tmp_path / dumps.FINAL)), dumps.replay(eng, init)["final"])


def test_trap_report_cli(tmp_path, capsys):
    """python -m ldpc_mi355x.dumps on captured ids: dumps written, the failure
    reproduced (not converged, 20 iterations) and its (a, b) signature equal to
    the one computed from the replayed final state."""
    with decoder.Engine(1024, "1/2", max_iter=20) as eng:
        eng.set_capture(4)
        eng.run_point(1.0, 400, seed=6)
        ids, _ = eng.captures()
        assert len(ids) >= 2
        rep = dumps.trap_dumps(eng, 6, 1.0, int(ids[1]), str(tmp_path / "direct"))
    exp = dumps.trapping_set(codes.ar4ja(1024, "1/2"), rep["final"][rep["failing_cw"]])
    assert dumps.main(["--seed", "6", "--ebn0", "1.0", "--max-iter", "20", "--id", str(ids[0]), "--id", str(ids[1]),
                       "--out", str(tmp_path / "cli")]) == 0
    import json
    lines = [json.loads(x) for x in capsys.readouterr().out.splitlines() if x.startswith("{")]
    assert [ln["id"] for ln in lines] == [int(ids[0]), int(ids[1])]
    for ln in lines:
        assert not ln["converged"] and ln["iterations"] == 20 and ln["a"] >= 1 and ln["b"] >= 1
        assert os.path.getsize(os.path.join(ln["dir"], dumps.FINAL)) == 20480
    assert (lines[1]["a"], lines[1]["b"], lines[1]["info_bits_in_error"]) == (exp["a"], exp["b"], exp["info_bits"])
    assert lines[1]["bit_errors"] == rep["bit_errors"][rep["failing_cw"]]
