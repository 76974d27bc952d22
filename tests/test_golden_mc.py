"""The oracle's Monte Carlo accounting against the reference golden's own
Monte Carlo run.

tests/golden/golden_mc.json holds what the reference stepper printed for
`--sfpi --monte-carlo <ebn0> 1000 1000000000 200 1` (stepper:708-856; one
host thread, so the run is deterministic): total blocks, frame errors, bit
errors and the iteration sum.  golden_mc_channel.cpp rebuilds its channel
draws (mt19937_64 seeded 1337, std::normal_distribution<float>, LLR =
bf16((2/sigma^2)(1 + n)), stepper:720-757) with the C++ standard library, the
oracle decodes them with the stepper's Monte Carlo damping (beta = 1,
stepper:708, 1029) and tallies them with the rule its own Monte Carlo uses
(oracle/ldpc_oracle.c tally_t), and all four numbers must match exactly.

This pins the tally rule every GPU Monte Carlo test is compared against:
frame error <=> still active after max_iter, bit errors = info bits k < 1024
with bf16 <= 0.0f of those codewords, iterations = convergence iteration or
max_iter.  A '< 0' for '<= 0', a count over the 2048 transmitted symbols, or
an off-by-one in the iteration count each changes a number (checked below by
planting each mistake in a numpy re-tally of the oracle's decode).  The
1.0 dB run ends 59 info LLRs of its failed codewords at exactly +0, and one
converged codeword holds a +0 info LLR, so even the '<' / '<=' choice and the
"failed codewords only" rule are pinned.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN_DIR

GOLDEN = json.load(open(os.path.join(GOLDEN_DIR, "golden_mc.json")))
N = 2560


@pytest.fixture(scope="module")
def channel_tool(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ is needed to build the golden channel generator")
    exe = str(tmp_path_factory.mktemp("gmc") / "golden_mc_channel")
    subprocess.run([gxx, "-O2", "-std=c++17", "-o", exe, os.path.join(GOLDEN_DIR, "golden_mc_channel.cpp")],
                   check=True)
    return exe


def golden_llrs(exe, ebn0, n_cw, tmp_path):
    out = tmp_path / f"llr_{ebn0}.bin"
    subprocess.run([exe, f"{ebn0:.2f}", str(n_cw), str(out)], check=True)
    return np.fromfile(out, dtype="<u2").reshape(n_cw, N)


@pytest.mark.parametrize("point", GOLDEN["points"], ids=lambda p: f"{p['ebn0_db']}dB")
def test_oracle_tally_reproduces_reference_monte_carlo(point, channel_tool, code12, oracle, tmp_path):
    llr = golden_llrs(channel_tool, point["ebn0_db"], point["blocks"], tmp_path)
    assert np.all(llr[:, 2048:] == 0)
    r = oracle.run_llrs_t(code12, llr, point["max_iter"], point["beta"])
    got = (r["blocks"], r["frame_errors"], r["bit_errors"], r["iter_sum"])
    exp = (point["blocks"], point["frame_errors"], point["bit_errors"], point["iter_sum"])
    assert got == exp
    assert int(r["hist"].sum()) == point["blocks"] and int(r["hist"][0]) == point["frame_errors"]


def test_planted_tally_mistakes_change_the_totals(channel_tool, code12, oracle, tmp_path):
    """The 1.0 dB run has frame errors, so each plausible slip in the rule
    moves a number away from the reference's (the slips are planted in a
    numpy re-tally of the oracle's decode, on the CPU)."""
    point = next(p for p in GOLDEN["points"] if p["frame_errors"] > 0)
    llr = golden_llrs(channel_tool, point["ebn0_db"], point["blocks"], tmp_path)
    out, iters, ok = oracle.decode_t(code12, llr, point["max_iter"], point["beta"])
    failed = ok == 0
    info = out[failed, :1024]
    nonpos = ((info & 0x8000) != 0) | (info == 0)
    assert int(nonpos.sum()) == point["bit_errors"]
    assert int(iters.sum()) == point["iter_sum"] and int(failed.sum()) == point["frame_errors"]

    def count_nonpos(a):
        return int((((a & 0x8000) != 0) | (a == 0)).sum())

    assert count_nonpos(out[failed, :2048]) != point["bit_errors"]   # over the 2048 symbols, not the 1024 info bits
    assert count_nonpos(out[:, :1024]) != point["bit_errors"]        # converged codewords counted too
    assert int((info > 0x8000).sum()) != point["bit_errors"]         # '< 0.0f' instead of '<= 0.0f'
    assert int((iters - 1).sum()) != point["iter_sum"]               # iteration index instead of count
    assert int((iters * ok).sum()) != point["iter_sum"]              # failures counted as 0 iterations


@pytest.mark.gpu
@pytest.mark.parametrize("point", GOLDEN["points"], ids=lambda p: f"{p['ebn0_db']}dB")
def test_gpu_decodes_reference_monte_carlo_draws(point, channel_tool, code12, oracle, tmp_path):
    """The HIP decoder on the reference run's own channel draws gives the
    oracle's final LLRs, iterations and convergence flags bit for bit, so its
    tallies are the reference's totals."""
    from ldpc_mi355x import decoder
    llr = golden_llrs(channel_tool, point["ebn0_db"], point["blocks"], tmp_path)
    print(f"GPU decode of the reference's {point['ebn0_db']} dB draws", flush=True)
    with decoder.Engine(1024, "1/2", max_iter=point["max_iter"], beta=point["beta"]) as eng:
        out, iters, ok = eng.decode_llrs(llr)
    o_out, o_it, o_ok = oracle.decode_t(code12, llr, point["max_iter"], point["beta"])
    assert np.array_equal(out, o_out) and np.array_equal(iters, o_it) and np.array_equal(ok, o_ok)
    assert int(iters.sum()) == point["iter_sum"] and int((ok == 0).sum()) == point["frame_errors"]
