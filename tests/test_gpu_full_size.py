"""Parity at the plain configurations' own sizes (AR4JA r1/2 K=1024, 200
iterations): 1e5 codewords per point at 1.0 / 1.5 / 2.0 dB through the fused
Monte Carlo kernel, counters and the whole iteration histogram equal to the
oracle's; a bit-exact decode of a 1e5-codeword LLR dump at 2.5 dB (final
LLRs, iterations, convergence); and the same 1e5-codeword Monte Carlo check
for Profile R (the scalar_5w arithmetic run_simulation.py drives) on the
reference's .chinn.out; and the large codes of the nine-code batch at their
batch Eb/N0 and 200 iterations."""
import os

import numpy as np
import pytest

from conftest import channel_params
from ldpc_mi355x import codes, decoder

pytestmark = pytest.mark.gpu

N_CW = 100_000


@pytest.mark.parametrize("ebn0", [1.0, 1.5, 2.0])
def test_monte_carlo_1e5_codewords_matches_oracle(oracle, code12, ebn0):
    mu, sg = channel_params(ebn0)
    print(f"1e5-codeword Monte Carlo at {ebn0} dB", flush=True)
    with decoder.Engine(1024, "1/2", max_iter=200) as eng:
        got = eng.run_point(ebn0, N_CW, seed=0x5EED2026, chunk_cw=1 << 20)
    exp = oracle.run_point_t(code12, 0x5EED2026, 0, N_CW, mu, sg, 200, 0.9375)
    assert (got.blocks, got.frame_errors, got.bit_errors, got.iter_sum) == \
        (exp["blocks"], exp["frame_errors"], exp["bit_errors"], exp["iter_sum"])
    np.testing.assert_array_equal(got.hist, exp["hist"])
    print(f"  FER {got.fer:.3e}  BER {got.ber:.3e}  mean iterations {got.mean_iterations:.3f}", flush=True)


def test_decode_1e5_codeword_dump_at_2p5db(oracle, code12):
    mu, sg = channel_params(2.5)
    with decoder.Engine(1024, "1/2", max_iter=200) as eng:
        llr = eng.gen_channel(0xD0D0, 0, N_CW, 2.5)
        np.testing.assert_array_equal(llr[:64], oracle.gen_channel(code12, 0xD0D0, 0, 64, mu, sg))
        out, it, ok = eng.decode_llrs(llr)
    e_out, e_it, e_ok = oracle.decode_t(code12, llr, 200, 0.9375)
    np.testing.assert_array_equal(ok, e_ok)
    np.testing.assert_array_equal(it, e_it)
    np.testing.assert_array_equal(out, e_out)


@pytest.mark.parametrize("ebn0", [1.5, 2.0])
def test_profile_r_monte_carlo_1e5_codewords_matches_oracle(oracle, ebn0):
    _, n, rows = codes.read_chinn_out(os.path.join(os.path.dirname(__file__), "golden",
                                                   "AR4JA_r45_4c_128c_r12.chinn.out"))
    print(f"Profile R 1e5-codeword Monte Carlo at {ebn0} dB", flush=True)
    with decoder.Engine(profile="R", matrix=(rows, n), n_punct=512, max_iter=200) as eng:
        got = eng.run_point(ebn0, N_CW, seed=0x5EED2026, chunk_cw=1 << 20)
        mu, sg = eng.channel_params(ebn0)
    exp = oracle.run_point_r(n, rows, 2048, 0x5EED2026, 0, N_CW, mu, sg, 0.0, 200)
    assert (got.blocks, got.frame_errors, got.bit_errors, got.iter_sum) == \
        (exp["blocks"], exp["frame_errors"], exp["bit_errors"], exp["iter_sum"])
    np.testing.assert_array_equal(got.hist, exp["hist"])
    print(f"  FER {got.fer:.3e}  BER {got.ber:.3e}  mean iterations {got.mean_iterations:.3f}", flush=True)


RATE = {"1/2": 0.5, "2/3": 2.0 / 3.0, "4/5": 0.8}


@pytest.mark.parametrize("k,rate,ebn0,n", [(16384, "1/2", 1.2, 2048), (16384, "2/3", 2.0, 2048),
                                           (16384, "4/5", 3.0, 2048), (4096, "4/5", 3.0, 8192)])
def test_large_codes_at_their_batch_points_match_oracle(oracle, k, rate, ebn0, n):
    """The nine-code batch's critical-path codes at their own Eb/N0 and 200
    iterations (library-default kernels): counters and histogram bit for bit."""
    mu, sg = channel_params(ebn0, RATE[rate])
    print(f"K={k} r{rate} {n} codewords at {ebn0} dB, 200 iterations", flush=True)
    with decoder.Engine(k, rate, max_iter=200) as eng:
        got = eng.run_point(ebn0, n, seed=0x5EED2026, first_cw=11, chunk_cw=n)
    exp = oracle.run_point_t(codes.ar4ja(k, rate), 0x5EED2026, 11, n, mu, sg, 200, 0.9375)
    assert (got.blocks, got.frame_errors, got.bit_errors, got.iter_sum) == \
        (exp["blocks"], exp["frame_errors"], exp["bit_errors"], exp["iter_sum"])
    np.testing.assert_array_equal(got.hist, exp["hist"])
    print(f"  mean iterations {got.mean_iterations:.3f}", flush=True)
