"""Parity at the plain configurations' own sizes (AR4JA r1/2 K=1024, 200
iterations): 1e5 codewords per point at 1.0 / 1.5 / 2.0 dB through the fused
Monte Carlo kernel, counters and the whole iteration histogram equal to the
oracle's; and a bit-exact decode of a 1e5-codeword LLR dump at 2.5 dB (final
LLRs, iterations, convergence)."""
import numpy as np
import pytest

from conftest import channel_params
from ldpc_mi355x import decoder

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
