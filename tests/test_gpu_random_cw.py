"""Random-info-bit codewords on the GPU (ldpc_set_codeword_mode), against the
oracle bit for bit, and against the all-zero codeword statistically.

The kernel encodes each codeword itself (quasi-cyclic generator, 128-bit
blocks) before it draws the channel; the oracle encodes with a dense
generator from plain Gauss-Jordan elimination.  Equal channel LLRs therefore
check the device encoder, and equal Monte Carlo counters check the tally
against the sent bits.
"""
import numpy as np
import pytest

from conftest import channel_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng_rand():
    from ldpc_mi355x import decoder
    eng = decoder.Engine(1024, "1/2", max_iter=200, beta=0.9375)
    eng.set_codeword_mode(True)
    yield eng
    eng.close()


@pytest.mark.parametrize("seed,first,ebn0", [(0x5EED2026, 0, 1.0), (77, 123456789012, 2.5)])
def test_random_channel_bit_exact(eng_rand, oracle, code12, seed, first, ebn0):
    print(f"random-codeword channel seed={seed} first={first} {ebn0} dB", flush=True)
    got = eng_rand.gen_channel(seed, first, 301, ebn0)
    mu, sg = channel_params(ebn0)
    exp = oracle.gen_channel_cw(code12, seed, first, 301, mu, sg)
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("ebn0,n", [(1.0, 3000), (1.5, 6000), (2.0, 6000)])
def test_random_monte_carlo_counters_bit_exact(eng_rand, oracle, code12, ebn0, n):
    print(f"random-codeword Monte Carlo {ebn0} dB n={n}", flush=True)
    r = eng_rand.run_point(ebn0, n, seed=0x5EED2026, chunk_cw=2048)
    mu, sg = channel_params(ebn0)
    o = oracle.run_point_t_cw(code12, 0x5EED2026, 0, n, mu, sg, 200, 0.9375)
    assert (r.blocks, r.frame_errors, r.bit_errors, r.iter_sum) == \
        (o["blocks"], o["frame_errors"], o["bit_errors"], o["iter_sum"])
    np.testing.assert_array_equal(r.hist, o["hist"])


def test_random_decode_of_dumps_bit_exact(eng_rand, oracle, code12):
    """The decode of random-codeword channel LLRs (ldpc_decode_llrs) equals
    the oracle's, final LLRs and iterations."""
    mu, sg = channel_params(1.25)
    llr = oracle.gen_channel_cw(code12, 9, 0, 200, mu, sg)
    out, it, ok = eng_rand.decode_llrs(llr)
    o_out, o_it, o_ok = oracle.decode_t(code12, llr, 200, 0.9375)
    np.testing.assert_array_equal(out, o_out)
    np.testing.assert_array_equal(it, o_it)


def test_random_fer_matches_all_zero_at_1p5db(oracle):
    """2^20 codewords of each mode at 1.5 dB on the same ids: the frame error
    counts agree (exact two-sample test), and mean iterations within 2%."""
    from scipy.stats import binomtest

    from ldpc_mi355x import decoder
    n = 1 << 20
    with decoder.Engine(1024, "1/2", max_iter=200, beta=0.9375) as eng:
        z = eng.run_point(1.5, n, seed=0xC0DE)
        eng.set_codeword_mode(True)
        r = eng.run_point(1.5, n, seed=0xC0DE)
    p = binomtest(r.frame_errors, r.frame_errors + z.frame_errors, 0.5).pvalue
    print("random FE", r.frame_errors, "all-zero FE", z.frame_errors, "p", p,
          "mean it", r.mean_iterations, z.mean_iterations, "Gsps", r.kernel_gsps, z.kernel_gsps, flush=True)
    assert r.frame_errors > 1000 and z.frame_errors > 1000
    assert p > 1e-3
    assert abs(r.mean_iterations - z.mean_iterations) / z.mean_iterations < 0.02


def test_codeword_mode_only_where_built():
    from ldpc_mi355x import _lib as L
    from ldpc_mi355x import decoder
    with decoder.Engine(4096, "1/2", max_iter=20) as eng:
        with pytest.raises(L.LdpcError):
            eng.set_codeword_mode(True)
        eng.set_codeword_mode(False)  # the all-zero codeword is always available
