"""Profile D (the learned layered min-sum with a syndrome early exit,
decode_multi.hip) against the oracle, bit for bit: final LLRs (bf16),
iterations and convergence of channel dumps; Monte Carlo counters and
histograms; the channel is the Profile T stream.  K=16384 r1/2 has its own
kernel (L in all of the LDS; messages rebuilt from three words per check,
half of the layers' words in registers and half in an HBM workspace), so
its cases cover per-iteration parameter rows (the rebuild uses the previous
iteration's row), the correction and the cap."""
import numpy as np
import pytest

import dmin_oracle as DO
from conftest import channel_params
from ldpc_mi355x import codes, decoder

pytestmark = pytest.mark.gpu

CASES = [(1024, "1/2", 1.6), (1024, "4/5", 3.0), (4096, "2/3", 2.0), (1024, "2/3", 2.4), (16384, "1/2", 0.9)]


def params(code, G, seed):
    rng = np.random.default_rng(seed)
    P = DO.n_params(code)
    return rng.uniform(0.7, 0.95, (G, P)).astype(np.float32), rng.uniform(0.0, 0.4, (G, P)).astype(np.float32)


# corr: 0 / False min-sum, 1 / True two-minimum correction, 2 three-minimum correction
DECODE_CASES = [(c, m) for c in CASES for m in [(1, False, 0.0), (8, True, 0.0), (1, False, 6.0), (8, True, 6.0),
                                                (8, 2, 0.0), (1, 2, 6.0)]]


@pytest.mark.parametrize("case,mode", DECODE_CASES)
def test_decode_matches_oracle(oracle, case, mode):
    (k, rate, ebn0), (G, corr, r_max) = case, mode
    code = codes.ar4ja(k, rate)
    mu, sg = channel_params(ebn0, code.k / code.n_tx)
    n = {1024: 128, 4096: 24, 16384: 6}[k]
    llr = oracle.gen_channel(code, 5, 0, n, mu, sg)
    a, b = params(code, G, k + G)
    print(f"Profile D decode K={k} rate={rate} G={G} corr={corr} r_max={r_max}", flush=True)
    with decoder.Engine(k, rate, max_iter=25, profile="D") as eng:
        eng.set_dmin_params(a, b, corr, r_max)
        out, it, ok = eng.decode_llrs(llr)
    e_out, e_it, e_ok = DO.decode_d(code, llr, a, b, 25, corr, r_max)
    np.testing.assert_array_equal(ok, e_ok)
    np.testing.assert_array_equal(it, e_it)
    np.testing.assert_array_equal(out, e_out)


@pytest.mark.parametrize("k,rate,ebn0", CASES[:2])
@pytest.mark.parametrize("corr", [0, 2])
def test_monte_carlo_matches_oracle(oracle, k, rate, ebn0, corr):
    code = codes.ar4ja(k, rate)
    mu, sg = channel_params(ebn0, code.k / code.n_tx)
    n = 1536
    a, b = params(code, 1, 3)
    print(f"Profile D Monte Carlo K={k} rate={rate} corr={corr}", flush=True)
    with decoder.Engine(k, rate, max_iter=30, profile="D") as eng:
        eng.set_dmin_params(a, b, corr)
        got = eng.run_point(ebn0, n, seed=17, first_cw=9, chunk_cw=n)
    llr = oracle.gen_channel(code, 17, 9, n, mu, sg)
    exp = DO.run_point_d(code, llr, a, b, 30, corr=corr)
    assert (got.blocks, got.frame_errors, got.bit_errors, got.iter_sum) == \
        (exp["blocks"], exp["frame_errors"], exp["bit_errors"], exp["iter_sum"])
    np.testing.assert_array_equal(got.hist, exp["hist"])


def _params_k16384(code):
    rng = np.random.default_rng(7)
    p = DO.n_params(code)
    return rng.uniform(0.8, 0.9, (1, p)).astype(np.float32), rng.uniform(0.0, 0.1, (1, p)).astype(np.float32)


@pytest.mark.parametrize("corr,ebn0", [(1, 1.1), (2, 0.95)])
def test_monte_carlo_k16384_matches_oracle(oracle, corr, ebn0):
    """40 codewords (all through the queue: the workgroup-wide broadcast of
    the queued id), converged and failed codewords, two- and three-minimum
    corrections (the three-minimum kernel keeps four words per compressed
    check)."""
    code = codes.ar4ja(16384, "1/2")
    mu, sg = channel_params(ebn0)
    a, b = _params_k16384(code)
    print(f"Profile D Monte Carlo K=16384 r1/2, 40 codewords, corr={corr}", flush=True)
    with decoder.Engine(16384, "1/2", max_iter=30, profile="D") as eng:
        eng.set_dmin_params(a, b, corr)
        got = eng.run_point(ebn0, 40, seed=23, first_cw=4, chunk_cw=40)
    llr = oracle.gen_channel(code, 23, 4, 40, mu, sg)
    exp = DO.run_point_d(code, llr, a, b, 30, corr=corr)
    assert exp["frame_errors"] > 0 and exp["iter_sum"] < 30 * 40
    assert (got.blocks, got.frame_errors, got.bit_errors, got.iter_sum) == \
        (exp["blocks"], exp["frame_errors"], exp["bit_errors"], exp["iter_sum"])
    np.testing.assert_array_equal(got.hist, exp["hist"])


def test_monte_carlo_k16384_queue_agrees_with_decode():
    """320 codewords over 256 workgroups, each taken from the device queue
    (some workgroups take two).  The Monte Carlo counters and histogram equal the tallies of
    decoding the same channel (gen_channel + decode_llrs, whose arithmetic the
    decode test pins to the oracle)."""
    code = codes.ar4ja(16384, "1/2")
    a, b = _params_k16384(code)
    n, max_iter = 320, 30
    print("Profile D K=16384 r1/2: queue, 320 codewords", flush=True)
    with decoder.Engine(16384, "1/2", max_iter=max_iter, profile="D") as eng:
        eng.set_dmin_params(a, b, True)
        got = eng.run_point(1.1, n, seed=29, first_cw=0, chunk_cw=n)
        llr = eng.gen_channel(29, 0, n, 1.1)
        out, it, ok = eng.decode_llrs(llr)
    fail = ok == 0
    be = int(((out[fail, :code.k] & 0x8000) != 0).sum() + (out[fail, :code.k] == 0).sum())
    hist = np.bincount(np.where(ok == 1, it, 0), minlength=max_iter + 1).astype(np.uint64)
    assert (got.blocks, got.frame_errors, got.iter_sum) == (n, int(fail.sum()), int(it.sum()))
    assert got.bit_errors == be
    np.testing.assert_array_equal(got.hist, hist)
    assert 0 < got.frame_errors < n


def test_same_channel_as_profile_t():
    with decoder.Engine(1024, "1/2", profile="D", max_iter=20) as d, decoder.Engine(1024, "1/2") as t:
        np.testing.assert_array_equal(d.gen_channel(3, 100, 4, 1.5), t.gen_channel(3, 100, 4, 1.5))


def test_correction_modes_accepted():
    """Corrections 0, 1 and 2 are accepted; anything else is an argument error."""
    from ldpc_mi355x import _lib as L
    code = codes.ar4ja(16384, "1/2")
    a, b = _params_k16384(code)
    with decoder.Engine(16384, "1/2", profile="D", max_iter=20) as eng:
        for corr in (0, 1, 2):
            eng.set_dmin_params(a, b, corr)
        for corr in (3, -1):
            with pytest.raises(L.LdpcError) as e:
                eng.set_dmin_params(a, b, corr)
            assert e.value.status == L.LDPC_ERR_ARG


def test_parameters_required():
    from ldpc_mi355x import _lib as L
    with decoder.Engine(1024, "1/2", profile="D", max_iter=20) as eng:
        with pytest.raises(L.LdpcError):
            eng.run_point(2.0, 64)
