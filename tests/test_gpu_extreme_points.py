"""Monte Carlo far outside the waterfall, against the oracle (needs an MI355X).

Far below it every codeword runs to the iteration cap, so the failure path of
each kernel (bit-error count of every codeword, histogram bin 0, the cap in
the iteration sum) carries the whole tally; far above it nearly every
codeword converges at the first or second iteration, so the latch and the
refill run once per codeword.  Counters and histograms bit for bit, for the
specialised r1/2 K=1024 kernel, a wave-mode and a group-mode nine-code kernel,
Profile R (and NMS) and the learned decoder's kernels (Profile D, both
corrections).
"""
import os

import numpy as np
import pytest

from conftest import channel_params

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MATRIX = os.path.join(HERE, "golden", "AR4JA_r45_4c_128c_r12.chinn.out")
RATE = {"1/2": 0.5, "2/3": 2.0 / 3.0, "4/5": 0.8}


def _case(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _same(r, o):
    assert (r.blocks, r.frame_errors, r.bit_errors, r.iter_sum) == (o["blocks"], o["frame_errors"],
                                                                     o["bit_errors"], o["iter_sum"])
    np.testing.assert_array_equal(r.hist, o["hist"])


@pytest.mark.parametrize("k,rate", [(1024, "1/2"), (1024, "4/5"), (4096, "1/2")])
@pytest.mark.parametrize("ebn0,max_iter,n", [(-3.0, 12, 768), (12.0, 200, 2048)])
def test_profile_t_far_from_the_waterfall(oracle, k, rate, ebn0, max_iter, n):
    from ldpc_mi355x import codes, decoder
    n = n if k == 1024 else n // 4
    _case(f"Profile T K={k} r{rate} {ebn0} dB max_iter {max_iter} n {n}")
    with decoder.Engine(k, rate, max_iter=max_iter, beta=0.9375) as eng:
        r = eng.run_point(ebn0, n, seed=0xE1, first_cw=3, chunk_cw=n)
    mu, sg = channel_params(ebn0, RATE[rate])
    o = oracle.run_point_t(codes.ar4ja(k, rate), 0xE1, 3, n, mu, sg, max_iter, 0.9375)
    _same(r, o)
    if ebn0 < 0:
        assert r.frame_errors >= 0.99 * n and r.hist[0] == r.frame_errors
    else:
        assert r.frame_errors == 0 and r.hist[1] + r.hist[2] >= 0.99 * n


@pytest.mark.parametrize("profile", ["R", "R_NMS"])
@pytest.mark.parametrize("ebn0,max_iter,n", [(-3.0, 12, 768), (12.0, 60, 2048)])
def test_profile_r_far_from_the_waterfall(oracle, profile, ebn0, max_iter, n):
    from ldpc_mi355x import codes, decoder
    M, N, rows = codes.read_chinn_out(MATRIX)
    _case(f"Profile {profile} {ebn0} dB max_iter {max_iter} n {n}")
    with decoder.Engine(profile=profile, matrix=(rows, N), n_punct=512, max_iter=max_iter) as eng:
        r = eng.run_point(ebn0, n, seed=0xE2, first_cw=3, chunk_cw=n)
        mu, sg = eng.channel_params(ebn0)
    o = oracle.run_point_r(N, rows, 2048, 0xE2, 3, n, mu, sg, np.float32(0.0), max_iter, profile == "R_NMS", 0.0,
                           0.0)
    _same(r, o)
    if ebn0 < 0:
        assert r.frame_errors >= 0.99 * n
    else:
        # the syndrome is first checked after the third iteration
        assert r.frame_errors == 0 and r.hist[3] >= 0.99 * n


@pytest.mark.parametrize("k,ebn0,max_iter,n", [(1024, -3.0, 8, 256), (1024, 12.0, 30, 512),
                                               (16384, -3.0, 4, 8), (16384, 12.0, 20, 16)])
@pytest.mark.parametrize("corr", [1, 2])
def test_profile_d_far_from_the_waterfall(oracle, k, ebn0, max_iter, n, corr):
    """The learned decoder (Profile D, two- and three-minimum corrections)
    against its oracle; K=16384 runs its own kernel (messages partly in an
    HBM workspace)."""
    import dmin_oracle as DO
    from ldpc_mi355x import codes, decoder
    code = codes.ar4ja(k, "1/2")
    rng = np.random.default_rng(k)
    P = DO.n_params(code)
    a = rng.uniform(0.75, 0.95, (1, P)).astype(np.float32)
    b = rng.uniform(0.0, 0.3, (1, P)).astype(np.float32)
    _case(f"Profile D K={k} {ebn0} dB max_iter {max_iter} n {n} corr {corr}")
    with decoder.Engine(k, "1/2", max_iter=max_iter, profile="D") as eng:
        eng.set_dmin_params(a, b, corr)
        r = eng.run_point(ebn0, n, seed=0xE3, first_cw=3, chunk_cw=n)
    mu, sg = channel_params(ebn0, code.k / code.n_tx)
    llr = oracle.gen_channel(code, 0xE3, 3, n, mu, sg)
    o = DO.run_point_d(code, llr, a, b, max_iter, corr=corr)
    _same(r, o)
    if ebn0 < 0:
        assert r.frame_errors >= 0.99 * n
    else:
        assert r.frame_errors == 0


@pytest.mark.parametrize("ebn0,max_iter,n", [(-3.0, 12, 768), (12.0, 200, 2048)])
def test_random_codewords_far_from_the_waterfall(oracle, code12, ebn0, max_iter, n):
    """Random information bits (r1/2 K=1024): failures counted against the sent
    bits at -3 dB, clean convergence to the sent codeword at 12 dB."""
    from ldpc_mi355x import decoder
    _case(f"random codewords {ebn0} dB max_iter {max_iter} n {n}")
    with decoder.Engine(1024, "1/2", max_iter=max_iter, beta=0.9375) as eng:
        eng.set_codeword_mode(True)
        r = eng.run_point(ebn0, n, seed=0xE4, first_cw=3, chunk_cw=n)
    mu, sg = channel_params(ebn0)
    o = oracle.run_point_t_cw(code12, 0xE4, 3, n, mu, sg, max_iter, 0.9375)
    _same(r, o)
    if ebn0 < 0:
        # a failed frame keeps about the raw channel's error rate on its information bits: Q(mu / sigma) =
        # Q(sqrt(2 R g)) = Q(0.707) = 0.24 at -3 dB
        assert r.frame_errors >= 0.99 * n and 0.18 < r.bit_errors / (1024 * r.frame_errors) < 0.30
    else:
        assert r.frame_errors == 0 and r.bit_errors == 0


def test_batch_with_codes_far_from_the_waterfall():
    """One graph mixing codes far below, far above and inside their waterfalls
    equals the per-code runs counter for counter."""
    from ldpc_mi355x import decoder
    codes3 = [(1024, "4/5"), (4096, "1/2"), (16384, "2/3")]
    ebn0 = [-3.0, 12.0, 1.6]
    n_cw = [512, 96, 24]
    _case(f"batch {codes3} at {ebn0}")
    with decoder.Batch(codes3, max_iter=16) as b:
        got = b.run(ebn0, n_cw, seed=[7, 8, 9], first_cw=2, chunk_cw=[n // 2 for n in n_cw])
    for i, (k, rate) in enumerate(codes3):
        with decoder.Engine(k, rate, max_iter=16) as eng:
            ref = eng.run_point(ebn0[i], n_cw[i], seed=7 + i, first_cw=2, chunk_cw=n_cw[i])
        assert (got[i].blocks, got[i].frame_errors, got[i].bit_errors, got[i].iter_sum) == \
            (ref.blocks, ref.frame_errors, ref.bit_errors, ref.iter_sum), (k, rate)
        np.testing.assert_array_equal(got[i].hist, ref.hist)
    assert got[0].frame_errors == n_cw[0] and got[1].frame_errors == 0
