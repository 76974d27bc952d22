"""Profile R (scalar_5w engine) HIP kernels against the reference worker's
fixtures and the oracle, through the C ABI (needs an MI355X).  Bit-exact."""
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR, channel_params
from ldpc_mi355x import codes

pytestmark = pytest.mark.gpu
MATRIX = os.path.join(GOLDEN_DIR, "AR4JA_r45_4c_128c_r12.chinn.out")


def _case(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def chinn():
    return codes.read_chinn_out(MATRIX)


def _engine(chinn, **kw):
    from ldpc_mi355x import decoder
    M, N, rows = chinn
    return decoder.Engine(matrix=(rows, N), n_punct=512, **kw)


def test_reference_worker_fixtures(chinn):
    z = np.load(os.path.join(GOLDEN_DIR, "scalar_profile_r.npz"))
    for i in range(len(z["nms"])):
        _case(f"scalar fixture {i}")
        with _engine(chinn, profile="R_NMS" if z["nms"][i] else "R", max_iter=int(z["max_iter"][i]),
                     r_max=float(z["r_max"][i]), l_post_max=float(z["l_post_max"][i])) as eng:
            out, iters, ok = eng.decode_llrs(z["llr_in"][i])
        np.testing.assert_array_equal(out, z["llr_out"][i], err_msg=f"case {i}")
        be = [int((out[c, :2048] >> 15).sum()) for c in range(2)]
        assert sum(be) == z["stats"][i][0] and sum(b > 0 for b in be) == z["stats"][i][1]
        assert int(iters[0]) == z["stats"][i][2]


@pytest.mark.parametrize("profile", ["R", "R_NMS"])
@pytest.mark.parametrize("ebn0", [0.8, 1.5, 2.2])
def test_decode_matches_oracle(chinn, oracle, profile, ebn0):
    M, N, rows = chinn
    _case(f"profile {profile} decode ebn0={ebn0}")
    mu, sg = channel_params(ebn0)
    llr = oracle.gen_channel_clamped(N, 2048, 0xABC, int(ebn0 * 100), 400, mu, sg)
    with _engine(chinn, profile=profile, max_iter=50) as eng:
        out, iters, ok = eng.decode_llrs(llr)
    o_out, o_it, o_ok = oracle.decode_r(N, rows, llr, 50, profile == "R_NMS")
    np.testing.assert_array_equal(out, o_out)
    np.testing.assert_array_equal(iters, o_it)
    np.testing.assert_array_equal(ok, o_ok)


def test_clamps_match_oracle(chinn, oracle):
    M, N, rows = chinn
    mu, sg = channel_params(1.2)
    llr = oracle.gen_channel_clamped(N, 2048, 5, 0, 300, mu, sg, 6.0)
    for prof, rmax, lpost in (("R", 1.0, 0.0), ("R", 0.0, 5.0), ("R_NMS", 2.0, 9.0)):
        _case(f"clamps {prof} r_max={rmax} l_post_max={lpost}")
        with _engine(chinn, profile=prof, max_iter=40, r_max=rmax, l_post_max=lpost) as eng:
            out, iters, ok = eng.decode_llrs(llr)
        o_out, o_it, _ = oracle.decode_r(N, rows, llr, 40, prof == "R_NMS", rmax, lpost)
        np.testing.assert_array_equal(out, o_out)
        np.testing.assert_array_equal(iters, o_it)


def test_gen_channel_clamped_bit_exact(chinn, oracle):
    M, N, rows = chinn
    for l_max, gamma in ((0.0, 0.0), (5.0, 0.0), (0.0, 1.5)):
        _case(f"gen_channel l_max={l_max} gamma={gamma}")
        with _engine(chinn, profile="R", max_iter=16, l_max=l_max, gamma=gamma) as eng:
            got = eng.gen_channel(77, 1000, 100, 1.7)
            mu, sg = eng.channel_params(1.7)
        lm = np.float32(np.float32(gamma) * np.float32(mu)) if gamma else l_max
        np.testing.assert_array_equal(got, oracle.gen_channel_clamped(N, 2048, 77, 1000, 100, mu, sg,
                                                                      np.float32(lm)))


@pytest.mark.parametrize("profile,ebn0,n,kw", [("R", 1.2, 3000, {}), ("R", 2.0, 4000, {}),
                                                ("R_NMS", 1.5, 3000, {}),
                                                ("R", 1.2, 2000, dict(l_max=8.0, r_max=6.0, l_post_max=20.0)),
                                                ("R", 1.2, 2000, dict(gamma=2.0))])
def test_monte_carlo_counters_exact(chinn, oracle, profile, ebn0, n, kw):
    M, N, rows = chinn
    _case(f"profile {profile} run_point ebn0={ebn0} n={n} {kw}")
    with _engine(chinn, profile=profile, max_iter=30, **kw) as eng:
        r = eng.run_point(ebn0, n, seed=31, first_cw=5, chunk_cw=1000)
        mu, sg = eng.channel_params(ebn0)
    lm = np.float32(np.float32(kw["gamma"]) * np.float32(mu)) if kw.get("gamma") else kw.get("l_max", 0.0)
    o = oracle.run_point_r(N, rows, 2048, 31, 5, n, mu, sg, np.float32(lm), 30, profile == "R_NMS",
                           kw.get("r_max", 0.0), kw.get("l_post_max", 0.0))
    assert (r.blocks, r.frame_errors, r.bit_errors, r.iter_sum) == (o["blocks"], o["frame_errors"], o["bit_errors"],
                                                                     o["iter_sum"])
    np.testing.assert_array_equal(r.hist, o["hist"])


def _random_rows(rng, n, degs):
    return [sorted(rng.choice(n, size=int(d), replace=False).tolist()) for d in degs]


@pytest.mark.parametrize("mix", ["exact36", "masked"])
def test_two_codeword_kernel_on_random_matrices(oracle, mix):
    """Random matrices through the two-codeword kernel: rows of degree 3 and 6
    (one-degree passes) and of degree 1..6 (masked passes, partly filled).
    Odd codeword counts leave a half without work at the end.  Decode and
    Monte Carlo are bit-exact against the oracle, for Approx-Min* and NMS."""
    from ldpc_mi355x import decoder
    rng = np.random.default_rng(11 if mix == "exact36" else 12)
    n, m = 700, 350
    degs = rng.choice([3, 6], size=m) if mix == "exact36" else rng.integers(1, 7, size=m)
    rows = _random_rows(rng, n, degs)
    ps, kind = decoder.row_passes_by_degree(rows, n, 32)
    assert ps.max() + 1 <= 64 and (set(kind.tolist()) <= {3, 6} if mix == "exact36" else set(kind.tolist()) == {7})
    mu, sg = channel_params(1.5)
    llr = oracle.gen_channel_clamped(n, n, 0x51, 0, 37, mu, sg)
    for nms in (False, True):
        _case(f"two-codeword kernel {mix} nms={nms} decode 37")
        with decoder.Engine(matrix=(rows, n), n_punct=0, profile="R_NMS" if nms else "R", max_iter=25) as eng:
            out, iters, ok = eng.decode_llrs(llr)
            _case(f"two-codeword kernel {mix} nms={nms} run_point 1001")
            r = eng.run_point(1.5, 1001, seed=9, first_cw=3, chunk_cw=1001)
            emu, esg = eng.channel_params(1.5)
        o_out, o_it, o_ok = oracle.decode_r(n, rows, llr, 25, nms)
        np.testing.assert_array_equal(out, o_out)
        np.testing.assert_array_equal(iters, o_it)
        np.testing.assert_array_equal(ok, o_ok)
        o = oracle.run_point_r(n, rows, n, 9, 3, 1001, emu, esg, np.float32(0.0), 25, nms, 0.0, 0.0)
        assert (r.blocks, r.frame_errors, r.bit_errors, r.iter_sum) == \
            (o["blocks"], o["frame_errors"], o["bit_errors"], o["iter_sum"])
        np.testing.assert_array_equal(r.hist, o["hist"])


@pytest.mark.parametrize("max_iter", [1, 2, 3])
def test_two_codeword_kernel_iteration_caps(chinn, oracle, max_iter):
    """Caps at and below the third iteration, where the syndrome check starts:
    one codeword (one half of the wave idle), three codewords (a refill), and
    a Monte Carlo point, against the oracle."""
    M, N, rows = chinn
    mu, sg = channel_params(2.5)
    llr = oracle.gen_channel_clamped(N, 2048, 0x77, 0, 3, mu, sg)
    for nms in (False, True):
        _case(f"two-codeword kernel max_iter={max_iter} nms={nms}")
        with _engine(chinn, profile="R_NMS" if nms else "R", max_iter=max_iter) as eng:
            for k in (1, 3):
                out, iters, ok = eng.decode_llrs(llr[:k])
                o_out, o_it, o_ok = oracle.decode_r(N, rows, llr[:k], max_iter, nms)
                np.testing.assert_array_equal(out, o_out)
                np.testing.assert_array_equal(iters, o_it)
                np.testing.assert_array_equal(ok, o_ok)
            r = eng.run_point(2.5, 257, seed=4, first_cw=1, chunk_cw=257)
            emu, esg = eng.channel_params(2.5)
        o = oracle.run_point_r(N, rows, 2048, 4, 1, 257, emu, esg, np.float32(0.0), max_iter, nms, 0.0, 0.0)
        assert (r.blocks, r.frame_errors, r.bit_errors, r.iter_sum) == \
            (o["blocks"], o["frame_errors"], o["bit_errors"], o["iter_sum"])
        np.testing.assert_array_equal(r.hist, o["hist"])


def test_one_codeword_kernel_still_exact(chinn, oracle, monkeypatch):
    """LDPC_R_ONE_CW=1 keeps the one-codeword kernel (messages in LDS); it stays
    bit-exact on the shipped matrix, and a matrix whose dependency chain needs
    more than 64 passes falls back to it by itself."""
    from ldpc_mi355x import decoder
    M, N, rows = chinn
    mu, sg = channel_params(1.5)
    llr = oracle.gen_channel_clamped(N, 2048, 0xABD, 0, 130, mu, sg)
    monkeypatch.setenv("LDPC_R_ONE_CW", "1")
    _case("one-codeword kernel, shipped matrix")
    with _engine(chinn, profile="R", max_iter=40) as eng:
        out, iters, ok = eng.decode_llrs(llr)
    o_out, o_it, _ = oracle.decode_r(N, rows, llr, 40)
    np.testing.assert_array_equal(out, o_out)
    np.testing.assert_array_equal(iters, o_it)
    monkeypatch.delenv("LDPC_R_ONE_CW")
    chain = [[i, i + 1, (i + 7) % 101] for i in range(100)]
    assert decoder.row_passes_by_degree(chain, 101, 32)[0].max() + 1 > 64
    llr = oracle.gen_channel_clamped(101, 101, 3, 0, 20, 2.0, 1.5)
    _case("long-chain matrix")
    with decoder.Engine(matrix=(chain, 101), n_punct=0, profile="R", max_iter=12) as eng:
        out, iters, ok = eng.decode_llrs(llr)
    o_out, o_it, _ = oracle.decode_r(101, chain, llr, 12)
    np.testing.assert_array_equal(out, o_out)
    np.testing.assert_array_equal(iters, o_it)


def test_profile_r_on_ar4ja_families(oracle):
    """Profile R on makeH-order AR4JA codes, including rate 2/3 (degree 10,
    the 12-wide row table)."""
    from ldpc_mi355x import decoder
    for rate in ("1/2", "2/3"):
        c = codes.ar4ja(1024, rate)
        rows = [c.rows[i] for i in range(c.m)]
        mu, sg = channel_params(2.0, c.rate)
        llr = oracle.gen_channel_clamped(c.n, c.n_tx, 3, 0, 200, mu, sg)
        _case(f"AR4JA r{rate} K=1024 Profile R")
        with decoder.Engine(1024, rate, max_iter=40, profile="R") as eng:
            assert (eng.n, eng.m, eng.n_tx) == (c.n, c.m, c.n_tx)
            out, iters, ok = eng.decode_llrs(llr)
        o_out, o_it, _ = oracle.decode_r(c.n, rows, llr, 40)
        np.testing.assert_array_equal(out, o_out)
        np.testing.assert_array_equal(iters, o_it)


def test_unsupported_sizes_fail_loudly():
    from ldpc_mi355x import _lib as L, decoder
    with pytest.raises(L.LdpcError):
        decoder.Engine(16384, "1/2", profile="R")   # LLRs + messages exceed one wave's LDS
    with pytest.raises(L.LdpcError):
        decoder.Engine(1024, "4/5", profile="R")    # degree 18 > 12
