"""HIP kernels against the oracle, through the C ABI (needs an MI355X).

Bit-exact for everything: channel LLRs, final LLRs, iteration counts,
convergence flags, Monte Carlo counters and the iteration histogram.
"""
import numpy as np
import pytest

from conftest import channel_params

pytestmark = pytest.mark.gpu


def _print_case(msg):
    print(f"[gpu-case] {msg}", flush=True)


def test_golden_fixtures(engine, golden, oracle, code12):
    """Every reference-stepper fixture codeword, decoded on the GPU."""
    from ldpc_mi355x import decoder
    for beta in (0.9375, 1.0):
        for mi in (200, 40, 30):
            sel = [i for i in range(len(golden["label"])) if golden["beta"][i] == np.float32(beta)
                   and golden["max_iter"][i] == mi]
            if not sel:
                continue
            _print_case(f"golden beta={beta} max_iter={mi} n={len(sel)}")
            with decoder.Engine(1024, "1/2", max_iter=mi, beta=beta) as eng:
                out, iters, ok = eng.decode_llrs(golden["llr_in"][sel])
            np.testing.assert_array_equal(out, golden["llr_out"][sel])
            np.testing.assert_array_equal(np.where(ok == 1, iters, 0), golden["conv_iter"][sel])


@pytest.mark.parametrize("ebn0", [0.5, 1.0, 1.5, 2.0, 2.5])
def test_decode_matches_oracle(engine, oracle, code12, ebn0):
    _print_case(f"decode_llrs ebn0={ebn0}")
    mu, sg = channel_params(ebn0)
    llr = oracle.gen_channel(code12, 0xD1CE, int(ebn0 * 1000), 512, mu, sg)
    out, iters, ok = engine.decode_llrs(llr)
    o_out, o_it, o_ok = oracle.decode_t(code12, llr, 200, 0.9375)
    np.testing.assert_array_equal(iters, o_it)
    np.testing.assert_array_equal(ok, o_ok)
    np.testing.assert_array_equal(out, o_out)


def test_decode_random_bit_patterns(oracle, code12):
    """Arbitrary finite bf16 inputs, including -0, tiny and huge magnitudes."""
    from ldpc_mi355x import decoder
    rng = np.random.default_rng(7)
    n = 256
    e = rng.integers(0x50, 0xB0, size=(n, code12.n))
    m = rng.integers(0, 128, size=(n, code12.n))
    s = rng.integers(0, 2, size=(n, code12.n))
    llr = ((s << 15) | (e << 7) | m).astype(np.uint16)
    llr[rng.random(llr.shape) < 0.02] = 0x8000
    llr[rng.random(llr.shape) < 0.02] = 0x0000
    _print_case("random bit patterns max_iter=25")
    with decoder.Engine(1024, "1/2", max_iter=25, beta=0.9375) as eng:
        out, iters, ok = eng.decode_llrs(llr)
    o_out, o_it, o_ok = oracle.decode_t(code12, llr, 25, 0.9375)
    np.testing.assert_array_equal(out, o_out)
    np.testing.assert_array_equal(iters, o_it)


def test_decode_edge_sizes(engine, oracle, code12):
    mu, sg = channel_params(1.5)
    for n in (1, 3, 64, 1000):
        _print_case(f"decode n_cw={n}")
        llr = oracle.gen_channel(code12, 5, 0, n, mu, sg)
        out, iters, ok = engine.decode_llrs(llr)
        o_out, o_it, _ = oracle.decode_t(code12, llr, 200, 0.9375)
        np.testing.assert_array_equal(out, o_out)
        np.testing.assert_array_equal(iters, o_it)
    out, iters, ok = engine.decode_llrs(np.zeros((0, code12.n), dtype=np.uint16))
    assert out.shape == (0, code12.n)


def test_gen_channel_bit_exact(engine, oracle, code12):
    for ebn0 in (1.0, 2.0, 2.5):
        _print_case(f"gen_channel ebn0={ebn0}")
        got = engine.gen_channel(0x5EED2026, 12345, 300, ebn0)
        mu, sg = engine.channel_params(ebn0)
        assert (np.float32(mu), np.float32(sg)) == channel_params(ebn0)
        np.testing.assert_array_equal(got, oracle.gen_channel(code12, 0x5EED2026, 12345, 300, mu, sg))
    big = engine.gen_channel(3, 1 << 33, 4, 2.0)  # codeword ids above 2^32
    np.testing.assert_array_equal(big, oracle.gen_channel(code12, 3, 1 << 33, 4, *channel_params(2.0)))


@pytest.mark.parametrize("ebn0,n", [(1.0, 3000), (1.5, 4000), (2.0, 6000)])
def test_monte_carlo_counters_exact(engine, oracle, code12, ebn0, n):
    """Fused GPU Monte Carlo == oracle Monte Carlo on the same seed, counter for counter."""
    _print_case(f"run_point ebn0={ebn0} n={n}")
    r = engine.run_point(ebn0, n, seed=0x5EED2026, first_cw=7, chunk_cw=1024)
    mu, sg = channel_params(ebn0)
    o = oracle.run_point_t(code12, 0x5EED2026, 7, n, mu, sg, 200, 0.9375)
    assert (r.blocks, r.frame_errors, r.bit_errors, r.iter_sum) == (o["blocks"], o["frame_errors"],
                                                                     o["bit_errors"], o["iter_sum"])
    np.testing.assert_array_equal(r.hist, o["hist"])
    assert r.next_cw == 7 + n and r.launches == -(-n // 1024)


def test_monte_carlo_stopping_rule(engine):
    _print_case("run_point target_fe at 0.8 dB")
    r = engine.run_point(0.8, 1 << 16, target_fe=5, chunk_cw=512)
    assert r.frame_errors >= 5 and r.blocks < (1 << 16) and r.blocks % 512 == 0


def test_monte_carlo_chunking_invariance(engine):
    """Counters depend only on (seed, codeword ids): chunking and resume do not matter."""
    _print_case("chunking invariance 2^15 codewords at 1.2 dB")
    a = engine.run_point(1.2, 1 << 15, seed=11, chunk_cw=1 << 15)
    b = engine.run_point(1.2, 1 << 14, seed=11, chunk_cw=3000)
    b = engine.run_point(1.2, 1 << 14, seed=11, first_cw=b.next_cw, chunk_cw=5000, result=b)
    assert (a.blocks, a.frame_errors, a.bit_errors, a.iter_sum) == (b.blocks, b.frame_errors, b.bit_errors,
                                                                     b.iter_sum)
    np.testing.assert_array_equal(a.hist, b.hist)


def test_monte_carlo_full_launch_invariants(engine):
    """One full 2^20-codeword launch at 2.0 dB: size-independent properties."""
    _print_case("2^20 launch at 2.0 dB")
    r = engine.run_point(2.0, 1 << 20, seed=0xBEEF)
    assert r.blocks == 1 << 20 and r.hist.sum() == r.blocks and r.hist[0] == r.frame_errors
    assert r.iter_sum == int((np.arange(201) * r.hist)[1:].sum()) + 200 * int(r.hist[0])
    assert 6.4 < r.mean_iterations < 7.3
    assert r.frame_errors < 60  # FER at 2.0 dB ~2.4e-6 (README.md:107): ~2.5 expected
    r2 = engine.run_point(2.0, 1 << 20, seed=0xBEEF)
    assert (r2.frame_errors, r2.bit_errors, r2.iter_sum) == (r.frame_errors, r.bit_errors, r.iter_sum)


def test_device_pointer_entry_point(engine, oracle, code12):
    """ldpc_decode_llrs_device_async on device buffers allocated with the same
    HIP runtime libldpc links (a second runtime in one process, such as the
    one bundled with torch, cannot open the GPU once this one has)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    H2D, D2H = 1, 2

    def dmalloc(nbytes):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        return p

    mu, sg = channel_params(1.0)
    n = 200
    llr = oracle.gen_channel(code12, 77, 0, n, mu, sg)
    d_in, d_out = dmalloc(llr.nbytes), dmalloc(llr.nbytes)
    d_it, d_ok = dmalloc(4 * n), dmalloc(n)
    try:
        assert hip.hipMemcpy(d_in, llr.ctypes.data_as(C.c_void_p), C.c_size_t(llr.nbytes), H2D) == 0
        _print_case("device async decode n=200")
        engine.decode_llrs_device_async(d_in.value, n, d_out.value, d_it.value, d_ok.value)
        engine.sync()
        out = np.empty_like(llr)
        it = np.empty(n, dtype=np.uint32)
        ok = np.empty(n, dtype=np.uint8)
        for host, dev in ((out, d_out), (it, d_it), (ok, d_ok)):
            assert hip.hipMemcpy(host.ctypes.data_as(C.c_void_p), dev, C.c_size_t(host.nbytes), D2H) == 0
    finally:
        for p in (d_in, d_out, d_it, d_ok):
            hip.hipFree(p)
    o_out, o_it, o_ok = oracle.decode_t(code12, llr, 200, 0.9375)
    np.testing.assert_array_equal(out, o_out)
    np.testing.assert_array_equal(it, o_it)
    np.testing.assert_array_equal(ok, o_ok)
