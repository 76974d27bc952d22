"""GPU parity of the nine-code Profile T kernels (decode_multi.hip) against the
oracle, bit for bit: decode of channel dumps and of random bit patterns for
every kernel variant, channel generation, Monte Carlo counters and
histograms; the generic kernel against the specialised r1/2 K=1024 kernel;
the mixed-code batch graph against per-code runs; the autotuner."""
import numpy as np
import pytest

from conftest import channel_params
from ldpc_mi355x import _lib as L
from ldpc_mi355x import codes, decoder

pytestmark = pytest.mark.gpu

RATE = {"1/2": 0.5, "2/3": 2.0 / 3.0, "4/5": 0.8}
# Eb/N0 in each code's waterfall: a mix of converged and failed codewords
WATERFALL = {("1/2", 1024): 1.4, ("2/3", 1024): 2.2, ("4/5", 1024): 3.0,
             ("1/2", 4096): 1.0, ("2/3", 4096): 1.8, ("4/5", 4096): 2.6,
             ("1/2", 16384): 0.8, ("2/3", 16384): 1.6, ("4/5", 16384): 2.4}
N_DUMP = {1024: 96, 4096: 24, 16384: 8}
ORDERS = {0: codes.BLOCK_ROW_ORDER_T, 1: (0, 1, 2)}


def code_of(k, rate, order=0):
    return codes.ar4ja(k, rate, block_row_order=ORDERS[order])


def random_patterns(rng, n_cw, n, n_tx):
    """Random bf16 bit patterns with exponents 2^-47 .. 2^48 (sums of them stay
    finite), -0 and +0 sprinkled in; punctured variables +0."""
    e = rng.integers(0x50, 0xB0, size=(n_cw, n))
    m = rng.integers(0, 128, size=(n_cw, n))
    s = rng.integers(0, 2, size=(n_cw, n))
    x = ((s << 15) | (e << 7) | m).astype(np.uint16)
    x[rng.random(x.shape) < 0.02] = 0x8000
    x[rng.random(x.shape) < 0.02] = 0x0000
    x[:, n_tx:] = 0
    return x


@pytest.mark.parametrize("k,rate", decoder.AR4JA_CODES)
def test_decode_every_variant_matches_oracle(oracle, k, rate):
    rng = np.random.default_rng(k + int(RATE[rate] * 100))
    mu, sg = channel_params(WATERFALL[(rate, k)], RATE[rate])
    max_iter = 30 if k < 16384 else 12
    base = code_of(k, rate)
    llr = oracle.gen_channel(base, 2026, 0, N_DUMP[k], mu, sg)
    llr = np.concatenate([llr, random_patterns(rng, 4, base.n, base.n_tx)])
    expect = {}
    for v in decoder.list_variants(k, rate):
        c, order = L.variant_fields(v)
        print(f"decode K={k} rate={rate} C={c} order={order}", flush=True)
        if order not in expect:
            expect[order] = oracle.decode_t(code_of(k, rate, order), llr, max_iter=max_iter)
        e_out, e_it, e_ok = expect[order]
        with decoder.Engine(k, rate, max_iter=max_iter, variant=v) as eng:
            out, it, ok = eng.decode_llrs(llr)
        np.testing.assert_array_equal(ok, e_ok, err_msg=f"ok C={c} order={order}")
        np.testing.assert_array_equal(it, e_it, err_msg=f"iters C={c} order={order}")
        np.testing.assert_array_equal(out, e_out, err_msg=f"llr C={c} order={order}")
    print(f"K={k} rate={rate}: {int(expect[0][2][:N_DUMP[k]].sum())}/{N_DUMP[k]} channel dumps converged", flush=True)


@pytest.mark.parametrize("k,rate", decoder.AR4JA_CODES)
def test_gen_channel_matches_oracle(oracle, k, rate):
    mu, sg = channel_params(2.0, RATE[rate])
    with decoder.Engine(k, rate) as eng:
        got = eng.gen_channel(99, 1 << 33, 3, 2.0)
        assert eng.channel_params(2.0) == (mu, sg)
    np.testing.assert_array_equal(got, oracle.gen_channel(code_of(k, rate), 99, 1 << 33, 3, mu, sg))


@pytest.mark.parametrize("k,rate", decoder.AR4JA_CODES)
def test_monte_carlo_default_variant_matches_oracle(oracle, k, rate):
    n_cw = {1024: 1024, 4096: 192, 16384: 32}[k]
    max_iter = 40 if k < 16384 else 16
    ebn0 = WATERFALL[(rate, k)]
    mu, sg = channel_params(ebn0, RATE[rate])
    exp = oracle.run_point_t(code_of(k, rate), 0x5EED2026, 7, n_cw, mu, sg, max_iter=max_iter)
    for v in (0, L.variant(2, 0)):  # library default and C = 2
        with decoder.Engine(k, rate, max_iter=max_iter, variant=v) as eng:
            got = eng.run_point(ebn0, n_cw, seed=0x5EED2026, first_cw=7, chunk_cw=n_cw)
        assert (got.blocks, got.frame_errors, got.bit_errors, got.iter_sum) == \
            (exp["blocks"], exp["frame_errors"], exp["bit_errors"], exp["iter_sum"]), (k, rate, v)
        np.testing.assert_array_equal(got.hist, exp["hist"])


def test_layer_order_1_monte_carlo_matches_oracle(oracle):
    k, rate = 1024, "2/3"
    mu, sg = channel_params(2.2, RATE[rate])
    exp = oracle.run_point_t(code_of(k, rate, 1), 3, 0, 2048, mu, sg, max_iter=50)
    for c in (1, 2):
        with decoder.Engine(k, rate, max_iter=50, variant=L.variant(c, 1)) as eng:
            got = eng.run_point(2.2, 2048, seed=3, chunk_cw=2048)
        assert (got.blocks, got.frame_errors, got.bit_errors, got.iter_sum) == \
            (exp["blocks"], exp["frame_errors"], exp["bit_errors"], exp["iter_sum"])
        np.testing.assert_array_equal(got.hist, exp["hist"])


def test_generic_kernel_equals_specialised_kernel_r12_k1024():
    """2^16 codewords at 1.5 dB through the specialised kernel and both generic
    variants: identical counters and iteration histograms."""
    runs = []
    for v in (0, L.variant(2, 0), L.variant(1, 0)):
        with decoder.Engine(1024, "1/2", max_iter=200, variant=v) as eng:
            r = eng.run_point(1.5, 1 << 16, seed=41, chunk_cw=1 << 16)
        runs.append((r.blocks, r.frame_errors, r.bit_errors, r.iter_sum, r.hist.tobytes()))
        print(f"variant {v:#x}: fe {r.frame_errors} mean it {r.mean_iterations:.3f} {r.kernel_gsps:.2f} Gsps",
              flush=True)
    assert runs[0] == runs[1] == runs[2]
    assert runs[0][1] > 0


def test_batch_of_nine_codes_equals_per_code_runs():
    codes9 = decoder.AR4JA_CODES
    ebn0 = [WATERFALL[(r, k)] for k, r in codes9]
    n_cw = [{1024: 768, 4096: 160, 16384: 40}[k] for k, _ in codes9]
    chunk = [max(8, n // 3) for n in n_cw]  # several rounds (graph replays) per code
    with decoder.Batch(codes9, max_iter=24) as b:
        got = b.run(ebn0, n_cw, seed=[100 + i for i in range(9)], first_cw=5, chunk_cw=chunk)
        again = b.run(ebn0, n_cw, seed=[100 + i for i in range(9)], first_cw=5, chunk_cw=chunk)
    for i, (k, rate) in enumerate(codes9):
        with decoder.Engine(k, rate, max_iter=24) as eng:
            ref = eng.run_point(ebn0[i], n_cw[i], seed=100 + i, first_cw=5, chunk_cw=n_cw[i])
        for g in (got[i], again[i]):
            assert (g.blocks, g.frame_errors, g.bit_errors, g.iter_sum) == \
                (ref.blocks, ref.frame_errors, ref.bit_errors, ref.iter_sum), (k, rate)
            np.testing.assert_array_equal(g.hist, ref.hist)
            assert g.next_cw == 5 + n_cw[i]
        assert got[i].launches == -(-n_cw[i] // chunk[i])


def test_batch_stops_each_code_at_its_target():
    codes2 = [(1024, "1/2"), (1024, "4/5")]
    with decoder.Batch(codes2, max_iter=20) as b:
        res = b.run([0.8, 2.0], [1 << 14, 1 << 14], target_fe=[5, 1 << 20], chunk_cw=1024)
    assert res[0].frame_errors >= 5 and res[0].blocks < 1 << 14
    assert res[1].blocks == 1 << 14


def test_autotune_reports_every_variant():
    best, rows = decoder.autotune(1024, "2/3", 2.2, n_cw=4096, reps=1, max_iter=50)
    assert {r["variant"] for r in rows} == set(decoder.list_variants(1024, "2/3"))
    assert best in {r["variant"] for r in rows if r["layer_order"] == 0}
    by_order = {}
    for r in rows:
        by_order.setdefault(r["layer_order"], set()).add((r["frame_errors"], round(r["mean_iterations"], 9)))
        assert r["gsps"] > 0 and r["wg_per_cu"] >= 1
    assert all(len(s) == 1 for s in by_order.values())
