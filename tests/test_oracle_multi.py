"""Profile T oracle on the nine AR4JA codes.

The reference's golden tree update exists for degree-3 and degree-6 checks
only (tests/test_circulant_dma_stepper.cpp:107-327).  The rate-2/3 and
rate-4/5 codes have degree-10 and degree-18 checks; for those the oracle uses
the degree-6 tree generalised by pairs (j, j + d/2) and a prefix/suffix pass
over the pairs.  These tests pin that generalisation:

  * at degree 6 it is the golden's own tree, bit for bit;
  * where every min* correction vanishes (inputs spaced >= 2.5 apart) it is
    exactly leave-one-out min-sum, beta * min over the other edges;
  * with two small inputs among huge ones, each small edge receives exactly
    the other small value and every other edge min*(a, b) -- which fails for a
    wrong partner index or a dropped pair;
  * whole decodes that report convergence satisfy H x = 0 (checked with a
    dense parity-check product), a clean channel converges in one iteration.
"""
import numpy as np
import pytest

from conftest import channel_params
from ldpc_mi355x import codes


def bf16_grid(rng, n, lo=0.0, hi=12.0):
    """Random non-negative values on the bf16 grid."""
    x = rng.uniform(lo, hi, n).astype(np.float32)
    return (x.view(np.uint32) & 0xFFFF0000).view(np.float32)


def test_even_tree_at_degree_6_is_the_golden_tree(oracle):
    rng = np.random.default_rng(11)
    for trial in range(20000):
        m = bf16_grid(rng, 6, 0.0, [0.5, 3.0, 20.0][trial % 3])
        if trial % 7 == 0:
            m[rng.integers(0, 6)] = m[rng.integers(0, 6)]  # ties
        if trial % 11 == 0:
            m[rng.integers(0, 6)] = 0.0
        for beta in (0.9375, 1.0):
            np.testing.assert_array_equal(oracle.tree_even(m, beta).view(np.uint32),
                                          oracle.tree6(m, beta).view(np.uint32))


@pytest.mark.parametrize("deg", [6, 10, 18])
def test_separated_inputs_give_leave_one_out_min(oracle, deg):
    rng = np.random.default_rng(deg)
    for _ in range(500):
        m = (rng.permutation(deg) * 2.5 + bf16_grid(rng, 1, 0.0, 2.0)[0]).astype(np.float32)
        m = (m.view(np.uint32) & 0xFFFF0000).view(np.float32)
        assert len(set(m.tolist())) == deg and np.min(np.diff(np.sort(m))) >= 2.0
        r = oracle.tree_even(m, 0.9375)
        for j in range(deg):
            others = np.delete(m, j)
            assert r[j] == np.float32(np.min(others)) * np.float32(0.9375), (deg, j)


@pytest.mark.parametrize("deg", [6, 10, 18])
def test_two_small_inputs_among_huge_ones(oracle, deg):
    rng = np.random.default_rng(100 + deg)
    for _ in range(300):
        m = bf16_grid(rng, deg, 100.0, 200.0)
        a, b = rng.choice(deg, 2, replace=False)
        m[a], m[b] = bf16_grid(rng, 2, 0.0, 10.0)
        r = oracle.tree_even(m, 1.0)
        assert r[a] == m[b] and r[b] == m[a]
        ab = np.float32(oracle.amin_t(float(m[a]), float(m[b])))
        for j in set(range(deg)) - {a, b}:
            assert r[j] == ab, (deg, a, b, j)


@pytest.mark.parametrize("deg", [3, 6, 10, 18])
def test_check_update_signs_and_rounding(oracle, deg):
    """Outputs: R' = RNE(copysign(r, s ^ sign q)), L' = RNE(q + R')."""
    rng = np.random.default_rng(deg)
    for _ in range(200):
        L = (rng.normal(0, 4, deg).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
        R = (rng.normal(0, 2, deg).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
        Lo, Ro = oracle.check_update_t(L, R, iter0=False)
        f = lambda h: (h.astype(np.uint32) << 16).view(np.float32)  # noqa: E731
        q = f(L) - f(R)
        q = f(((q.view(np.uint32) + 0x7FFF + ((q.view(np.uint32) >> 16) & 1)) >> 16).astype(np.uint16))
        s = np.bitwise_xor.reduce((q.view(np.uint32) >> 31).astype(np.uint32))
        sign_r = (f(Ro).view(np.uint32) >> 31).astype(np.uint32)
        nz = f(Ro) != 0
        np.testing.assert_array_equal(sign_r[nz], (s ^ (q.view(np.uint32) >> 31))[nz])
        lq = q + f(Ro)
        exp = ((lq.view(np.uint32) + 0x7FFF + ((lq.view(np.uint32) >> 16) & 1)) >> 16).astype(np.uint16)
        np.testing.assert_array_equal(Lo, exp)


def test_unsupported_degree_raises(oracle):
    with pytest.raises(ValueError):
        oracle.check_update_t(np.zeros(5, np.uint16), np.zeros(5, np.uint16))


def dense_h(code):
    H = np.zeros((code.m, code.n), dtype=np.uint8)
    for i, r in enumerate(code.rows):
        H[i, r] = 1
    return H


@pytest.mark.parametrize("rate", ["1/2", "2/3", "4/5"])
@pytest.mark.parametrize("order", [(1, 2, 0), (0, 1, 2)])
def test_decode_k1024_converged_words_are_codewords(oracle, rate, order):
    code = codes.ar4ja(1024, rate, block_row_order=order)
    H = dense_h(code)
    num, den = {"1/2": (1, 2), "2/3": (2, 3), "4/5": (4, 5)}[rate]
    ebn0 = {"1/2": 1.6, "2/3": 2.4, "4/5": 3.2}[rate]
    mu, sg = channel_params(ebn0, rate=num / den)
    llr = oracle.gen_channel(code, 77, 0, 48, mu, sg)
    out, iters, ok = oracle.decode_t(code, llr, max_iter=50)
    assert ok.sum() >= 24, "most words converge at this Eb/N0"
    hard = ((out > 0x8000)).astype(np.uint8)
    for c in range(len(ok)):
        syn = (H.astype(np.int32) @ hard[c].astype(np.int32)) & 1
        assert (syn.sum() == 0) == bool(ok[c]), c
    clean = np.full((1, code.n), 0x4100, dtype=np.uint16)
    clean[0, code.n_tx:] = 0
    _, it, okc = oracle.decode_t(code, clean, max_iter=5)
    assert okc[0] == 1 and it[0] == 1


@pytest.mark.parametrize("k,rate", [(4096, "1/2"), (4096, "4/5")])
def test_decode_k4096_converges_at_high_snr(oracle, k, rate):
    code = codes.ar4ja(k, rate)
    num, den = {"1/2": (1, 2), "2/3": (2, 3), "4/5": (4, 5)}[rate]
    mu, sg = channel_params(4.0, rate=num / den)
    llr = oracle.gen_channel(code, 5, 0, 4, mu, sg)
    out, iters, ok = oracle.decode_t(code, llr, max_iter=30)
    assert ok.all()
    info_err = ((out[:, :code.k] & 0x8000) != 0) | (out[:, :code.k] == 0)
    assert not info_err.any()
