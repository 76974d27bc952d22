"""The AR4JA encoder (libldpc's ldpc_encode / ldpc_get_generator) and the
oracle's random-codeword path, on the CPU.

A codeword c of info bits u is pinned by the mathematics alone: it is
systematic (c starts with u) and satisfies every parity check (H c = 0).  The
parity part of every AR4JA H (its last 3M columns) is invertible, which the
oracle's dense elimination confirms for the K=1024 codes and the encoder's
unit-pivot elimination for all nine, so these two properties leave exactly
one codeword per u: any encoder that meets them is the encoder.
"""
import numpy as np
import pytest

from conftest import channel_params
from ldpc_mi355x import codes, decoder

ALL = [(k, r) for k in (1024, 4096, 16384) for r in ("1/2", "2/3", "4/5")]


def parity_ok(code, x):
    """H x = 0 for every row of x (sparse rows of H, exact integers)."""
    x = np.asarray(x, dtype=np.int64)
    for row in code.rows:
        if np.any(x[:, row].sum(axis=1) & 1):
            return False
    return True


@pytest.mark.parametrize("k,rate", ALL, ids=[f"K{k}-r{r.replace('/', '')}" for k, r in ALL])
def test_encoder_codewords_satisfy_every_check(k, rate):
    code = codes.ar4ja(k, rate)
    rng = np.random.default_rng(k + len(rate))
    n_cw = 3 if k < 16384 else 2
    u = rng.integers(0, 2, (n_cw, code.k), dtype=np.uint8)
    u[0] = 0                      # the all-zero word encodes to the all-zero codeword
    x = decoder.encode(k, rate, u)
    assert x.shape == (n_cw, code.n)
    assert np.array_equal(x[:, :code.k], u)
    assert not x[0].any()
    assert parity_ok(code, x)
    # random words are not trivially zero in any parity block
    z = code.z
    par = x[1:, code.k:].reshape(n_cw - 1, 12, z)
    assert np.all(par.any(axis=2))
    # a flipped parity bit breaks a check (the test can fail)
    bad = x[1:2].copy()
    bad[0, code.k + 5] ^= 1
    assert not parity_ok(code, bad)


@pytest.mark.parametrize("k,rate", [(1024, "1/2"), (4096, "2/3")])
def test_encoder_is_linear(k, rate):
    code = codes.ar4ja(k, rate)
    rng = np.random.default_rng(5)
    a, b = rng.integers(0, 2, (2, code.k), dtype=np.uint8)
    xa, xb, xab = decoder.encode(k, rate, np.stack([a, b, a ^ b]))
    assert np.array_equal(xa ^ xb, xab)


def test_generator_matches_encoder_and_shape():
    """p_r = sum_i g_{r,i} u_i mod x^z + 1 with the generator words as
    documented: a unit info vector e_{i,0} (bit 0 of block i) yields g_{r,i}
    itself as parity block r."""
    k, rate = 1024, "2/3"
    code = codes.ar4ja(k, rate)
    z = code.z
    g = decoder.generator(k, rate)
    assert g.shape == (12, k // z, z // 32)
    for i in (0, 3, k // z - 1):
        u = np.zeros((1, k), np.uint8)
        u[0, i * z] = 1
        x = decoder.encode(k, rate, u)[0]
        par = x[k:].reshape(12, z)
        bits = ((g[:, i, :, None] >> np.arange(32, dtype=np.uint32)) & 1).reshape(12, z)
        assert np.array_equal(par, bits)


def test_encoder_rejects_bad_input():
    import ctypes as C
    from ldpc_mi355x import _lib as L
    lib = L.load()
    d = decoder.ar4ja_desc(1024, "1/2")
    u = np.full((1, 1024), 2, np.uint8)
    x = np.zeros((1, 2560), np.uint8)
    assert lib.ldpc_encode(C.byref(d), u.ctypes.data, 1, x.ctypes.data) == L.LDPC_ERR_ARG
    assert not x.any()
    bad = L.CodeDesc(L.LDPC_CODE_AR4JA, 2048, 1, 2)
    assert lib.ldpc_encode(C.byref(bad), u.ctypes.data, 1, x.ctypes.data) == L.LDPC_ERR_UNSUPPORTED
    assert lib.ldpc_get_generator(C.byref(bad), None, 0) == L.LDPC_ERR_UNSUPPORTED


@pytest.mark.parametrize("rate", ["1/2", "2/3", "4/5"])
def test_oracle_dense_encoder_equals_product_encoder(oracle, rate):
    """The oracle encodes with a dense generator from plain Gauss-Jordan
    elimination of [H_p | H_u]; the product with the quasi-cyclic ring
    elimination.  Both give the same codewords (and satisfy H c = 0)."""
    code = codes.ar4ja(1024, rate)
    llr, bits = oracle.gen_channel_cw(code, 0x5EED, 100, 40, 4.0, 1e-6, with_bits=True)
    assert parity_ok(code, bits)
    assert np.array_equal(decoder.encode(1024, rate, bits[:, :code.k]), bits)
    # at negligible noise the LLR signs are the sent bits; punctured stay +0
    assert np.array_equal((llr[:, :code.n_tx] >> 15).astype(np.uint8), bits[:, :code.n_tx])
    assert not llr[:, code.n_tx:].any()
    # the info bits are Philox(ctr = (id, 0, j, 1)): distinct per codeword, about half ones
    assert len({bits[c, :code.k].tobytes() for c in range(40)}) == 40
    assert 0.45 < bits[:, :code.k].mean() < 0.55


def test_oracle_random_codewords_match_all_zero_fer(oracle, code12):
    """Random codewords and the all-zero codeword see the same channel law, and
    the decoder treats both signs alike (up to the +-0 rules), so the frame
    error rates agree at 1.5 dB (exact two-sample test on the same ids)."""
    from scipy.stats import binomtest
    mu, sg = channel_params(1.5)
    n = 40000
    z = oracle.run_point_t(code12, 0x5EED2026, 0, n, mu, sg, 200, 0.9375)
    r = oracle.run_point_t_cw(code12, 0x5EED2026, 0, n, mu, sg, 200, 0.9375)
    assert r["blocks"] == n and int(r["hist"].sum()) == n
    assert r["frame_errors"] >= int(r["hist"][0])
    k1, k2 = r["frame_errors"], z["frame_errors"]
    assert k1 > 50 and k2 > 50
    p = binomtest(k1, k1 + k2, 0.5).pvalue
    print("random", k1, "all-zero", k2, "p", p, "mean it", r["iter_sum"] / n, z["iter_sum"] / n)
    assert p > 1e-3
    assert abs(r["iter_sum"] - z["iter_sum"]) / z["iter_sum"] < 0.02
