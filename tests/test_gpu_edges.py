"""Edge cases through the C-ABI on the GPU, for every kernel family
(specialised Profile T, the generic nine-code kernel in wave and group mode,
Profile R, Profile D and its K=16384 kernel): empty calls return without work and leave the
codeword cursor where it was; ragged batch sizes (a single codeword, a
partly filled wave of codeword slots, one more than a multiple of the slots)
decode bit for bit like the oracle; a Monte Carlo run split at arbitrary
points adds up to the same counters as one call."""
import os

import numpy as np
import pytest

from ldpc_mi355x import _lib as L
from ldpc_mi355x import codes, decoder

pytestmark = pytest.mark.gpu

CHINN = os.path.join(os.path.dirname(__file__), "golden", "AR4JA_r45_4c_128c_r12.chinn.out")


def _case(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _engines():
    """(name, factory) for one engine of every kernel family."""
    _, n, rows = codes.read_chinn_out(CHINN)
    p = 60  # parameters of r1/2 (sum of the layer degrees), K=1024 and 16384 alike
    a, b = np.full((1, p), 0.85, np.float32), np.full((1, p), 0.1, np.float32)

    def prof_d(k):
        def make():
            e = decoder.Engine(k, "1/2", max_iter=30, profile="D")
            e.set_dmin_params(a, b)
            return e
        return make
    return [
        ("T specialised r1/2 K=1024", lambda: decoder.Engine(1024, "1/2", max_iter=30)),
        ("T generic wave K=1024 r4/5", lambda: decoder.Engine(1024, "4/5", max_iter=30)),
        ("T generic group K=4096 r2/3", lambda: decoder.Engine(4096, "2/3", max_iter=30)),
        ("R .chinn.out", lambda: decoder.Engine(profile="R", matrix=(rows, n), n_punct=512, max_iter=30)),
        ("D r1/2 K=1024", prof_d(1024)),
        ("D r1/2 K=16384", prof_d(16384)),
    ]


@pytest.mark.parametrize("name,make", [pytest.param(n, f, id=n) for n, f in _engines()])
def test_empty_calls(name, make):
    _case(f"empty calls: {name}")
    with make() as eng:
        out, it, ok = eng.decode_llrs(np.zeros((0, eng.n), np.uint16))
        assert out.shape == (0, eng.n) and it.shape == (0,) and ok.shape == (0,)
        assert eng.gen_channel(5, 0, 0, 2.0).shape == (0, eng.n)
        r = eng.run_point(2.0, 0, seed=5, first_cw=1234)
        assert (r.blocks, r.frame_errors, r.bit_errors, r.iter_sum, r.launches, r.next_cw) == (0, 0, 0, 0, 0, 1234)
        assert int(np.sum(r.hist)) == 0


def _oracle_decode(oracle, name, eng, llr):
    if name.startswith("R"):
        _, n, rows = codes.read_chinn_out(CHINN)
        return oracle.decode_r(n, rows, llr, 30)
    if name.startswith("D"):
        import dmin_oracle as DO
        code = codes.ar4ja(eng.k_info, "1/2")
        p = DO.n_params(code)
        return DO.decode_d(code, llr, np.full((1, p), 0.85, np.float32), np.full((1, p), 0.1, np.float32), 30)
    return oracle.decode_t(codes.ar4ja(eng.k_info, eng.rate), llr, 30, 0.9375)


@pytest.mark.parametrize("name,make", [pytest.param(n, f, id=n) for n, f in _engines()])
def test_ragged_batch_sizes_decode_like_the_oracle(oracle, name, make):
    with make() as eng:
        for n_cw in (1, 3, 65):
            _case(f"ragged decode: {name} n_cw={n_cw}")
            ebn0 = 1.0 if eng.k_info == 16384 else (1.6 if eng.rate == "1/2" else (2.2 if eng.rate == "2/3" else 3.0))
            llr = eng.gen_channel(0xED6E + n_cw, 17, n_cw, ebn0)
            out, it, ok = eng.decode_llrs(llr)
            e_out, e_it, e_ok = _oracle_decode(oracle, name, eng, llr)
            np.testing.assert_array_equal(ok, e_ok)
            np.testing.assert_array_equal(it, e_it)
            np.testing.assert_array_equal(out, e_out)


@pytest.mark.parametrize("name,make", [pytest.param(n, f, id=n) for n, f in _engines()])
def test_monte_carlo_split_at_odd_points_adds_up(name, make):
    with make() as eng:
        ebn0 = 0.9 if eng.k_info == 16384 else (1.4 if eng.rate == "1/2" else (2.2 if eng.rate == "2/3" else 3.0))
        _case(f"split Monte Carlo: {name}")
        whole = eng.run_point(ebn0, 3001, seed=99, first_cw=10, chunk_cw=3001)
        acc = None
        first = 10
        for n in (1, 64, 999, 1937):
            acc = eng.run_point(ebn0, n, seed=99, first_cw=first, chunk_cw=700, result=acc)
            first = acc.next_cw
        assert first == whole.next_cw == 3011
        assert (acc.blocks, acc.frame_errors, acc.bit_errors, acc.iter_sum) == \
            (whole.blocks, whole.frame_errors, whole.bit_errors, whole.iter_sum)
        np.testing.assert_array_equal(acc.hist, whole.hist)
        assert whole.frame_errors > 0 or name.startswith("T generic group")


def test_abi_rejects_oversized_chunk():
    with decoder.Engine(1024, "1/2", max_iter=10) as eng:
        with pytest.raises(RuntimeError, match="chunk_cw"):
            eng.run_point(2.0, 10, seed=1, chunk_cw=(1 << 31) + 1)
    assert L.LDPC_ERR_ARG < 0
