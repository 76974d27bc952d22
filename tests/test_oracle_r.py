"""Profile R oracle (scalar_5w recursive Approx-Min* / normalized min-sum),
pinned against the reference's own worker (tests/golden/scalar_profile_r.npz,
written by oracle/_ref/scalar_driver around ldpc_decoder_core.h)."""
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR, REFERENCE
from ldpc_mi355x import codes

MATRIX = os.path.join(GOLDEN_DIR, "AR4JA_r45_4c_128c_r12.chinn.out")


@pytest.fixture(scope="module")
def scalar_golden():
    return np.load(os.path.join(GOLDEN_DIR, "scalar_profile_r.npz"))


@pytest.fixture(scope="module")
def chinn():
    return codes.read_chinn_out(MATRIX)


def test_matrix_copy_is_the_reference_file(scalar_golden):
    assert hashlib.sha256(open(MATRIX, "rb").read()).hexdigest() == str(scalar_golden["matrix_sha256"])
    if os.path.isdir(REFERENCE):
        ref = os.path.join(REFERENCE, "matrices", "AR4JA_r45_4c_128c_r12.chinn.out")
        assert open(ref, "rb").read() == open(MATRIX, "rb").read()


def test_chinn_is_an_ar4ja_r12_labelling(chinn, code12):
    """Same degree profile as the CCSDS construction (a relabelled H)."""
    M, N, rows = chinn
    H = np.zeros((M, N), dtype=np.uint8)
    for i, r in enumerate(rows):
        H[i, r] = 1
    assert np.array_equal(np.sort(H.sum(1)), np.sort(code12.row_degrees()))
    assert np.array_equal(np.bincount(H[:, 2048:].sum(0)), np.bincount(code12.col_degrees()[2048:]))
    assert all(r == sorted(r) for r in rows)


def test_oracle_matches_reference_worker(scalar_golden, chinn, oracle):
    M, N, rows = chinn
    z = scalar_golden
    for i in range(len(z["nms"])):
        out, iters, ok = oracle.decode_r(N, rows, z["llr_in"][i], int(z["max_iter"][i]), bool(z["nms"][i]),
                                         float(z["r_max"][i]), float(z["l_post_max"][i]))
        np.testing.assert_array_equal(out, z["llr_out"][i], err_msg=f"case {i}")
        be = [int((out[c, :2048] >> 15).sum()) for c in range(2)]
        assert sum(be) == z["stats"][i][0] and sum(b > 0 for b in be) == z["stats"][i][1]
        assert int(iters[0]) == z["stats"][i][2]


def test_fixture_covers_variants(scalar_golden):
    z = scalar_golden
    assert set(z["nms"]) == {0, 1}
    assert (z["l_max"] > 0).any() and (z["r_max"] > 0).any() and (z["l_post_max"] > 0).any()
    assert (z["max_iter"] <= 3).any() and (z["stats"][:, 1] > 0).any() and (z["stats"][:, 1] == 0).any()


def test_syndrome_starts_at_third_iteration(chinn, oracle):
    """A clean channel converges in one iteration but is only detected at 3."""
    M, N, rows = chinn
    llr = np.full((1, N), 0x4100, dtype=np.uint16)  # +8.0
    llr[0, 2048:] = 0
    out, iters, ok = oracle.decode_r(N, rows, llr, 10)
    assert ok[0] == 1 and iters[0] == 3


def test_r_levels_are_sequentially_equivalent(chinn):
    """ASAP levels of the file row order (37 for the shipped matrix): for every
    variable, the rows touching it have strictly increasing levels in file
    order, so level-parallel decoding reads what the sequential order reads;
    libldpc's C++ levels equal the host ones."""
    from ldpc_mi355x import decoder, schedule_r
    M, N, rows = chinn
    lv = schedule_r.row_levels(rows, N)
    assert max(lv) == 37
    last = {}
    for i, r in enumerate(rows):
        for v in r:
            assert last.get(v, 0) < lv[i]
            last[v] = lv[i]
    assert np.array_equal(decoder.row_levels(rows, N), lv)
    assert sum(len(x) for x in schedule_r.levels(rows, N)) == M


def test_r_passes_are_sequentially_equivalent(chinn):
    """The kernel's list schedule (passes of <= 64 rows, longest remaining
    dependency chain first) takes 37 passes for the shipped matrix, the length
    of its chain; for every variable the rows touching it sit in strictly
    increasing passes in file order (so no pass holds two rows sharing a
    variable); libldpc's C++ schedule equals the host one, and at width 32 it
    packs the 1536 rows into 48 full passes."""
    from ldpc_mi355x import decoder, schedule_r
    M, N, rows = chinn
    ps = schedule_r.passes(rows, N, 64)
    assert max(ps) + 1 == 37 and np.bincount(ps).max() <= 64
    assert np.array_equal(decoder.row_passes(rows, N, 64), ps)
    last = {}
    for i, r in enumerate(rows):
        for v in r:
            assert last.get(v, -1) < ps[i]
            last[v] = ps[i]
    p32 = decoder.row_passes(rows, N, 32)
    assert np.array_equal(p32, schedule_r.passes(rows, N, 32))
    assert p32.max() + 1 == 48 and np.bincount(p32).min() == 32


def test_r_passes_on_random_matrices():
    """The pass schedule on random matrices (degrees 1..12, widths 1, 3, 7 and
    64): libldpc's C++ equals the host mirror, no pass is wider than asked,
    and the rows touching a variable sit in strictly increasing passes in
    file order (which also keeps two rows of a pass off a common variable)."""
    from ldpc_mi355x import decoder, schedule_r
    rng = np.random.default_rng(2026)
    for trial in range(40):
        n = int(rng.integers(12, 200))
        m = int(rng.integers(1, 120))
        rows = [sorted(rng.choice(n, size=int(rng.integers(1, min(12, n) + 1)), replace=False).tolist())
                for _ in range(m)]
        for width in (1, 3, 7, 64):
            ps = schedule_r.passes(rows, n, width)
            assert np.array_equal(decoder.row_passes(rows, n, width), ps), (trial, width)
            assert np.bincount(ps).max() <= width
            last = {}
            for i, r in enumerate(rows):
                for v in r:
                    assert last.get(v, -1) < ps[i], (trial, width, i)
                    last[v] = ps[i]
            if width == 1:
                assert max(ps) + 1 == m


def _check_passes(rows, ps, width, tag):
    assert np.bincount(ps).max() <= width, tag
    last = {}
    for i, r in enumerate(rows):
        for v in r:
            assert last.get(v, -1) < ps[i], (tag, i)
            last[v] = ps[i]


def test_r2_passes_by_degree_on_the_shipped_matrix(chinn, oracle):
    """The two-codeword kernel's schedule of the .chinn.out: passes of at most
    32 rows, each pass of one degree, the minimum 32 degree-6 passes plus 18
    degree-3 ones; sequentially equivalent (per-variable pass order follows
    file order), and the oracle's decode in that pass order equals the
    file-order decode bit for bit."""
    from conftest import channel_params
    from ldpc_mi355x import decoder
    M, N, rows = chinn
    ps, kind = decoder.row_passes_by_degree(rows, N, 32)
    assert len(kind) == ps.max() + 1 == 50
    assert (kind == 6).sum() == 32 and (kind == 3).sum() == 18
    for i, r in enumerate(rows):
        assert len(r) == kind[ps[i]]
    _check_passes(rows, ps, 32, "chinn")
    mu, sg = channel_params(1.2)
    llr = oracle.gen_channel_clamped(N, 2048, 23, 0, 48, mu, sg)
    order = sorted(range(M), key=lambda i: (ps[i], i))
    for nms in (False, True):
        seq = oracle.decode_r(N, rows, llr, 30, nms=nms)
        par = oracle.decode_r(N, [rows[i] for i in order], llr, 30, nms=nms)
        for a, b in zip(seq, par):
            np.testing.assert_array_equal(a, b)


def test_r2_passes_by_degree_on_random_matrices():
    """Random matrices whose rows all have degree 3 or 6 get one-degree passes
    (kinds 3 and 6); any other degree mix gets kind-7 passes equal to the plain
    list schedule.  Every schedule keeps the per-variable file order."""
    from ldpc_mi355x import decoder
    rng = np.random.default_rng(321)
    for trial in range(30):
        n = int(rng.integers(16, 300))
        m = int(rng.integers(1, 200))
        exact = trial % 2 == 0
        degs = rng.choice([3, 6], size=m) if exact else rng.integers(1, 7, size=m)
        rows = [sorted(rng.choice(n, size=int(d), replace=False).tolist()) for d in degs]
        for width in (5, 32):
            ps, kind = decoder.row_passes_by_degree(rows, n, width)
            assert len(kind) == ps.max() + 1
            _check_passes(rows, ps, width, (trial, width))
            if exact:
                assert set(kind.tolist()) <= {3, 6}
                for i, r in enumerate(rows):
                    assert len(r) == kind[ps[i]], (trial, width, i)
            elif not set(int(d) for d in degs) <= {3, 6}:
                assert set(kind.tolist()) == {7}
                assert np.array_equal(ps, decoder.row_passes(rows, n, width))


def test_pass_order_decode_equals_sequential_on_random_matrices(oracle):
    """On random matrices and random bf16 LLRs, the oracle's Profile R decode
    with the rows in the kernel's pass order (file order inside a pass)
    equals the file-order decode bit for bit: final LLRs, iterations and
    convergence, for Approx-Min* and NMS."""
    from ldpc_mi355x import schedule_r
    rng = np.random.default_rng(77)
    for trial in range(12):
        n = int(rng.integers(24, 120))
        m = int(rng.integers(8, n // 2 + 1))
        rows = [sorted(rng.choice(n, size=int(rng.integers(2, 13)), replace=False).tolist()) for _ in range(m)]
        llr = (rng.normal(1.0, 1.2, (6, n)).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
        ps = schedule_r.passes(rows, n, 64)
        order = sorted(range(m), key=lambda i: (ps[i], i))
        for nms in (False, True):
            seq = oracle.decode_r(n, rows, llr, 20, nms=nms)
            par = oracle.decode_r(n, [rows[i] for i in order], llr, 20, nms=nms)
            for a, b in zip(seq, par):
                np.testing.assert_array_equal(a, b)


def test_level_parallel_decode_equals_sequential(chinn, oracle):
    """Re-ordering rows level by level, or pass by pass as the GPU does,
    leaves the oracle's sequential Profile R result unchanged."""
    from ldpc_mi355x import schedule_r
    from conftest import channel_params
    M, N, rows = chinn
    mu, sg = channel_params(1.2)
    llr = oracle.gen_channel_clamped(N, 2048, 21, 0, 64, mu, sg)
    seq, it1, _ = oracle.decode_r(N, rows, llr, 40)
    order = [i for lv in schedule_r.levels(rows, N) for i in lv]
    par, it2, _ = oracle.decode_r(N, [rows[i] for i in order], llr, 40)
    np.testing.assert_array_equal(seq, par)
    np.testing.assert_array_equal(it1, it2)
    # the kernel's pass order (file order inside a pass)
    ps = schedule_r.passes(rows, N, 64)
    order = sorted(range(M), key=lambda i: (ps[i], i))
    par, it3, _ = oracle.decode_r(N, [rows[i] for i in order], llr, 40)
    np.testing.assert_array_equal(seq, par)
    np.testing.assert_array_equal(it1, it3)


def test_reference_worker_monte_carlo_mode_matches_fixture_mode(tmp_path):
    """The CPU baseline mode of oracle/_ref/scalar_driver runs the same worker
    call as the fixture mode: one thread, two codewords, same seed -> the same
    frame and bit errors as the fixture mode's stats; and at 1.0 dB the
    reference worker's FER agrees with the oracle's Profile R within sampling
    error (different channel streams)."""
    import json
    import subprocess
    exe = os.path.join(os.path.dirname(GOLDEN_DIR), "..", "oracle", "_ref", "scalar_driver")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/scalar_driver not built (needs the reference sources)")
    for ebn0, seed in ((0.5, 7), (1.0, 11)):
        out = str(tmp_path / f"f{seed}.bin")
        subprocess.run([exe, MATRIX, str(ebn0), "512", "30", str(seed), "0", "0", "0", out], check=True, timeout=60)
        stats = np.fromfile(out, dtype="<u4")[-4:]
        p = subprocess.run([exe, "--monte-carlo", MATRIX, str(ebn0), "512", "30", "2", "1", str(seed)],
                           capture_output=True, text=True, check=True, timeout=60)
        r = json.loads(p.stdout)
        assert (r["codewords"], r["bit_errors"], r["frame_errors"]) == (2, int(stats[0]), int(stats[1]))
    import math

    from scipy.stats import binomtest
    import oracle as O
    p = subprocess.run([exe, "--monte-carlo", MATRIX, "1.0", "512", "30", "500", "8", "5"],
                       capture_output=True, text=True, check=True, timeout=120)
    r = json.loads(p.stdout)
    M, N, rows = codes.read_chinn_out(MATRIX)
    g = 10 ** 0.1
    o = O.run_point_r(N, rows, 2048, 5, 0, 4000, 2 * g, math.sqrt(4 * g), 0.0, 30)
    k1, n1, k2, n2 = r["frame_errors"], r["codewords"], o["frame_errors"], o["blocks"]
    assert k1 > 50 and k2 > 50
    assert binomtest(k1, k1 + k2, n1 / (n1 + n2)).pvalue > 1e-3
