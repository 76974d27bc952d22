"""The CPU oracle, pinned against the reference and against mathematics.

  * Profile T decode == the reference golden stepper (tests/golden fixtures,
    written by oracle/_ref/stepper built from the reference's own source);
  * min* == its closed form on exhaustive / random inputs, and the oracle's
    min* differs from plausible mistakes (untruncated constants, missing
    clamp, swapped lines);
  * Philox4x32-10 == Random123's published known-answer vectors;
  * Box-Muller noise: moments, tail counts against Q(x), raw BPSK sign-error
    rate against Q(mu/sigma) (uncoded_bpsk_validation.md:28-53 methodology).
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import channel_params


def bf16_to_f32(h):
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32)


def f32_to_bf16(f):
    u = np.asarray(f, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


# ---------------------------------------------------------------- golden ---

def _groups(golden):
    for i in range(0, len(golden["label"]), 4):
        yield i, str(golden["label"][i]), int(golden["max_iter"][i]), float(golden["beta"][i])


def test_fixture_matches_code(golden, code12):
    assert str(golden["code_digest"]) == code12.digest()


def test_oracle_matches_reference_stepper(golden, code12, oracle):
    for i, label, mi, beta in _groups(golden):
        out, iters, ok = oracle.decode_t(code12, golden["llr_in"][i:i + 4], mi, beta)
        np.testing.assert_array_equal(out, golden["llr_out"][i:i + 4], err_msg=label)
        conv = np.where(ok == 1, iters, 0)
        np.testing.assert_array_equal(conv, golden["conv_iter"][i:i + 4], err_msg=label)
        np.testing.assert_array_equal(iters[ok == 0], mi)


def test_oracle_trace_matches_unsat_counts(golden, code12, oracle):
    """Per-iteration unsatisfied-check counts equal the stepper's "Unsat=" lines."""
    for i, label, mi, beta in _groups(golden):
        for j in range(4):
            unsat, _ = oracle.trace_t(code12, golden["llr_in"][i + j], mi, beta)
            ref = golden["unsat"][i + j]
            ref = ref[ref >= 0]
            conv = int(golden["conv_iter"][i + j])
            assert len(unsat) == (conv if conv else mi), label
            np.testing.assert_array_equal(unsat, ref[:len(unsat)], err_msg=f"{label} cw{j}")


def test_fixture_covers_edge_cases(golden):
    llr = golden["llr_in"]
    assert (llr == 0x8000).any() and (llr == 0).any()          # -0 and +0 inputs
    assert (golden["conv_iter"] == 0).any()                      # non-converged codewords
    assert set(np.unique(golden["beta"])) == {np.float32(0.9375), np.float32(1.0)}


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "stepper")),
                    reason="reference stepper not built (make -f oracle/Makefile.ref)")
def test_oracle_matches_live_reference_on_new_dump(tmp_path, code12, oracle):
    """A dump never seen by the fixture script, through the built reference."""
    mu, sg = channel_params(1.25)
    llr = oracle.gen_channel(code12, 0xC0FFEE, 777, 4, mu, sg)
    dump = tmp_path / "d.bin"
    dump.write_bytes(oracle.to_packed4(llr))
    subprocess.run([oracle.REF_STEPPER, "--sfpi", "--load-init", str(dump), "60", "0.9375"], cwd=tmp_path,
                   check=False, capture_output=True)
    ref = oracle.from_packed4((tmp_path / "stepper_final_vnodes.bin").read_bytes(), code12.n)
    out, _, _ = oracle.decode_t(code12, llr, 60, 0.9375)
    np.testing.assert_array_equal(out, ref)


# ----------------------------------------------------------------- min* ---

C1, K1, C2, K2 = 0.69140625, -0.3984375, 0.65625, -0.326171875


def amin_closed(a, b):
    a, b = np.float32(a), np.float32(b)
    mn, d = np.minimum(a, b), np.abs(a - b)
    l1 = np.float32(C1) + np.float32(K1) * d
    l2 = np.float32(C2) + np.float32(K2) * d
    corr = np.maximum(np.maximum(l1, l2), np.float32(0))
    return np.maximum(mn - corr, np.float32(0))


def test_amin_constants_are_truncated_bf16():
    for full, trunc in ((0.69315, C1), (-0.4, K1), (0.65715, C2), (-0.328, K2)):
        bits = np.float32(full).view(np.uint32) & np.uint32(0xFFFF0000)
        assert bits.view(np.float32) == np.float32(trunc)


def test_amin_closed_form_on_bf16_grid(oracle):
    """All pairs of a dense bf16 magnitude grid (0 .. 64) against the closed form."""
    grid = bf16_to_f32(np.arange(0x0000, 0x4281, 7, dtype=np.uint16))
    rng = np.random.default_rng(1)
    a = rng.choice(grid, 20000)
    b = rng.choice(grid, 20000)
    got = np.array([oracle.amin_t(float(x), float(y)) for x, y in zip(a, b)], dtype=np.float32)
    exp = amin_closed(a, b)
    np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))


def test_amin_properties(oracle):
    # symmetric, bounded by min, zero when the correction exceeds the minimum
    for a, b in ((3.0, 1.0), (0.25, 0.3), (10.0, 10.0), (0.0, 5.0)):
        r = oracle.amin_t(a, b)
        assert r == oracle.amin_t(b, a) and 0.0 <= r <= min(a, b)
    assert oracle.amin_t(0.5, 0.5) == 0.0             # corr(0) = C1 > 0.5
    assert oracle.amin_t(100.0, 1.0) == np.float32(1.0)  # far apart: no correction
    # plausible mistakes are distinguishable: untruncated constant, no clamp
    assert oracle.amin_t(2.0, 2.1) != np.float32(2.0) - (np.float32(0.69315) + np.float32(-0.4) * np.float32(0.1))
    assert oracle.amin_t(0.2, 0.3) == 0.0


# --------------------------------------------------------------- Philox ---

# Random123 kat_vectors, philox4x32 with 10 rounds: (ctr, key) -> out
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, out):
    assert tuple(int(x) for x in oracle.philox4x32_10(ctr, key)) == out


# ---------------------------------------------------------------- noise ---

def test_log_sqrt_sincos_accuracy(oracle):
    rng = np.random.default_rng(2)
    for u in np.concatenate([rng.random(2000), [2.0 ** -33, 0.5, 0.70710677, 0.7071068, 1.0]]).astype(np.float32):
        u = max(float(u), 2.0 ** -33)
        assert abs(oracle.log_unit(u) - math.log(u)) <= 2e-7 * max(1.0, abs(math.log(u)))
    for t in rng.random(2000) * 50:
        assert abs(oracle.sqrt_newton(float(t)) - math.sqrt(t)) <= 1e-6 * max(1.0, math.sqrt(t))
    for x in rng.integers(0, 2 ** 32, 4000, dtype=np.uint64):
        s, c = oracle.sincos_turn(int(x))
        th = 2 * math.pi * int(x) / 2.0 ** 32
        assert abs(s - math.sin(th)) < 4e-7 and abs(c - math.cos(th)) < 4e-7


def _q(x):
    return 0.5 * math.erfc(x / math.sqrt(2.0))


@pytest.fixture(scope="module")
def normals(oracle):
    """2M standard normals of the Philox/Box-Muller stream, in channel order."""
    return oracle.normals(0x5EED2026, 0, 1024, 2048).astype(np.float64).ravel()


def test_noise_moments(normals):
    n = len(normals)
    sem = 1.0 / math.sqrt(n)
    assert abs(normals.mean()) < 4 * sem
    assert abs(normals.var() - 1.0) < 4 * math.sqrt(2.0 / n)
    assert abs((normals ** 3).mean()) < 4 * math.sqrt(15.0 / n)
    assert abs((normals ** 4).mean() - 3.0) < 4 * math.sqrt(96.0 / n)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_noise_tails_vs_q(normals, k):
    """Two-sided tail counts beyond k sigma against 2 Q(k), 5-sigma Poisson band
    (test_noise_models_comparison.cpp:243-289 methodology)."""
    n = len(normals)
    exp = 2 * _q(k) * n
    got = int((np.abs(normals) > k).sum())
    assert abs(got - exp) < 5 * math.sqrt(exp) + 1


def test_channel_llr_statistics(oracle, code12):
    """LLRs at 1.0 / 2.0 / 2.5 dB: mean mu, std sigma, sign-error rate Q(mu/sigma)."""
    for ebn0, q_ref in ((1.0, 0.1309), (2.0, 0.1040), (2.5, 0.0912)):
        mu, sg = channel_params(ebn0)
        llr = oracle.gen_channel(code12, 0xA11CE + int(ebn0 * 10), 0, 400, mu, sg)
        f = bf16_to_f32(llr[:, :code12.n_tx]).astype(np.float64).ravel()
        n = f.size
        assert abs(f.mean() - mu) < 5 * sg / math.sqrt(n) + 0.01
        assert abs(f.std() - sg) < 0.01 * sg
        p = (f < 0).mean()
        assert abs(p - _q(float(mu) / float(sg))) < 5 * math.sqrt(q_ref / n)
        assert abs(_q(float(mu) / float(sg)) - q_ref) < 5e-5
        assert (llr[:, code12.n_tx:] == 0).all()


def test_gen_channel_is_counter_based(oracle, code12):
    """Codeword w's LLRs depend only on (seed, w): any split gives the same stream."""
    mu, sg = channel_params(2.0)
    a = oracle.gen_channel(code12, 99, 10, 8, mu, sg)
    b = np.concatenate([oracle.gen_channel(code12, 99, 10 + i, 1, mu, sg) for i in range(8)])
    np.testing.assert_array_equal(a, b)
    c = oracle.gen_channel(code12, 100, 10, 8, mu, sg)
    assert (a[:, :code12.n_tx] != c[:, :code12.n_tx]).mean() > 0.9


# --------------------------------------------------------- Monte Carlo ---

def test_run_point_equals_decode_of_generated(oracle, code12):
    """The fused oracle Monte Carlo equals gen_channel + decode + tally."""
    mu, sg = channel_params(1.0)
    n = 96
    r = oracle.run_point_t(code12, 4242, 1000, n, mu, sg, 200, 0.9375)
    llr = oracle.gen_channel(code12, 4242, 1000, n, mu, sg)
    out, iters, ok = oracle.decode_t(code12, llr, 200, 0.9375)
    fail = ok == 0
    info = out[:, :code12.k]
    be = int(((info & 0x8000) | (info == 0)).astype(bool)[fail].sum())
    assert r["blocks"] == n and r["frame_errors"] == int(fail.sum()) and r["bit_errors"] == be
    assert r["iter_sum"] == int(iters.sum())
    assert r["hist"][0] == int(fail.sum()) and r["hist"].sum() == n
    for i in range(1, 201):
        assert r["hist"][i] == int(((iters == i) & ~fail).sum())


def test_mean_iterations_near_golden(oracle, code12):
    """Mean iterations at 2.0 dB are about 6.8: the reference golden stepper's Monte Carlo
    reports 6.81 on 40k codewords at that point."""
    mu, sg = channel_params(2.0)
    r = oracle.run_point_t(code12, 1, 0, 3000, mu, sg, 200, 0.9375)
    assert 6.4 < r["iter_sum"] / r["blocks"] < 7.3
