"""Packed4CW dump format (the reference stepper's --load-init input,
stepper:879-880): our writer equals the oracle's, which produced the dumps
the reference binary replayed for tests/golden/stepper_profile_t.npz."""
import numpy as np

from ldpc_mi355x import dumps


def test_packed4_round_trip_and_layout(oracle):
    rng = np.random.default_rng(4)
    x = rng.integers(0, 1 << 16, size=(4, 2560), dtype=np.uint32).astype(np.uint16)
    b = dumps.to_packed4(x)
    assert len(b) == 20480
    assert b == oracle.to_packed4(x)
    np.testing.assert_array_equal(dumps.from_packed4(b), x)
    # variable-major, four little-endian bf16 per variable
    assert b[:8] == bytes([x[0, 0] & 255, x[0, 0] >> 8, x[1, 0] & 255, x[1, 0] >> 8,
                           x[2, 0] & 255, x[2, 0] >> 8, x[3, 0] & 255, x[3, 0] >> 8])


def test_packed4_rejects_wrong_sizes():
    import pytest
    with pytest.raises(ValueError):
        dumps.to_packed4(np.zeros((3, 2560), np.uint16))
    with pytest.raises(ValueError):
        dumps.from_packed4(b"\0" * 100)


def test_trapping_set_signature():
    """(a, b) from the definition: a single wrong variable leaves exactly its
    own checks unsatisfied; two wrong variables satisfy again the checks they
    share; -0 (0x8000) is not a hard 1 (bf16 < 0 is false for it)."""
    from ldpc_mi355x import codes
    code = codes.ar4ja(1024, "1/2")
    pos = np.full(code.n, 0x3F80, np.uint16)  # +1.0 everywhere: the all-zero codeword
    ts = dumps.trapping_set(code, pos)
    assert (ts["a"], ts["b"], ts["info_bits"]) == (0, 0, 0)
    deg = code.col_degrees()
    checks_of = [set() for _ in range(code.n)]
    for i, r in enumerate(code.rows):
        for v in r:
            checks_of[v].add(i)
    # a punctured variable of the highest column degree, and an information bit sharing a check with it
    v1 = int(code.n_tx + np.argmax(deg[code.n_tx:]))
    v2 = next(v for v in range(code.k) if checks_of[v] & checks_of[v1])
    x = pos.copy()
    x[v1] = 0xBF80  # -1.0
    ts = dumps.trapping_set(code, x)
    assert (ts["a"], ts["b"], ts["info_bits"]) == (1, int(deg[v1]), 0)
    assert ts["checks"] == sorted(checks_of[v1])
    x[v2] = 0xC000  # -2.0
    shared = len(checks_of[v1] & checks_of[v2])
    ts = dumps.trapping_set(code, x)
    assert (ts["a"], ts["info_bits"]) == (2, 1) and shared >= 1
    assert ts["b"] == int(deg[v1] + deg[v2]) - 2 * shared
    assert ts["checks"] == sorted(checks_of[v1] ^ checks_of[v2])
    x[v2] = 0x8000  # -0: not a hard 1
    assert dumps.trapping_set(code, x)["a"] == 1
    import pytest
    with pytest.raises(ValueError):
        dumps.trapping_set(code, pos[:100])


def test_trap_cli_needs_ids():
    import pytest
    with pytest.raises(SystemExit):
        dumps.main(["--seed", "1", "--ebn0", "2.5"])
