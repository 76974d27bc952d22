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
