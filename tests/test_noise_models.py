"""Does the silicon's channel generator explain the error-floor gap?

At 2.5 dB the reference's silicon reports 61 frame errors in 4e10 codewords
(README.md:120) where this build's GPU run of the golden arithmetic finds 34
(p = 0.0073).  The two runs share the decoder arithmetic (golden fixtures)
but not the channel: the silicon draws xoshiro128+ uniforms and runs
Box-Muller with a bf16-coefficient log polynomial, Newton square roots and
Remez sin/cos (compute_trisc.cpp:30-65, 179-294).

The reference ships its own study of that generator,
tests/test_noise_models_comparison.cpp, which emulates the Tensix Box-Muller
(sfpi_host_emu.h) next to the host's polar method.  Built from source by
oracle/Makefile.ref and run here, it reports the silicon's normals with
variance about 0.995 (0.021 dB less noise than N(0, 1)) and 3.5-4.5 sigma
tails no heavier than the normal law's (0.99 of it at 4 sigma over 4e8
samples): such a channel lowers the floor, it cannot raise it.  So the
channel does not explain the silicon's higher count, and the floor's parity
with the silicon stays open (its decoder arithmetic is pinned only by the
host emulator).  This build's own stream (Philox4x32-10 + Box-Muller with
full-precision polynomials) sits on N(0, 1) within sampling error.
"""
import os
import re
import subprocess

import numpy as np
import pytest

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
STUDY = os.path.join(ORACLE_DIR, "_ref", "noise_models")
N_STUDY = 100_000_000


def parse_row(text, label):
    line = next(ln for ln in text.splitlines() if ln.startswith(label))
    return [float(x) for x in re.findall(r"-?\d+\.\d+(?:e[-+]\d+)?", line)]


@pytest.fixture(scope="module")
def study():
    if not os.path.exists(STUDY):
        pytest.skip("oracle/_ref/noise_models is built from the reference tree (make -f oracle/Makefile.ref)")
    p = subprocess.run([STUDY, str(N_STUDY), "8"], capture_output=True, text=True, check=True, timeout=600)
    return p.stdout


def test_silicon_noise_is_not_heavier_than_normal(study):
    var = parse_row(study, "Variance (sigma^2)")
    # columns: ideal, polar (host), Tensix both, Tensix z0 (cos), Tensix z1 (sin), ...
    ideal, polar, tensix, t_cos, t_sin = var[:5]
    se = np.sqrt(2.0 / N_STUDY)                      # standard error of a sample variance
    print("variance: polar", polar, "Tensix", tensix, "(cos", t_cos, "sin", t_sin, ") SE", se)
    assert abs(polar - 1.0) < 5 * se                 # the study's reference generator is sound
    assert tensix < 1.0 - 10 * se                    # the silicon's normals carry less noise
    for label, ideal_p in (("|Z| >= 3.5 sigma", 4.6526e-04), ("|Z| >= 4.0 sigma", 6.3342e-05),
                           ("|Z| >= 4.5 sigma", 6.7953e-06)):
        row = parse_row(study, label)
        t_both = row[2]
        se_p = np.sqrt(ideal_p / N_STUDY)
        print(label, "ideal", ideal_p, "Tensix", t_both)
        assert t_both < ideal_p + 3 * se_p            # no heavier tail where trapping sets live


def test_own_stream_is_standard_normal(oracle):
    """Moments and tails of this build's channel normals (the oracle's
    bit-identical restatement of the device stream), 4e7 of them."""
    from scipy.stats import norm
    n_cw, n_tx = 20000, 2048
    z = oracle.normals(0x5EED2026, 0, n_cw, n_tx).astype(np.float64).ravel()
    n = z.size
    var = z.var()
    print("own stream: mean", z.mean(), "variance", var, "SNR shift dB", -10 * np.log10(var))
    assert abs(z.mean()) < 5 / np.sqrt(n)
    assert abs(var - 1.0) < 5 * np.sqrt(2.0 / n)
    for k in (1.0, 2.0, 3.0, 3.5, 4.0):
        p = np.mean(np.abs(z) >= k)
        exp = 2 * norm.sf(k)
        assert abs(p - exp) < 5 * np.sqrt(exp / n), (k, p, exp)
