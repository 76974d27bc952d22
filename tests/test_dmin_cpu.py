"""Differentiable min-sum decoder on the CPU: the numpy oracle and the
torch-autograd restatement agree (forward bit for bit in float32, gradient
within finite-difference error in float64), and the parameterisation reduces
to plain layered min-sum at alpha = 1, beta = 0."""
import numpy as np
import pytest
import torch

import dmin_oracle as DO
import dmin_reference as REF
from conftest import channel_params
from ldpc_mi355x import codes


@pytest.fixture(scope="module")
def code12():
    return codes.ar4ja(1024, "1/2")


def llr_batch(code, B, ebn0, seed):
    rng = np.random.default_rng(seed)
    mu, sg = channel_params(ebn0)
    x = np.zeros((B, code.n), np.float32)
    x[:, :code.n_tx] = (mu + sg * rng.standard_normal((B, code.n_tx))).astype(np.float32)
    return x


def rand_params(code, G, seed):
    rng = np.random.default_rng(seed)
    P = DO.n_params(code)
    return (rng.uniform(0.7, 1.0, (G, P)).astype(np.float32), rng.uniform(0.05, 0.3, (G, P)).astype(np.float32))


@pytest.mark.parametrize("corr,r_max", [(False, 0.0), (True, 0.0), (True, 2.0), (2, 0.0), (2, 2.0)])
def test_oracle_and_torch_reference_forward_agree(code12, corr, r_max):
    llr = llr_batch(code12, 8, 1.5, 1)
    a, b = rand_params(code12, 3, 2)
    got = DO.forward(code12, llr, a, b, 3, corr, r_max=r_max)
    ref = REF.forward(code12, torch.from_numpy(llr), torch.from_numpy(a), torch.from_numpy(b), 3, corr,
                      r_max).numpy()
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_unit_parameters_are_plain_layered_min_sum(code12):
    """alpha = 1, beta = 0: every message is the leave-one-out minimum with
    the product sign -- checked on one layer against a direct computation."""
    llr = llr_batch(code12, 4, 1.0, 3)
    P = DO.n_params(code12)
    L = DO.forward(code12, llr, np.ones((1, P), np.float32), np.zeros((1, P), np.float32), 1)
    # re-run the first layer by hand
    first, slots = code12.layers[0]
    z = code12.z
    j = np.arange(z)
    var = np.stack([s.base + (j + s.shift) % z for s in slots])
    q = llr[:, var]
    r = np.empty_like(q)
    for d in range(len(slots)):
        others = np.delete(q, d, axis=1)
        r[:, d] = np.prod(np.sign(others), axis=1) * np.min(np.abs(others), axis=1)
    L1 = llr.copy()
    L1[:, var] = q + r
    # layer 0 variables are not touched again before layer 4 (same block row); compare them after layer 0 only
    L_one = DO.forward(code12, llr, np.ones((1, P), np.float32), np.zeros((1, P), np.float32), 1)
    sch = DO.schedule(code12)
    Lc = llr.copy().astype(np.float32)
    Rc = np.zeros((4, sum(l[0] * l[1] for l in sch)), np.float32)
    DO._layer(Lc, Rc, sch[0], np.ones(P, np.float32), np.zeros(P, np.float32), False, np.float32)
    np.testing.assert_array_equal(Lc[:, var], L1[:, var].astype(np.float32))
    assert L.shape == L_one.shape


@pytest.mark.parametrize("corr,r_max", [(False, 0.0), (True, 0.0), (False, 2.0), (2, 0.0)])
def test_torch_reference_gradient_matches_finite_differences(code12, corr, r_max):
    llr = llr_batch(code12, 4, 1.5, 5)
    a, b = rand_params(code12, 1, 6)
    T = 3
    at = torch.tensor(a, dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, requires_grad=True)
    L = REF.forward(code12, torch.tensor(llr, dtype=torch.float64), at, bt, T, corr, r_max)
    REF.loss(L, code12.n_tx).backward()
    # The loss is piecewise smooth (relu gates, argmin swaps, correction
    # lines); a difference step can straddle a kink that lies within it.  A
    # parameter passes if the central difference agrees at one of several
    # small steps -- a wrong derivative rule disagrees at all of them.
    idx = [(0, p) for p in (0, 7, 19, 33, 45, 59)]
    for which, grad in (("alpha", at.grad), ("beta", bt.grad)):
        fds = [DO.fd_grad(code12, llr, a, b, T, which, idx, eps, corr, r_max) for eps in (1e-6, 1e-7, 1e-8)]
        for k, (g, p) in enumerate(idx):
            an = grad[g, p].item()
            assert any(abs(f[k] - an) <= 1e-4 * abs(an) + 1e-10 for f in fds), (which, p, an, [f[k] for f in fds])
    assert at.grad.abs().max() > 1e-4 and bt.grad.abs().max() > 1e-4


def _amin_scalar(lo, hi):
    """min*(lo, hi) for lo <= hi with the correction lines, float32 per op."""
    f = np.float32
    x = f(hi) - f(lo)
    c = max(f(f(DO.K1) * x) + f(DO.C1), f(f(DO.K2) * x) + f(DO.C2), f(0))
    return max(f(f(lo) - c), f(0))


@pytest.mark.parametrize("k,rate", [(1024, "1/2"), (1024, "4/5")])
def test_three_minima_give_each_edge_the_two_smallest_others(k, rate):
    """Three-minimum correction (corr = 2), one layer at alpha = 1, beta = 0,
    against a direct per-edge computation: the edge's magnitude is min* of
    the two smallest magnitudes among the OTHER edges of its check (so never
    its own), with the product sign of the others.  The magnitudes take few
    values, so ties are common: the result must not depend on which of two
    tied edges counts as the first, second or third minimum."""
    code = codes.ar4ja(k, rate)
    sch = DO.schedule(code)
    deg, z, var, _, _ = sch[1]
    rng = np.random.default_rng(7)
    B = 3
    llr = np.zeros((B, code.n), np.float32)
    llr[:, :] = rng.choice(np.array([-2.5, -1.0, -0.5, 0.5, 1.0, 1.25, 2.5, 4.0], np.float32), (B, code.n))
    P = DO.n_params(code)
    L = llr.copy()
    R = np.zeros((B, sum(l[0] * l[1] for l in sch)), np.float32)
    DO._layer(L, R, sch[1], np.ones(P, np.float32), np.zeros(P, np.float32), 2, np.float32)
    q = llr[:, var]                              # R = 0 before the first pass: q = L
    for b in range(B):
        for j in range(z):
            for d in range(deg):
                others = [q[b, e, j] for e in range(deg) if e != d]
                m = sorted(abs(float(v)) for v in others)
                mag = _amin_scalar(m[0], m[1])
                neg = sum(v < 0 or (v == 0 and np.signbit(v)) for v in others) % 2
                exp = np.float32(q[b, d, j] + (-mag if neg else mag))
                assert L[b, var[d, j]] == exp, (b, j, d)


@pytest.mark.parametrize("k,rate", [(1024, "1/2"), (1024, "4/5")])
def test_three_minima_select_by_magnitude_equals_select_by_index(k, rate):
    """The Profile D kernels pick each edge's pair by its magnitude (equal to
    min1: min*(min2, min3); equal to min2: min*(min1, min3); else
    min*(min1, min2)) where the oracle goes by first-index argmins.  On
    magnitudes drawn from four values (ties in nearly every check) both
    rules give the same layer output, bit for bit."""
    code = codes.ar4ja(k, rate)
    sch = DO.schedule(code)
    rng = np.random.default_rng(11)
    B = 16
    P = DO.n_params(code)
    al = rng.uniform(0.8, 1.0, P).astype(np.float32)
    be = rng.uniform(0.0, 0.2, P).astype(np.float32)
    for li, lay in enumerate(sch[:3]):
        deg, z, var, e_off, p_off = lay
        llr = rng.choice(np.array([-1.5, -0.5, 0.5, 1.5, 3.0], np.float32), (B, code.n))
        L = llr.copy()
        R = np.zeros((B, sum(l[0] * l[1] for l in sch)), np.float32)
        DO._layer(L, R, lay, al, be, 2, np.float32)
        q = llr[:, var]
        a = np.abs(q)
        srt = np.sort(a, axis=1)
        m1, m2, m3 = srt[:, 0:1], srt[:, 1:2], srt[:, 2:3]
        o12, o13, o23 = DO._amin(m1, m2, np.float32), DO._amin(m1, m3, np.float32), DO._amin(m2, m3, np.float32)
        g = np.where(a == m1, o23, np.where(a == m2, o13, o12)).astype(np.float32)
        zz = (al[p_off:p_off + deg, None] * g).astype(np.float32) - be[p_off:p_off + deg, None]
        mag = np.maximum(zz, np.float32(0))
        neg = np.bitwise_xor.reduce(np.signbit(q), axis=1)[:, None, :] ^ np.signbit(q)
        r = np.where(neg, -mag, mag).astype(np.float32)
        assert (a == m1).sum() > (a[:, :1] >= 0).sum(), "ties present"
        np.testing.assert_array_equal(L[:, var], (q + r).astype(np.float32), err_msg=f"layer {li}")


def test_three_minima_on_a_degree_3_check_is_the_golden_pairwise_min_star():
    """On a degree-3 layer each edge's two smallest others are simply the two
    other edges, so corr = 2 at alpha = 1, beta = 0 equals the golden tree's
    degree-3 update without damping: r_d = min*(|q_e|, |q_f|), e, f != d."""
    code = codes.ar4ja(1024, "1/2")
    sch = DO.schedule(code)
    lay = next(l for l in sch if l[0] == 3)
    deg, z, var, _, _ = lay
    rng = np.random.default_rng(9)
    llr = (rng.standard_normal((4, code.n)) * 3).astype(np.float32)
    P = DO.n_params(code)
    L = llr.copy()
    R = np.zeros((4, sum(l[0] * l[1] for l in sch)), np.float32)
    DO._layer(L, R, lay, np.ones(P, np.float32), np.zeros(P, np.float32), 2, np.float32)
    q = llr[:, var]
    a = np.abs(q)
    for d, (e, f) in enumerate([(1, 2), (0, 2), (0, 1)]):
        lo, hi = np.minimum(a[:, e], a[:, f]), np.maximum(a[:, e], a[:, f])
        mag = np.vectorize(_amin_scalar, otypes=[np.float32])(lo, hi)
        sign = np.where(np.signbit(q[:, e]) ^ np.signbit(q[:, f]), np.float32(-1), np.float32(1))
        np.testing.assert_array_equal(L[:, var[d]], (q[:, d] + sign * mag).astype(np.float32))


def test_schedule_table_matches_compiled_schedule(code12):
    from ldpc_mi355x import decoder, dmin
    tab = dmin.schedule_table(code12).numpy()
    sched = decoder.compiled_schedule(1024, "1/2")
    for i, (first, slots) in enumerate(sched):
        assert tab[i, 0] == len(slots) and tab[i, 1] == 128
        assert [(tab[i, 2 + 2 * d], tab[i, 3 + 2 * d]) for d in range(len(slots))] == slots


def test_soft_fer_is_one_minus_the_product_of_bit_probabilities():
    """soft_fer (1 - exp(-sum softplus(-L))) against the product of the
    per-bit probabilities sigmoid(L), formed in float64; positions past n_tx
    do not count; hopeless and clean codewords sit at 1 and 0."""
    from ldpc_mi355x import dmin as D
    rng = np.random.default_rng(4)
    n_tx, N = 40, 48
    L = torch.from_numpy(rng.normal(3.0, 4.0, (6, N))).float()
    L[:, n_tx:] = -50.0  # punctured positions: ignored
    direct = (1.0 - torch.sigmoid(L[:, :n_tx].double()).prod(dim=1)).mean()
    assert abs(float(D.soft_fer(L, n_tx)) - float(direct)) < 1e-6
    clean = torch.full((3, N), 30.0)
    hopeless = torch.full((3, N), -30.0)
    assert float(D.soft_fer(clean, n_tx)) < 1e-9
    assert abs(float(D.soft_fer(hopeless, n_tx)) - 1.0) < 1e-9
    # [T, B, N]: the mean over iterations as well
    stacked = torch.stack([clean, hopeless])
    assert abs(float(D.soft_fer(stacked, n_tx)) - 0.5) < 1e-9


def test_shipped_parameters_load_and_fit_their_codes():
    """Every shipped parameter file loads without pickling, names its code,
    holds finite float32 rows of the code's parameter count, and a
    correction the Profile D kernels implement; the three-minimum files are
    the ones with correction 2."""
    import glob
    import os
    from ldpc_mi355x import dmin as D
    files = sorted(glob.glob(os.path.join(os.path.dirname(D.__file__), "data", "dmin_*.npz")))
    assert len(files) >= 5
    for f in files:
        z = np.load(f, allow_pickle=False)
        code = codes.ar4ja(int(z["k_info"]), str(z["rate"]))
        P = DO.n_params(code)
        assert z["alpha"].dtype == np.float32 and z["alpha"].shape[1] == P and z["beta"].shape == z["alpha"].shape
        assert np.isfinite(z["alpha"]).all() and np.isfinite(z["beta"]).all()
        assert int(z["correction"]) in (0, 1, 2) and float(z["r_max"]) >= 0.0
        assert (int(z["correction"]) == 2) == f.endswith("_min3.npz"), f
