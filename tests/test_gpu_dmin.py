"""Differentiable min-sum decoder on the GPU (csrc/dmin_torch.hip): forward
bit for bit against the oracle, backward against the PyTorch-autograd
restatement and against float64 central finite differences, checkpointing
that changes nothing, and a short training run."""
import numpy as np
import pytest
import torch

import dmin_oracle as DO
import dmin_reference as REF
from conftest import channel_params
from ldpc_mi355x import codes

pytestmark = pytest.mark.gpu


def dmin():
    from ldpc_mi355x import dmin as D
    return D


def llr_batch(code, B, ebn0, seed):
    rng = np.random.default_rng(seed)
    mu, sg = channel_params(ebn0, code.k / code.n_tx)
    x = np.zeros((B, code.n), np.float32)
    x[:, :code.n_tx] = (mu + sg * rng.standard_normal((B, code.n_tx))).astype(np.float32)
    return x


def rand_params(code, G, seed):
    rng = np.random.default_rng(seed)
    P = DO.n_params(code)
    return (rng.uniform(0.7, 1.0, (G, P)).astype(np.float32), rng.uniform(0.05, 0.3, (G, P)).astype(np.float32))


def run_ext(code, llr, a, b, T, corr, ckpt_every=4, grad=True, r_max=0.0):
    D = dmin()
    tab = D.schedule_table(code)
    x = torch.tensor(llr, device="cuda", requires_grad=grad)
    at = torch.tensor(a, device="cuda", requires_grad=grad)
    bt = torch.tensor(b, device="cuda", requires_grad=grad)
    L = D.DMinSumFunction.apply(x, at, bt, tab, T, ckpt_every, corr, r_max)
    if grad:
        D.soft_bce(L, code.n_tx).backward()
        torch.cuda.synchronize()
        return L.detach().cpu().numpy(), x.grad.cpu().numpy(), at.grad.cpu().numpy(), bt.grad.cpu().numpy()
    return L.detach().cpu().numpy()


@pytest.mark.parametrize("k,rate,B,T", [(1024, "1/2", 64, 4), (1024, "4/5", 32, 3), (4096, "2/3", 8, 2),
                                        (16384, "1/2", 4, 2)])
@pytest.mark.parametrize("corr,r_max", [(False, 0.0), (True, 0.0), (False, 3.0), (2, 0.0)])
def test_forward_matches_oracle_bit_for_bit(k, rate, B, T, corr, r_max):
    code = codes.ar4ja(k, rate)
    llr = llr_batch(code, B, 1.0 + 0.8 * (rate != "1/2"), 3)
    a, b = rand_params(code, T, 4)
    print(f"forward K={k} rate={rate} B={B} T={T} corr={corr} r_max={r_max}", flush=True)
    got = run_ext(code, llr, a, b, T, corr, grad=False, r_max=r_max)
    exp = DO.forward(code, llr, a, b, T, corr, r_max=r_max)
    np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))


@pytest.mark.parametrize("corr,r_max", [(False, 0.0), (True, 0.0), (True, 2.5), (2, 0.0), (2, 2.5)])
@pytest.mark.parametrize("G", [1, 5])
def test_backward_matches_torch_autograd(corr, r_max, G):
    code = codes.ar4ja(1024, "1/2")
    T = 5
    llr = llr_batch(code, 48, 1.4, 7)
    a, b = rand_params(code, G, 8)
    L, gx, ga, gb = run_ext(code, llr, a, b, T, corr, r_max=r_max)
    x = torch.tensor(llr, device="cuda", requires_grad=True)
    at = torch.tensor(a, device="cuda", requires_grad=True)
    bt = torch.tensor(b, device="cuda", requires_grad=True)
    Lr = REF.forward(code, x, at, bt, T, corr, r_max)
    REF.loss(Lr, code.n_tx).backward()
    np.testing.assert_array_equal(L.view(np.uint32), Lr.detach().cpu().numpy().view(np.uint32))
    np.testing.assert_allclose(gx, x.grad.cpu().numpy(), rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(ga, at.grad.cpu().numpy(), rtol=1e-4, atol=1e-8)
    np.testing.assert_allclose(gb, bt.grad.cpu().numpy(), rtol=1e-4, atol=1e-8)
    assert np.abs(ga).max() > 0 and np.abs(gb).max() > 0


@pytest.mark.parametrize("corr", [False, 2])
def test_gradient_against_float64_finite_differences(corr):
    """64 codewords, T = 20, tied parameters, central differences of the
    float64 oracle with step 1e-3; plain min-sum and the three-minimum
    correction.

    The operating point is 2.0 dB with messages capped at 8, where every
    codeword of the batch converges.  A codeword that fails makes the
    unrolled decoder chaotic: argmin and sign decisions flip under 1e-3
    perturbations and float32 / float64 trajectories part, so a difference
    quotient there measures nothing (the pathwise gradient is still what the
    PyTorch-autograd test above pins)."""
    code = codes.ar4ja(1024, "1/2")
    T = 20
    llr = llr_batch(code, 64, 2.0, 11)
    a, b = rand_params(code, 1, 12)
    print(f"finite differences K=1024 T=20 corr={corr}", flush=True)
    L, _, ga, gb = run_ext(code, llr, a, b, T, corr, r_max=8.0)
    assert ((L[:, :code.n_tx] <= 0).sum(1) == 0).all(), "every codeword of the batch converges"
    idx = [(0, p) for p in (0, 5, 11, 17, 26, 33, 41, 52, 59)]
    an, fd = [], []
    for which, g in (("alpha", ga), ("beta", gb)):
        f = DO.fd_grad(code, llr, a, b, T, which, idx, 1e-3, corr=corr, r_max=8.0)
        for k, (gg, p) in enumerate(idx):
            an.append(float(g[gg, p]))
            fd.append(float(f[k]))
            print(which, p, an[-1], fd[-1], flush=True)
    an, fd = np.array(an), np.array(fd)
    # entries far below the largest sit at float32's summation noise floor,
    # so the error is measured against the gradient's scale
    err = np.abs(an - fd).max() / np.abs(fd).max()
    cos = an @ fd / (np.linalg.norm(an) * np.linalg.norm(fd))
    print("max error / max |fd|", err, "cos", cos, flush=True)
    assert np.abs(fd).max() > 0
    assert cos > 0.9999 and err < 1e-3


@pytest.mark.parametrize("trunc", [0, 1, 3])
def test_per_iteration_outputs_gradient_matches_torch_autograd(trunc):
    """Multi-loss: the BCE averaged over every iteration's soft output, with
    backpropagation through all iterations or truncated every `trunc`."""
    D = dmin()
    code = codes.ar4ja(1024, "1/2")
    T = 4
    llr = llr_batch(code, 32, 1.4, 31)
    a, b = rand_params(code, 1, 32)
    tab = D.schedule_table(code)
    x = torch.tensor(llr, device="cuda", requires_grad=True)
    at = torch.tensor(a, device="cuda", requires_grad=True)
    bt = torch.tensor(b, device="cuda", requires_grad=True)
    Lall = D.DMinSumFunction.apply(x, at, bt, tab, T, 5, True, 0.0, True, trunc)
    assert tuple(Lall.shape) == (T, 32, code.n)
    D.soft_bce(Lall, code.n_tx).backward()
    x2 = torch.tensor(llr, device="cuda", requires_grad=True)
    a2 = torch.tensor(a, device="cuda", requires_grad=True)
    b2 = torch.tensor(b, device="cuda", requires_grad=True)
    Lr = REF.forward(code, x2, a2, b2, T, True, 0.0, all_iters=True, trunc=trunc)
    REF.loss(Lr, code.n_tx).backward()
    np.testing.assert_array_equal(Lall.detach().cpu().numpy(), Lr.detach().cpu().numpy())
    np.testing.assert_allclose(x.grad.cpu().numpy(), x2.grad.cpu().numpy(), rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(at.grad.cpu().numpy(), a2.grad.cpu().numpy(), rtol=1e-4, atol=1e-8)
    np.testing.assert_allclose(bt.grad.cpu().numpy(), b2.grad.cpu().numpy(), rtol=1e-4, atol=1e-8)


def test_checkpoint_interval_changes_nothing():
    code = codes.ar4ja(1024, "1/2")
    T = 6
    llr = llr_batch(code, 32, 1.3, 21)
    a, b = rand_params(code, 1, 22)
    runs = [run_ext(code, llr, a, b, T, True, ckpt_every=k) for k in (1, 4, 12, T * 12)]
    for r in runs[1:]:
        for x, y in zip(runs[0], r):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("corr", [0, 1, 2])
def test_no_grad_forward_equals_the_autograd_forward(corr):
    """Under torch.no_grad the module skips the autograd.Function (and its
    checkpoints); the soft output is the same, bit for bit, for the last
    iteration and for every iteration."""
    D = dmin()
    model = D.DMinSum(1024, "1/2", T=5, ckpt_every=4, correction=corr)
    code = model.code
    a, b = rand_params(code, 1, 51)
    with torch.no_grad():
        model.alpha.copy_(torch.from_numpy(a))
        model.beta.copy_(torch.from_numpy(b))
    llr = torch.from_numpy(llr_batch(code, 16, 1.4, 52)).cuda()
    for all_iters in (False, True):
        with torch.no_grad():
            quick = model(llr, all_iters=all_iters)
        full = model(llr, all_iters=all_iters)
        assert full.requires_grad and not quick.requires_grad
        np.testing.assert_array_equal(quick.cpu().numpy(), full.detach().cpu().numpy())


def test_training_draws_mined_codewords():
    """hard_frac = 1 with a one-codeword hard set: every row of the batch is
    that codeword, so the first step's loss (taken before the update) is its
    own BCE."""
    D = dmin()
    torch.manual_seed(0)
    model = D.DMinSum(1024, "1/2", T=6, ckpt_every=12)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(9)
    one = D.channel_llr(model.code, torch.full((1,), 1.2, device="cuda"), gen)
    with torch.no_grad():
        own = float(D.soft_bce(model(one), model.n_tx))
    log = D.train(model, steps=2, batch=2048, micro_batch=1024, lr=1e-3, ebn0_range=(2.0, 2.0), seed=3,
                  log_every=1, hard=one, hard_frac=1.0)
    assert abs(log[0].loss - own) <= 1e-5 * abs(own), (log[0].loss, own)
    # and without it the loss is the channel's at 2.0 dB, far from this 1.2 dB codeword's
    model2 = D.DMinSum(1024, "1/2", T=6, ckpt_every=12)
    log2 = D.train(model2, steps=1, batch=2048, micro_batch=1024, lr=1e-3, ebn0_range=(2.0, 2.0), seed=3,
                   log_every=1)
    assert abs(log2[0].loss - own) > 1e-3 * abs(own)


def test_training_steps_run():
    D = dmin()
    torch.manual_seed(0)
    model = D.DMinSum(1024, "1/2", T=10, ckpt_every=12)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    eb = torch.full((4096,), 1.5, device="cuda")
    llr = D.channel_llr(model.code, eb, gen)
    with torch.no_grad():
        before = float(D.soft_bce(model(llr), model.n_tx))
    log = D.train(model, steps=30, batch=4096, micro_batch=2048, lr=1e-2, ebn0_range=(1.5, 1.5), seed=1,
                  log_every=10)
    with torch.no_grad():
        after = float(D.soft_bce(model(llr), model.n_tx))
    print("loss", before, "->", after, [(l.step, l.loss) for l in log], flush=True)
    assert np.isfinite(after) and all(np.isfinite(l.loss) for l in log)
    assert float((model.alpha.detach() - 1).abs().max()) > 0 and float(model.beta.detach().abs().max()) > 0


@pytest.mark.parametrize("k,rate,B,T,ckpt_every,ebn0", [
    (1024, "4/5", 4, 3, 5, 2.6),       # degree-18 checks, checkpoints every 5 of 12 layer steps
    (4096, "2/3", 3, 3, 7, 2.0),       # degree-10 checks, z = 256
    (16384, "1/2", 2, 3, 5, 1.0),      # z = 2048, the large training configuration's code
    (16384, "1/2", 2, 6, 60, 1.0),     # checkpoints every 5 iterations (60 layer steps), as trained at T = 50
])
@pytest.mark.parametrize("corr", [1, 2])
def test_backward_other_codes_matches_torch_autograd(k, rate, B, T, ckpt_every, ebn0, corr):
    """The backward for degree 10 and 18 and for K=16384 (z = 2048), with the
    min* correction (two and three minima) and checkpoint intervals that are
    not multiples of the 12 layers per iteration, so that every recompute
    starts mid-iteration."""
    code = codes.ar4ja(k, rate)
    llr = llr_batch(code, B, ebn0, 41)
    a, b = rand_params(code, 1, 42)
    print(f"backward K={k} rate={rate} B={B} T={T} ckpt_every={ckpt_every} corr={corr}", flush=True)
    L, gx, ga, gb = run_ext(code, llr, a, b, T, corr, ckpt_every=ckpt_every)
    x = torch.tensor(llr, device="cuda", requires_grad=True)
    at = torch.tensor(a, device="cuda", requires_grad=True)
    bt = torch.tensor(b, device="cuda", requires_grad=True)
    Lr = REF.forward(code, x, at, bt, T, corr, 0.0)
    REF.loss(Lr, code.n_tx).backward()
    np.testing.assert_array_equal(L.view(np.uint32), Lr.detach().cpu().numpy().view(np.uint32))
    np.testing.assert_allclose(gx, x.grad.cpu().numpy(), rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(ga, at.grad.cpu().numpy(), rtol=1e-4, atol=1e-8)
    np.testing.assert_allclose(gb, bt.grad.cpu().numpy(), rtol=1e-4, atol=1e-8)
    assert np.abs(ga).max() > 0 and np.abs(gb).max() > 0
    # the checkpoint interval changes nothing
    L1, gx1, ga1, gb1 = run_ext(code, llr, a, b, T, corr, ckpt_every=1)
    for p, q in ((L, L1), (gx, gx1), (ga, ga1), (gb, gb1)):
        np.testing.assert_array_equal(p, q)


def test_gradient_k16384_against_float64_finite_differences():
    """K=16384 r1/2 with the shipped trained parameters (min* correction),
    T = 12 at 2.0 dB, messages capped at 8, 4 codewords that all converge:
    the GPU gradient against float64 central differences.

    With 131,072 bits in the batch some message sits within 1e-3 of a kink
    (a cap, a correction branch, a first/second minimum swap) for most
    parameters, so the step-1e-3 quotient of the K=1024 test crosses kinks
    here (measured: cosine 0.72).  The float64 oracle allows steps of 1e-6
    and 1e-7; each parameter is compared with the quotient at the step that
    lies closest, as a quotient across a kink is no derivative."""
    import os
    from ldpc_mi355x import dmin as D
    code = codes.ar4ja(16384, "1/2")
    T = 12
    z = np.load(os.path.join(os.path.dirname(D.__file__), "data", "dmin_r12_k16384_t50_corr.npz"))
    a, b = z["alpha"].astype(np.float32), z["beta"].astype(np.float32)
    rng = np.random.default_rng(13)
    mu, sg = channel_params(2.0, 0.5)
    llr = np.zeros((4, code.n), np.float32)
    llr[:, :code.n_tx] = (mu + sg * rng.standard_normal((4, code.n_tx))).astype(np.float32)
    print("K=16384 finite differences, T=12, 4 codewords", flush=True)
    L, _, ga, gb = run_ext(code, llr, a, b, T, True, ckpt_every=60, r_max=8.0)
    assert ((L[:, :code.n_tx] <= 0).sum(1) == 0).all(), "every codeword of the batch converges"
    idx = [(0, p) for p in (0, 7, 19, 30, 44, 59)]
    an, fd = [], []
    for which, g in (("alpha", ga), ("beta", gb)):
        quot = [DO.fd_grad(code, llr, a, b, T, which, idx, eps, corr=True, r_max=8.0) for eps in (1e-6, 1e-7)]
        for kk, (gg, p) in enumerate(idx):
            an.append(float(g[gg, p]))
            fd.append(min((float(q[kk]) for q in quot), key=lambda v: abs(v - an[-1])))
            print(which, p, an[-1], [float(q[kk]) for q in quot], flush=True)
    an, fd = np.array(an), np.array(fd)
    err = np.abs(an - fd).max() / np.abs(fd).max()
    cos = an @ fd / (np.linalg.norm(an) * np.linalg.norm(fd))
    print("max error / max |fd|", err, "cos", cos, flush=True)
    assert np.abs(fd).max() > 0
    assert cos > 0.9999 and err < 1e-3
