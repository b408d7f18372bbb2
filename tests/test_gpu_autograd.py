"""torch.autograd binding of the rollout (lenia.autograd) on the GPU.

The Function's gradients are the C ABI's own (bit-identical to
World.rollout_grad for the same upstream gradient), and an Adam fit of
mu / sigma / dt / shell peaks / initial state from perturbed values lowers
the MSE to a target produced by hidden parameters.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from lenia_cases import FFT, INIT_NONE, INIT_PATCH, config4, ring  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import lenia.autograd  # noqa: F401  (torch first: one HIP runtime)
    import lenia

    c = lenia.Context(0)
    yield c
    c.close()


def small_desc():
    return dict(width=24, height=20, convMode=FFT, growthMode=0, dt=0.1, radius=4,
                initMode=INIT_NONE,
                kernels=[ring(4, [1.0], 0.2, 0.05, 0.6), ring(5, [1.0, 0.6], 0.25, 0.06, 0.4)])


def params_of(desc, dev):
    ks = desc["kernels"]
    f64 = torch.float64
    mu = torch.tensor([k["mu"] for k in ks], dtype=f64, device=dev)
    sigma = torch.tensor([k["sigma"] for k in ks], dtype=f64, device=dev)
    weight = torch.tensor([k["weight"] for k in ks], dtype=f64, device=dev)
    betas = torch.tensor([(list(k["betas"]) + [0.0] * 4)[:4] for k in ks], dtype=f64, device=dev)
    dt = torch.tensor(desc["dt"], dtype=f64, device=dev)
    return mu, sigma, weight, betas, dt


def test_function_matches_c_abi(ctx):
    import lenia
    from lenia.autograd import rollout

    dev = torch.device("cuda:0")
    desc = small_desc()
    w = lenia.World(ctx, desc)
    rng = np.random.default_rng(11)
    A0n = (0.2 + 0.6 * rng.random((20, 24))).astype(np.float32)
    tgt = rng.random((20, 24)).astype(np.float32)
    T = 5
    A0 = torch.tensor(A0n, device=dev, requires_grad=True)
    mu, sigma, weight, betas, dt = [p.requires_grad_() for p in params_of(desc, dev)]
    AT = rollout(w, A0, mu, sigma, weight, betas, dt, T, ckpt_every=2)
    assert np.array_equal(AT.detach().cpu().numpy(), w.rollout(A0n, T))
    loss = ((AT - torch.tensor(tgt, device=dev)) ** 2).sum()
    loss.backward()
    gT = 2.0 * (AT.detach().cpu().numpy() - tgt)
    ref = w.rollout_grad(A0n, T, gT.astype(np.float32), ckpt_every=2)
    assert np.array_equal(A0.grad.cpu().numpy(), ref["A0"])
    assert np.array_equal(mu.grad.cpu().numpy(), ref["mu"])
    assert np.array_equal(sigma.grad.cpu().numpy(), ref["sigma"])
    assert np.array_equal(weight.grad.cpu().numpy(), ref["weight"])
    assert float(dt.grad) == ref["dt"]
    assert np.array_equal(betas.grad.cpu().numpy()[1, :2], ref["betas"][1, :2])
    w.close()


def test_adam_fit_lowers_loss(ctx):
    """256x256, two ring kernels R=16, T=16, checkpoints every 4: target from
    hidden parameters; 40 Adam steps from perturbed mu / sigma / dt / peaks.
    (The FP64 oracle driving the same Adam settings lowers the loss 40x.)"""
    import lenia
    from lenia.autograd import rollout

    dev = torch.device("cuda:0")
    hidden = dict(config4(), initMode=INIT_PATCH)
    w0 = lenia.World(ctx, hidden)
    w0.reset(77)
    w0.frame(0)
    A0n = w0.read()
    w0.close()
    desc = dict(hidden, initMode=INIT_NONE)
    w = lenia.World(ctx, desc)
    T = 16
    target = torch.tensor(w.rollout(A0n, T), device=dev)
    A0 = torch.tensor(A0n, device=dev)
    mu, sigma, weight, betas, dt = params_of(desc, dev)
    mu = (mu + torch.tensor([0.004, -0.006], dtype=torch.float64, device=dev)).requires_grad_()
    sigma = (sigma * 1.08).requires_grad_()
    dt = (dt * 0.9).requires_grad_()
    betas = betas.clone()
    betas[1, 1] = 0.4
    betas.requires_grad_()
    opt = torch.optim.Adam([mu, sigma, dt, betas], lr=5e-4)
    losses = []
    for it in range(40):
        opt.zero_grad()
        AT = rollout(w, A0, mu, sigma, weight, betas, dt, T, ckpt_every=4)
        loss = ((AT - target) ** 2).mean()
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    print("loss", losses[0], "->", losses[-1])
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.1 * losses[0], losses
    w.close()
