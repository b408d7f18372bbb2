"""torch.autograd binding of the rollout (lenia.autograd) on the GPU.

The Function's gradients are the C ABI's own (bit-identical to
World.rollout_grad for the same upstream gradient), and 200 Adam iterations
of mu / sigma / dt / mixing weights / shell peaks / initial state from
perturbed values lower the MSE to a T=64 target produced by hidden ones.
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


def test_adam_fit_config4(ctx):
    """The differentiable-rollout fit: 256x256, two ring kernels R=16, T=64
    with checkpoints every 8.  The target is the T=64 state from hidden
    parameters (mu, sigma, dt, mixing weights, shell peaks) and a hidden
    initial state (the world's seeded noise after one step).  200 Adam
    iterations fit all of them, the full initial state included, from
    perturbed values (mu +-0.005, sigma +8%, dt -10%, weights +-5%, one shell
    peak 0.5 -> 0.4, initial state + 0.02 N(0, 1) clipped to [0, 1]).
    Measured on MI355X: the MSE falls 1.5e-2 -> 4.7e-5 (325x) and the growth
    parameters move towards the hidden ones; the gate is 100x and halved
    mu / sigma errors."""
    import lenia
    from lenia.autograd import rollout

    dev = torch.device("cuda:0")
    f64 = torch.float64
    hidden = dict(config4(), initMode=INIT_PATCH)
    w0 = lenia.World(ctx, hidden)
    w0.reset(77)
    w0.frame(0)
    A0n = w0.read()
    w0.close()
    desc = dict(hidden, initMode=INIT_NONE)
    w = lenia.World(ctx, desc)
    T, C = 64, 8
    target = torch.tensor(w.rollout(A0n, T, C), device=dev)
    hmu, hsg, hw, hb, hdt = params_of(desc, dev)
    hA = torch.tensor(A0n, device=dev)
    noise = torch.tensor(np.random.default_rng(3).standard_normal(A0n.shape).astype(np.float32),
                         device=dev)
    mu = (hmu + torch.tensor([0.004, -0.006], dtype=f64, device=dev)).requires_grad_()
    sigma = (hsg * 1.08).requires_grad_()
    dt = (hdt * 0.9).requires_grad_()
    weight = (hw * torch.tensor([1.05, 0.95], dtype=f64, device=dev)).requires_grad_()
    betas = hb.clone()
    betas[1, 1] = 0.4
    betas.requires_grad_()
    A0 = (hA + 0.02 * noise).clamp(0, 1).requires_grad_()
    start = {"mu": float((mu.detach() - hmu).abs().max()),
             "sigma": float((sigma.detach() - hsg).abs().max())}
    opt = torch.optim.Adam([{"params": [mu, sigma, dt, weight, betas], "lr": 5e-4},
                            {"params": [A0], "lr": 2e-3}])
    losses = []
    for it in range(200):
        opt.zero_grad()
        AT = rollout(w, A0, mu, sigma, weight, betas, dt, T, ckpt_every=C)
        loss = ((AT - target) ** 2).mean()
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    end = {"mu": float((mu.detach() - hmu).abs().max()),
           "sigma": float((sigma.detach() - hsg).abs().max())}
    print(f"Adam 200 iterations: loss {losses[0]:.3e} -> {losses[-1]:.3e} "
          f"({losses[0] / losses[-1]:.0f}x); |mu - hidden| {start['mu']:.2e} -> {end['mu']:.2e}, "
          f"|sigma - hidden| {start['sigma']:.2e} -> {end['sigma']:.2e}, "
          f"|dt - hidden| {abs(float(dt.detach() - hdt)):.2e}, "
          f"A0 rms to hidden {float(((A0.detach() - hA) ** 2).mean().sqrt()):.2e}")
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.01 * losses[0], losses[::20]
    assert end["mu"] < 0.5 * start["mu"] and end["sigma"] < 0.5 * start["sigma"], (start, end)
    assert A0.grad is not None and float(A0.grad.abs().max()) > 0
    w.close()
